"""Multifrontal tangent solve over a nested-dissection order (gf_front.hip) on
the device.

* One Newton solve: the increment equals numpy's FP64 solution of the device's
  own assembled tangent (read back before the solve) to ~cond(K) eps, for a
  plate, a corners-first numbered plate and a beam chain (each forced through
  the fronts), and the band LU's increment on the same state.
* Static solves through the fronts match the oracle (MKL PARDISO where present):
  same converged steps, iterations within one, nodes within 1e-8, FP64 and
  mixed precision (FP32 fronts, FP64 refinement).
* A zero tangent fails the solve (info_fail), as in the band and block solvers.
"""
import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime
from renumber import corners_first, renumber
from test_gpu_configs import check_against_oracle, oracle_result

pytestmark = pytest.mark.gpu


def _one_solve(m, p, solver, seed=3, scale=1e-3):
    dm = runtime.DeviceModel(m, solver=solver)
    dm.set_window(p.start_time, p.end_time)
    dm.begin_step(0.0, 0.1)
    rng = np.random.default_rng(seed)
    free = runtime.front_plan(m)["dof"] >= 0
    d = rng.normal(size=(m.n_nodes, 6)) * scale
    d[~free] = 0.0
    dm.write_disp(d)
    dm.assemble()
    K = dm.tangent_dense()
    R = dm.residual()
    assert dm.solve() == 0
    dm.update()
    return dm, K, R, dm.increment()


def _cases():
    plate, pp, _ = M.shell_plate(nq=10, n_steps=4, p_max=4e6)
    cat, pc, _ = M.catenary(n_elem=80)
    yield "plate10", plate, pp
    yield "plate8_corners_first", renumber(M.shell_plate(nq=8)[0], corners_first(8)), M.shell_plate(nq=8)[1]
    yield "catenary80", cat, pc


@pytest.mark.parametrize("name", [c[0] for c in _cases()])
def test_fronts_solve_the_assembled_tangent(name):
    _, m, p = next(c for c in _cases() if c[0] == name)
    print(f"fronts one solve: {name}", flush=True)
    dm, K, R, inc = _one_solve(m, p, abi.GF_SOLVER_FRONTS)
    assert dm.info()["layout"] == 2
    free = runtime.front_plan(m)["dof"].ravel() >= 0
    Kf = K[np.ix_(free, free)]
    x = np.linalg.solve(Kf, R.ravel()[free])
    got = inc.ravel()[free]
    cond = np.linalg.cond(Kf)
    assert np.abs(got - x).max() <= 50 * cond * 2.2e-16 * np.abs(x).max(), (np.abs(got - x).max(), cond)
    # the band LU on the same state (BTD for the narrow catenary)
    _, Kb, Rb, incb = _one_solve(m, p, abi.GF_SOLVER_BAND)
    assert np.abs(Kb - K).max() <= 1e-14 * np.abs(K).max() and np.abs(Rb - R).max() <= 1e-14 * np.abs(R).max()
    assert np.abs(incb.ravel()[free] - got).max() <= 50 * cond * 2.2e-16 * np.abs(x).max()


@pytest.mark.parametrize("precision", [abi.GF_PREC_FP64, abi.GF_PREC_MIXED])
def test_fronts_static_plate_matches_oracle(oracle_lib, precision):
    m, p, _ = M.shell_plate(nq=16, n_steps=6, p_max=8e6)
    print(f"fronts static plate16 precision {precision}", flush=True)
    ref = oracle_result(oracle_lib, m, p)
    dm = runtime.DeviceModel(m, precision=precision, solver=abi.GF_SOLVER_FRONTS)
    assert dm.info()["layout"] == 2
    rep, iters = dm.static_solve(p)
    check_against_oracle(m, rep, iters, dm.nodes(), ref)
    if precision == abi.GF_PREC_MIXED:
        st = dm.stats()
        assert st["refine_sweeps"] > 0 and st["fallbacks"] == 0


def test_fronts_static_catenary_matches_oracle(oracle_lib):
    m, p, _ = M.catenary(n_elem=200)
    print("fronts static catenary200", flush=True)
    ref = oracle_result(oracle_lib, m, p)
    dm = runtime.DeviceModel(m, solver=abi.GF_SOLVER_FRONTS)
    rep, iters = dm.static_solve(p)
    check_against_oracle(m, rep, iters, dm.nodes(), ref)


def test_fronts_zero_tangent_fails():
    m, p, _ = M.shell_plate(nq=6, E=0.0)
    print("fronts zero tangent", flush=True)
    dm = runtime.DeviceModel(m, solver=abi.GF_SOLVER_FRONTS)
    dm.set_window(0.0, 1.0)
    dm.begin_step(0.0, 0.1)
    dm.assemble()
    assert dm.solve() == 1


def test_auto_picks_fronts_for_the_wide_plate():
    m, _, _ = M.shell_plate(nq=32)
    dm = runtime.DeviceModel(m)
    assert dm.info()["layout"] == 2
    m2, _, _ = M.catenary(n_elem=100)
    assert runtime.DeviceModel(m2).info()["layout"] == 0
