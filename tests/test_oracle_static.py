"""Pins the oracle's Static loop (src/Static.cpp:140-311) end to end by closed forms
and by the reference's own solver (MKL PARDISO, called as sparsesystem does)."""
import numpy as np
import pytest

from giraffe_amd import abi
from giraffe_amd import model as M


@pytest.mark.parametrize("t_end,theta", [(0.25, np.pi / 2), (0.5, np.pi), (0.75, 1.5 * np.pi), (1.0, 2 * np.pi)])
def test_rollup_closed_form(oracle_lib, t_end, theta):
    """Follower end moment M: uniform curvature M/EI, tip at (L/th)(sin th, 1-cos th),
    tip rotation th about z (Simo & Vu-Quoc 1986)."""
    m, p, info = M.cantilever_rollup()
    p.end_time = t_end
    om = oracle_lib.OracleModel(m)
    rep, iters = om.static_solve(p)
    assert rep.aborted == 0 and rep.diverged_steps == 0
    assert rep.converged_steps == round(t_end / 0.05)
    tip = om.nodes()[info["tip"]]
    L = info["L"]
    exp = (L / theta) * np.array([np.sin(theta), 1 - np.cos(theta)])
    assert np.allclose(tip[:2], exp, atol=1e-9 * L)
    assert abs(tip[2]) < 1e-12
    # Euler vector (Node::WriteResults): angle wrapped to (-pi, pi]
    ang = np.arctan2(np.sin(theta), np.cos(theta))
    assert abs(tip[5] - ang) < 1e-9 or abs(abs(tip[5]) - np.pi) < 1e-9
    # every node lies on the circle of radius L/theta up to the discretisation
    # error of the interpolated incremental rotations (the tip is exact)
    nd = om.nodes()
    R = L / theta
    r = np.hypot(nd[:, 0], nd[:, 1] - R)
    assert np.abs(r - R).max() < 1e-7 * L


def test_timoshenko_cantilever_small_load(oracle_lib):
    """Small dead tip force: PL^3/(3EI) + PL/(GA) (shear factor 1, src/Beam_1.cpp:567)."""
    m, p, info = M.cantilever_rollup(n_elem=40)
    m.loads.clear()
    E, nu = 210e9, 0.3
    A, I1, I2, I12, It = M.tube_section(0.2, 0.18)
    L = info["L"]
    P = 1e-3 * 3 * E * I1 / L ** 3 * L * 1e-3  # tip deflection ~1e-6 L
    m.add_load(abi.GF_LOAD_NODAL, [info["tip"]], [0.0, 1.0], [[0] * 6, [0, P, 0, 0, 0, 0]])
    # the load is ~1e-6 of the section's axial scale, so the residual floor of FP64
    # rounding (~1e-9 N) sits above 1e-10 * P: judge the residual at 1e-4
    p.force_tol = p.moment_tol = 1e-4
    p.force_min, p.moment_min = P, P * L
    p.disp_min, p.rot_min = L, 1.0  # increments judged against the problem scale
    om = oracle_lib.OracleModel(m)
    rep, _ = om.static_solve(p)
    assert rep.aborted == 0
    G = E / (2 * (1 + nu))
    exp = P * L ** 3 / (3 * E * I2) + P * L / (G * A)
    tip = om.nodes()[info["tip"]]
    assert tip[1] == pytest.approx(exp, rel=1e-6)


def test_band_matches_pardiso(oracle_lib):
    if not oracle_lib.pardiso_available():
        pytest.skip("libmkl_rt not present")
    m, p, info = M.cantilever_rollup(n_elem=30)
    a = oracle_lib.OracleModel(m, oracle_lib.SOLVER_BAND)
    b = oracle_lib.OracleModel(m, oracle_lib.SOLVER_PARDISO)
    ra, ia = a.static_solve(p)
    rb, ib = b.static_solve(p)
    assert np.array_equal(ia, ib)
    assert np.abs(a.nodes() - b.nodes()).max() < 1e-10


def test_bisection_and_abort_sequence(oracle_lib):
    """MaxIt too small: every attempt diverges and is bisected until the step is
    below MinTimeStep; the aborted attempt runs one iteration and is committed
    (diverged stays false on abort, src/Static.cpp:244-256)."""
    m, p, _ = M.cantilever_rollup(n_elem=10)
    p.max_it = 3
    p.min_time_step = p.time_step / 8
    om = oracle_lib.OracleModel(m)
    rep, iters = om.static_solve(p)
    assert rep.aborted == 1
    assert list(iters) == [-3, -3, -3, -3, 1]
    assert rep.diverged_steps == 4


def test_step_growth(oracle_lib):
    """IncFactor growth when a step converges within MinIt (src/Static.cpp:262-268)."""
    m, p, _ = M.cantilever_rollup(n_elem=10)
    p.time_step = 0.05
    p.max_time_step = 0.2
    p.inc_factor = 2.0
    p.min_it = 100  # every converged step counts as easy -> dt doubles up to MaxTimeStep
    om = oracle_lib.OracleModel(m)
    rep, iters = om.static_solve(p)
    # t: 0.05, 0.15, 0.35, 0.55, 0.75, 0.95, 1.0
    assert rep.converged_steps == 7
    assert rep.final_time == 1.0


def test_catenary_reduced(oracle_lib):
    m, p, info = M.catenary(n_elem=400)
    om = oracle_lib.OracleModel(m)
    rep, iters = om.static_solve(p)
    assert rep.aborted == 0 and rep.diverged_steps == 0 and rep.converged_steps == 10
    assert om.n_free() == info["n_free_expected"]
    assert om.bandwidth() == 17
    # equilibrium of the support reactions is not available; check the sag grew
    nd = om.nodes()
    x0 = m.freeze()["xyz"]
    assert np.abs(nd[:, :3] - x0).max() > 1e-3
