"""Mixed-precision tangent solve (GF_PREC_MIXED): the FP64 tangent is
equilibrated, factored in FP32 and the increment recovered by FP64 iterative
refinement, falling back to the FP64 factorisation when the refinement
stagnates.  Each result is checked against the FP64 device path and the
oracle to the same bar as the FP64 path."""
import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime

pytestmark = pytest.mark.gpu


def _increments(m, p, prec, states):
    dm = runtime.DeviceModel(m, precision=prec)
    dm.set_window(p.start_time, p.end_time)
    out = []
    for t_last, d in states:
        dm.begin_step(t_last, p.time_step)
        dm.write_disp(d)
        dm.assemble()
        assert dm.solve() == 0
        out.append(dm.increment().copy())
    return out, dm.stats()


def _states(m, scale, seed):
    rng = np.random.default_rng(seed)
    return [(0.0, np.zeros((m.n_nodes, 6))), (0.2, rng.normal(size=(m.n_nodes, 6)) * scale)]


@pytest.mark.parametrize("case", ["plate", "rollup"])
def test_mixed_increment_matches_fp64(case):
    if case == "plate":
        m, p, _ = M.shell_plate(nq=8)
        states = _states(m, 1e-5, 3)
    else:
        m, p, _ = M.cantilever_rollup(n_elem=12)
        states = _states(m, 1e-3, 4)
    x64, _ = _increments(m, p, abi.GF_PREC_FP64, states)
    xmx, st = _increments(m, p, abi.GF_PREC_MIXED, states)
    for a, b in zip(x64, xmx):
        # both are FP64-accurate solutions of the same FP64 system: they agree to
        # cond(K) * eps64 (cond(DKD) <= 1e7 for these models)
        assert np.abs(a - b).max() <= 1e-8 * np.abs(a).max()
    assert st["solves"] == len(states)
    assert st["refine_sweeps"] >= len(states)
    if case == "plate":
        # cond(D K D) ~ 5e5: the FP32 factor converges without falling back
        assert st["fallbacks"] == 0


def test_mixed_plate_static_matches_oracle(oracle_lib):
    m, p, _ = M.shell_plate(nq=8)
    om = oracle_lib.OracleModel(m)
    ro, io = om.static_solve(p)
    dm = runtime.DeviceModel(m, precision=abi.GF_PREC_MIXED)
    rg, ig = dm.static_solve(p)
    assert rg.aborted == ro.aborted == 0
    assert rg.converged_steps == ro.converged_steps
    assert len(ig) == len(io) and np.abs(ig - io).max() <= 1
    no, ng = om.nodes(), dm.nodes()
    x0 = m.freeze()["xyz"]
    scale = np.abs(no[:, :3] - x0).max()
    assert np.abs(ng[:, :3] - no[:, :3]).max() <= 1e-8 * scale
    assert rg.max_rel_residual <= 1e-10
    assert dm.stats()["fallbacks"] == 0


@pytest.mark.parametrize("nq", [12, 24])
def test_mixed_reused_fp32_factor(nq):
    """With refine_reuse (the default) the FP32 factor of an earlier tangent is
    kept while the refinement against the current FP64 tangent converges fast;
    every increment is still refined to the same tolerance against the current
    tangent, so the Newton history is the one of fresh factorisations and the
    nodes agree to the refinement's accuracy, with fewer FP32 factorisations
    than tangent solves."""
    m, p, _ = M.shell_plate(nq=nq, n_steps=5, p_max=4e6)
    out = {}
    for reuse in (False, True):
        dm = runtime.DeviceModel(m, precision=abi.GF_PREC_MIXED, refine_reuse=reuse)
        rep, it = dm.static_solve(p)
        assert rep.aborted == 0
        out[reuse] = (list(it), dm.nodes(), dm.stats())
    assert out[True][0] == out[False][0]
    nf, nr = out[False][1], out[True][1]
    x0 = m.freeze()["xyz"]
    assert np.abs(nr - nf).max() <= 1e-9 * np.abs(nf[:, :3] - x0).max()
    sf, sr = out[False][2], out[True][2]
    assert sf["fp32_factorizations"] + sf["fp64_direct"] == sf["solves"]
    # a kept factor is replaced when it contracts less than 20x per sweep, so
    # reuse never factors more often; on the 12 x 12 plate it keeps most factors
    assert sr["fp32_factorizations"] <= sr["solves"]
    if nq == 12:
        assert sr["fp32_factorizations"] < sr["solves"]


def test_mixed_catenary_falls_back_to_fp64(oracle_lib):
    """A long slender riser has cond(D K D) ~ 1e13 (500 elements), far beyond
    1/eps32: the FP32 factor cannot drive the refinement and every solve must
    end on the FP64 factorisation, with results identical to the FP64 path.  After
    two fallbacks in a row the path factors in FP64 directly, for 16 solves and
    then for doubling stretches while each retry falls back again; a refinement
    that does not contract at all starts the longest back-off at once."""
    m, p, _ = M.catenary(n_elem=500)
    d64 = runtime.DeviceModel(m)
    r64, i64 = d64.static_solve(p)
    dmx = runtime.DeviceModel(m, precision=abi.GF_PREC_MIXED)
    rmx, imx = dmx.static_solve(p)
    assert rmx.aborted == 0 and rmx.converged_steps == r64.converged_steps
    assert np.abs(np.asarray(imx) - np.asarray(i64)).max() <= 1
    n64, nmx = d64.nodes(), dmx.nodes()
    assert np.abs(nmx - n64).max() <= 1e-8 * np.abs(n64).max()
    st = dmx.stats()
    assert st["fallbacks"] > 0 and st["fp64_direct"] > 0
    assert st["fallbacks"] + st["fp64_direct"] == st["solves"]
    # the back-off: two fallbacks, 16 direct FP64 solves, then one retry per
    # back-off while the back-offs double (16, 32, 64, ...)
    assert st["fallbacks"] <= 2 + int(np.log2(st["solves"] / 16 + 1)) + 1


def test_mixed_singular_tangent_reports_info_fail():
    m, p, _ = M.cantilever_rollup(n_elem=4, E=0.0)
    dm = runtime.DeviceModel(m, precision=abi.GF_PREC_MIXED)
    dm.set_window(0.0, 1.0)
    dm.begin_step(0.0, 0.05)
    dm.assemble()
    assert dm.solve() == 1


def test_mixed_fronts_backs_off_when_refinement_costs_more():
    """On the multifrontal path a converged refined solve whose sweeps and FP32
    factorisations cost more than an FP64 factorisation and solve starts the
    back-off a fallback starts: later tangents are factored in FP64 directly.
    The Newton history is the FP64 one and the nodes agree to the refinement's
    accuracy; no solve falls back."""
    m, p, _ = M.shell_plate(nq=32)  # 15 load steps, each converging
    print("mixed back-off plate32", flush=True)
    d64 = runtime.DeviceModel(m, precision=abi.GF_PREC_FP64)
    r64, i64 = d64.static_solve(p)
    dmx = runtime.DeviceModel(m, precision=abi.GF_PREC_MIXED)
    assert dmx.info()["layout"] == 2
    rmx, imx = dmx.static_solve(p)
    assert rmx.aborted == r64.aborted == 0
    assert list(imx) == list(i64)
    st = dmx.stats()
    print("stats", st, flush=True)
    assert st["fallbacks"] == 0
    assert st["fp64_direct"] > 0
    # every solve is FP64-direct or refined; refined ones need at least one sweep
    assert st["refine_sweeps"] >= st["solves"] - st["fp64_direct"]
    n64, nmx = d64.nodes(), dmx.nodes()
    x0 = m.freeze()["xyz"]
    assert np.abs(nmx[:, :3] - n64[:, :3]).max() <= 1e-8 * np.abs(n64[:, :3] - x0).max()
