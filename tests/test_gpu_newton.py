"""Newton-step pieces on the GPU (gf_assemble / gf_solve / gf_update / gf_commit)
against the oracle's phases on the same model and the same states."""
import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime

pytestmark = pytest.mark.gpu


def _pair(oracle_lib, m, p):
    om = oracle_lib.OracleModel(m)
    om.begin_solution(p.start_time, p.end_time)
    dm = runtime.DeviceModel(m)
    dm.set_window(p.start_time, p.end_time)
    return om, dm


def _models():
    m1, p1, _ = M.cantilever_rollup(n_elem=12)
    m2, p2, _ = M.catenary(n_elem=40)
    # add a dead moment and a follower force so every load path is exercised
    m1.add_load(abi.GF_LOAD_NODAL, [7], [0.0, 1.0], [[0] * 6, [1e3, -2e3, 5e2, 3e4, -1e4, 2e4]])
    m1.add_load(abi.GF_LOAD_NODAL_FOLLOWER, [5, 9], [0.0, 1.0], [[0] * 6, [2e3, 1e3, -3e3, 0, 0, 0]])
    return [(m1, p1), (m2, p2)]


@pytest.mark.parametrize("which", [0, 1])
def test_assembly_matches_oracle(oracle_lib, which):
    m, p = _models()[which]
    om, dm = _pair(oracle_lib, m, p)
    rng = np.random.default_rng(which)
    for t_last, scale in [(0.0, 0.0), (0.3, 1e-3), (0.55, 3e-2)]:
        om.begin_step(t_last, 0.1)
        dm.begin_step(t_last, 0.1)
        d = rng.normal(size=(m.n_nodes, 6)) * scale
        om.write_disp(d)
        dm.write_disp(d)
        ro = om.assemble()
        rg = dm.assemble()
        Ko = om.tangent_dense()
        Kg = dm.tangent_dense()
        assert np.abs(Kg - Ko).max() <= 1e-12 * np.abs(Ko).max()
        Ro, Rg = om.residual(), dm.residual()
        scale_r = max(np.abs(Ro).max(), 1e-12 * np.abs(Ko).max())
        assert np.abs(Rg - Ro).max() <= 1e-11 * scale_r
        assert np.allclose(rg[:2], ro[:2], rtol=1e-9, atol=1e-12 * scale_r)
        assert rg[2] == ro[2] == 0.0


def test_solve_update_commit_match_oracle(oracle_lib):
    m, p = _models()[0]
    om, dm = _pair(oracle_lib, m, p)
    for step in range(3):
        om.begin_step(0.1 * step, 0.1)
        dm.begin_step(0.1 * step, 0.1)
        for it in range(4):
            om.assemble()
            dm.assemble()
            assert om.solve() == 0
            assert dm.solve() == 0
            xo, xg = om.increment(), None
            io, ig = om.update(), dm.update()
            xg = dm.increment()
            xo = om.increment()
            # two FP64 direct solvers (cyclic reduction vs pivoted band LU) agree to
            # ~cond(K) * eps; the tangent here has cond ~1e7
            assert np.abs(xg - xo).max() <= 1e-7 * np.abs(xo).max() + 1e-300
            assert np.allclose(ig[:4], io[:4], rtol=1e-8, atol=1e-15)
        om.commit()
        dm.commit()
        no, ng = om.nodes(), dm.nodes()
        assert np.abs(ng - no).max() <= 1e-10 * np.abs(no).max()


def test_singular_tangent_reports_info_fail(oracle_lib):
    # zero material stiffness: the tangent is exactly zero, every pivot is zero
    m, p, _ = M.cantilever_rollup(n_elem=4, E=0.0)
    dm = runtime.DeviceModel(m)
    dm.set_window(0.0, 1.0)
    dm.begin_step(0.0, 0.05)
    dm.assemble()
    assert dm.solve() == 1
