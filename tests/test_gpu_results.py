"""Element results of the last converged configuration on the GPU (gf_read_element_results,
gf_write_results) against the oracle's last Mount (Beam_1::WriteResults
src/Beam_1.cpp:444-458, Shell_1::WriteResults src/Shell_1.cpp:624-666,
PostFiles::WriteConfigurationResults src/PostFiles.cpp:1429-1464)."""
import numpy as np
import pytest

from giraffe_amd import model as M, runtime

pytestmark = pytest.mark.gpu


def _groups_close(got, exp, width, tol):
    """Each 3-vector quantity (columns 3k..3k+2) within tol of its own scale, plus an
    absolute 1e-12 of the largest entry of its half (stresses, then strains): a quantity
    that vanishes physically (the axial force under a pure end moment) holds only the
    noise the 1e-10 Newton tolerance leaves."""
    half = width // 2
    for k in range(width // 3):
        g, e = got[:, 3 * k:3 * k + 3], exp[:, 3 * k:3 * k + 3]
        part = exp[:, :half] if 3 * k < half else exp[:, half:]
        bound = tol * np.abs(e).max() + 1e-12 * np.abs(part).max()
        err = np.abs(g - e).max()
        assert err <= bound, (k, err, bound)


def _solve_both(oracle_lib, m, p):
    om = oracle_lib.OracleModel(m)
    ro, _ = om.static_solve(p)
    dm = runtime.DeviceModel(m)
    rg, _ = dm.static_solve(p)
    assert ro.aborted == rg.aborted == 0 and ro.converged_steps == rg.converged_steps
    return om, dm


@pytest.mark.parametrize("case", ["rollup", "catenary", "plate"])
def test_element_results_match_oracle(oracle_lib, case):
    if case == "rollup":
        m, p, _ = M.cantilever_rollup(n_elem=20)
        p.end_time = 0.5
    elif case == "catenary":
        m, p, _ = M.catenary(n_elem=200)
    else:
        m, p, _ = M.shell_plate(nq=4)
    print(f"case {case}", flush=True)
    om, dm = _solve_both(oracle_lib, m, p)
    bo, so = om.element_results()
    bg, sg = dm.element_results()
    assert bg.shape == bo.shape and sg.shape == so.shape
    # the nodes agree to 1e-8 (Newton stops at 1e-10 residuals): the stresses follow
    if m.n_beams:
        _groups_close(bg, bo, 12, 1e-7)
    if m.n_shells:
        _groups_close(sg, so, 24, 1e-7)


def _floats(fields):
    return np.array([float(v) for v in fields])


def test_results_file_element_lines(tmp_path):
    """Element lines in number order, %.6e, the values of gf_read_element_results, the
    reference positions and Coor = converged + incremental coordinates."""
    m, p, _ = M.shell_plate(nq=4)
    p.end_time = 3 * p.time_step
    ns = m.n_shells
    rng = np.random.default_rng(5)
    m.shell_number = rng.permutation(ns) + 1
    dm = runtime.DeviceModel(m)
    dm.static_solve(p)
    path = tmp_path / "solution_1_configuration_3.txt"
    dm.write_results(path)
    lines = path.read_text().split("\n")
    k = lines.index("RESULTS - ELEMENTS")
    el = [ln.split("\t") for ln in lines[k + 2:] if ln]
    assert len(el) == ns
    assert [int(f[1]) for f in el] == list(range(1, ns + 1))
    _, sres = dm.element_results()
    nodes, disp = dm.nodes(), dm.disp()
    x0 = m.freeze()["xyz"]
    conn = m.freeze()["shell_conn"]
    order = np.argsort(m.shell_number)
    for f, e in zip(el, order):
        assert f[0] == "Shell_1"
        assert f[2::4][:8] == ["n1", "m1", "n2", "m2", "eta1", "kappa1", "eta2", "kappa2"]
        vals = np.concatenate([_floats(f[3 + 4 * q:6 + 4 * q]) for q in range(8)])
        assert all(len(v.split("e")[0].split(".")[1]) == 6 for v in f[3:6])  # %.6e
        np.testing.assert_allclose(vals, sres[e], rtol=6e-7, atol=1e-300)
        assert f[34] == "Ref_Coor" and f[38] == "Coor"
        c = conn[e, :3]
        np.testing.assert_allclose(_floats(f[35:38]), x0[c].mean(0), rtol=6e-7, atol=1e-12)
        cur = nodes[c, :3].sum(0) / 3 + disp[c, :3].sum(0) / 3
        np.testing.assert_allclose(_floats(f[39:42]), cur, rtol=6e-7, atol=1e-12)


def test_results_file_beam_lines(tmp_path):
    m, p, info = M.cantilever_rollup(n_elem=10)
    p.end_time = 0.25
    dm = runtime.DeviceModel(m)
    dm.static_solve(p)
    path = tmp_path / "r.txt"
    dm.write_results(path)
    lines = path.read_text().split("\n")
    k = lines.index("RESULTS - ELEMENTS")
    el = [ln.split("\t") for ln in lines[k + 2:] if ln]
    bres, _ = dm.element_results()
    x0 = m.freeze()["xyz"]
    conn = m.freeze()["beam_conn"]
    nodes, disp = dm.nodes(), dm.disp()
    assert [int(f[1]) for f in el] == list(range(1, 11))
    for e, f in enumerate(el):
        assert f[0] == "Beam_1" and f[2] == "Stress_r" and f[9] == "Strain_r"
        assert f[16] == "Ref_Coor" and f[20] == "Coor"
        np.testing.assert_allclose(_floats(f[3:9]), bres[e, :6], rtol=6e-7, atol=1e-300)
        np.testing.assert_allclose(_floats(f[10:16]), bres[e, 6:], rtol=6e-7, atol=1e-300)
        c = conn[e, 1]
        np.testing.assert_allclose(_floats(f[17:20]), x0[c], rtol=6e-7, atol=1e-12)
        np.testing.assert_allclose(_floats(f[21:24]), nodes[c, :3] + disp[c, :3], rtol=6e-7, atol=1e-12)
    # the follower end moment: m_r[1] = M(t) in every element
    np.testing.assert_allclose(bres[:, 4], 0.25 * info["M"], rtol=1e-8)
