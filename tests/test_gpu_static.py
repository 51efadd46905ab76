"""Whole Static::Solve on the GPU (gf_static_solve) against the oracle:
same Newton iteration counts per step (+-1), node results within 1e-8."""
import numpy as np
import pytest

from giraffe_amd import model as M, runtime

pytestmark = pytest.mark.gpu


def _compare(oracle_lib, m, p, node_tol=1e-8):
    om = oracle_lib.OracleModel(m)
    ro, io = om.static_solve(p)
    dm = runtime.DeviceModel(m)
    rg, ig = dm.static_solve(p)
    assert rg.aborted == ro.aborted == 0
    assert rg.converged_steps == ro.converged_steps
    assert len(ig) == len(io)
    assert np.abs(ig - io).max() <= 1, (ig, io)
    no, ng = om.nodes(), dm.nodes()
    x0 = m.freeze()["xyz"]
    scale = max(np.abs(no[:, :3] - x0).max(), 1e-12)
    err_u = np.abs(ng[:, :3] - no[:, :3]).max() / scale
    err_r = np.abs(ng[:, 3:] - no[:, 3:]).max() / max(np.abs(no[:, 3:]).max(), 1e-12)
    assert err_u <= node_tol and err_r <= node_tol, (err_u, err_r)
    assert rg.max_rel_residual <= 1e-10
    return ro, rg


def test_config1_rollup(oracle_lib):
    m, p, info = M.cantilever_rollup()
    ro, rg = _compare(oracle_lib, m, p)
    dm = runtime.DeviceModel(m)
    dm.static_solve(p)
    tip = dm.nodes()[info["tip"]]
    assert np.abs(tip[:3]).max() < 1e-9 * info["L"]  # full circle: back at the root


def test_catenary_reduced(oracle_lib):
    m, p, _ = M.catenary(n_elem=500)
    _compare(oracle_lib, m, p)


def test_catenary_full_size(oracle_lib):
    """Config 2 at its full size: 10,000 Beam_1, 120k DOFs."""
    m, p, _ = M.catenary(n_elem=10000)
    _compare(oracle_lib, m, p)


def test_results_file_layout(oracle_lib, tmp_path):
    """gf_write_results: the node section of PostFiles::WriteConfigurationResults
    (Node::WriteResults lines, %.6e) for the converged configuration."""
    m, p, info = M.cantilever_rollup()
    dm = runtime.DeviceModel(m)
    dm.static_solve(p)
    path = tmp_path / "solution_1_configuration_1.txt"
    dm.write_results(path)
    lines = path.read_text().split("\n")
    assert lines[0] == "RESULTS - NODES" and lines[1] == ""
    nodes = dm.nodes()
    for i in range(m.n_nodes):
        f = lines[2 + i].split("\t")
        assert f[0] == "Node" and int(f[1]) == i + 1
        assert f[2::2] == ["X", "Y", "Z", "RX", "RY", "RZ"]
        vals = np.array([float(v) for v in f[3::2]])
        assert all(len(v.split("e")[0].split(".")[1]) == 6 for v in f[3::2])  # %.6e
        np.testing.assert_allclose(vals, nodes[i], rtol=5e-7, atol=1e-300)
    tail = lines[2 + m.n_nodes:]
    assert tail[:4] == ["", "", "RESULTS - ELEMENTS", ""]
    # against the oracle's converged configuration at the file's precision
    om = oracle_lib.OracleModel(m)
    om.static_solve(p)
    no = om.nodes()
    got = np.array([[float(v) for v in lines[2 + i].split("\t")[3::2]] for i in range(m.n_nodes)])
    scale = np.abs(no).max()
    assert np.abs(got - no).max() <= 1e-6 * scale
