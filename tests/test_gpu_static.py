"""Whole Static::Solve on the GPU (gf_static_solve) against the oracle:
same Newton iteration counts per step (+-1), node results within 1e-8."""
import numpy as np
import pytest

from giraffe_amd import model as M, runtime

pytestmark = pytest.mark.gpu


def oracle_solver(oracle_lib):
    """The reference's own solver (MKL PARDISO, called as sparsesystem does) where
    libmkl_rt is present, else the oracle's band LU."""
    return oracle_lib.SOLVER_PARDISO if oracle_lib.pardiso_available() else oracle_lib.SOLVER_BAND


def _compare(oracle_lib, m, p, node_tol=1e-8):
    om = oracle_lib.OracleModel(m, oracle_solver(oracle_lib))
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
    om = oracle_lib.OracleModel(m, oracle_solver(oracle_lib))
    om.static_solve(p)
    no = om.nodes()
    got = np.array([[float(v) for v in lines[2 + i].split("\t")[3::2]] for i in range(m.n_nodes)])
    scale = np.abs(no).max()
    assert np.abs(got - no).max() <= 1e-6 * scale


def _branch_compare(oracle_lib, m, p):
    """Same step sequence (iterations per attempt, negative for bisected attempts),
    counts, end time and nodes on the device as in the oracle."""
    om = oracle_lib.OracleModel(m, oracle_solver(oracle_lib))
    ro, io = om.static_solve(p)
    dm = runtime.DeviceModel(m)
    rg, ig = dm.static_solve(p)
    ig, io = np.asarray(ig), np.asarray(io)
    assert len(ig) == len(io) and np.array_equal(np.sign(ig), np.sign(io)), (ig, io)
    assert np.abs(ig - io).max() <= 1, (ig, io)
    for k in ("aborted", "converged_steps", "diverged_steps"):
        assert getattr(rg, k) == getattr(ro, k), k
    assert rg.final_time == ro.final_time
    no, ng = om.nodes(), dm.nodes()
    assert np.abs(ng - no).max() <= 1e-8 * np.abs(no).max()
    return rg, ig


def test_bisection_and_abort_on_device(oracle_lib):
    """MaxIt too small: every attempt diverges, gf_rollback restores the step and
    the step is halved (src/Static.cpp:238-242, 286-294) until it falls below
    MinTimeStep; the aborted attempt runs one iteration and is committed
    (src/Static.cpp:244-256)."""
    m, p, _ = M.cantilever_rollup(n_elem=10)
    p.max_it = 3
    p.min_time_step = p.time_step / 8
    rg, ig = _branch_compare(oracle_lib, m, p)
    assert rg.aborted == 1 and list(ig) == [-3, -3, -3, -3, 1]


def test_bisection_then_recovery_on_device(oracle_lib):
    """A step too large to converge in MaxIt iterations is bisected, and the halved
    steps converge: rollback leaves the committed state untouched."""
    m, p, _ = M.cantilever_rollup(n_elem=10)
    p.time_step = p.max_time_step = 1.0
    p.min_time_step = 1.0 / 64
    p.max_it = 12  # the oracle: four bisections, then 16 steps of 9 iterations
    rg, ig = _branch_compare(oracle_lib, m, p)
    assert rg.aborted == 0 and rg.diverged_steps >= 1 and min(ig) < 0
    assert rg.final_time == 1.0


def test_step_growth_on_device(oracle_lib):
    """IncFactor growth when a step converges within MinIt (src/Static.cpp:262-268):
    t = 0.05, 0.15, 0.35, 0.55, 0.75, 0.95, 1.0."""
    m, p, _ = M.cantilever_rollup(n_elem=10)
    p.time_step = 0.05
    p.max_time_step = 0.2
    p.inc_factor = 2.0
    p.min_it = 100
    rg, ig = _branch_compare(oracle_lib, m, p)
    assert rg.converged_steps == 7 and rg.final_time == 1.0


def test_rollback_zeroes_increments():
    """RestoreConfiguration (src/Solution.cpp:532-574): the incremental displacements
    return to zero and the converged configuration is unchanged."""
    m, p, _ = M.cantilever_rollup(n_elem=6)
    dm = runtime.DeviceModel(m)
    dm.set_window(0.0, 1.0)
    dm.begin_step(0.0, 0.05)
    dm.assemble()
    assert dm.solve() == 0
    dm.update()
    before = dm.nodes()
    assert np.abs(dm.disp()).max() > 0.0
    dm.rollback()
    assert np.abs(dm.disp()).max() == 0.0
    assert np.array_equal(dm.nodes(), before)
