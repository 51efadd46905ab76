"""Pipe_1 with buoyancy below the ocean surface on the GPU against the oracle: a
partly submerged hanging pipe and a riser catenary half under water (same steps,
iterations per step within 1, nodes within 1e-8, element results)."""
import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime

pytestmark = pytest.mark.gpu


def _pipe_catenary(n_elem=200):
    """The riser catenary with its Beam_1 tubes as Pipe_1 of the same stiffness and
    weight, all under water (uniformly lighter: the same shape, less tension; a surface
    crossing the span would move Gauss points across it between iterations)."""
    m, p, info = M.catenary(n_elem=n_elem)
    De = 0.3
    for i, pr in enumerate(m.beam_prop):
        E, nu, rho, A, I11, I22, I12, It, T0 = pr
        G = E / (2 * (1 + nu))
        m.beam_prop[i] = [E * A, E * I11, G * It, G * A, rho * A, 0.0, 0.0, 0.0, 0.0]
        m.beam_kind[i] = abi.GF_BEAM_KIND_PIPE1
        m.beam_De[i] = De
    m.set_ocean(1025.0, [0.0, 0.0, 2000.0])
    m._frozen = None
    return m, p


def _hanging(n_air=3, n_elem=8, L=80.0):
    """A pipe hanging from a clamped top along -z, the surface n_air elements down."""
    m = M.Model("hanging pipe")
    t = np.linspace(0.0, L, 2 * n_elem + 1)
    m.add_nodes(np.outer(t, [0.0, 0.0, -1.0]))
    cs = M.cs_rows([1.0, 0.0, 0.0], [0.0, 0.0, -1.0])
    for e in range(n_elem):
        m.add_pipe(2 * e, 2 * e + 1, 2 * e + 2, cs, 2.0e9, 3.0e6, 2.4e6, 8.0e8, 120.0, De=0.3)
    m.fix_dofs(0, range(6))
    m.gravity, m.gravity_on = (0.0, 0.0, -9.81), 1
    m.set_ocean(1025.0, [0.0, 0.0, -L / n_elem * n_air])
    w = 120.0 * 9.81
    p = M.StaticParams(end_time=1.0, time_step=0.25, max_time_step=0.25, min_time_step=0.25 / 64, max_it=30,
                       min_it=2, conv_increase=4, inc_factor=1.0, force_tol=1e-10, moment_tol=1e-10,
                       disp_tol=1e-10, rot_tol=1e-10, force_min=w * L, moment_min=w * L * L, disp_min=1e-6,
                       rot_min=1e-6)
    return m, p


@pytest.mark.parametrize("case", ["hanging", "catenary"])
def test_pipe_static_matches_oracle(oracle_lib, case):
    m, p = _hanging() if case == "hanging" else _pipe_catenary()
    print(f"case {case}", flush=True)
    om = oracle_lib.OracleModel(m)
    ro, io = om.static_solve(p)
    dm = runtime.DeviceModel(m)
    rg, ig = dm.static_solve(p)
    assert ro.aborted == rg.aborted == 0 and ro.converged_steps == rg.converged_steps
    assert len(ig) == len(io) and np.abs(np.asarray(ig) - np.asarray(io)).max() <= 1
    no, ng = om.nodes(), dm.nodes()
    x0 = m.freeze()["xyz"]
    scale = max(np.abs(no[:, :3] - x0).max(), 1e-12)
    assert np.abs(ng[:, :3] - no[:, :3]).max() <= 1e-8 * scale
    assert np.abs(ng[:, 3:] - no[:, 3:]).max() <= 1e-8 * max(np.abs(no[:, 3:]).max(), 1e-12)
    assert rg.max_rel_residual <= 1e-10
    bo, _ = om.element_results()
    bg, _ = dm.element_results()
    assert np.abs(bg[:, 2] - bo[:, 2]).max() <= 1e-7 * np.abs(bo[:, 2]).max()


def test_pipe_results_lines(tmp_path):
    m, p = _hanging()
    dm = runtime.DeviceModel(m)
    dm.static_solve(p)
    path = tmp_path / "r.txt"
    dm.write_results(path)
    lines = path.read_text().split("\n")
    k = lines.index("RESULTS - ELEMENTS")
    el = [ln.split("\t") for ln in lines[k + 2:] if ln]
    assert len(el) == 8 and all(f[0] == "Pipe_1" and f[2] == "Stress_r" for f in el)
