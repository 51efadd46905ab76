"""Pins the oracle's element results -- the Gauss-point means that Beam_1::WriteResults
(src/Beam_1.cpp:444-458) and Shell_1::WriteResults (src/Shell_1.cpp:624-666) print --
by closed forms: a follower end moment gives a uniform internal moment and curvature
M/EI with no force, an axial end force gives n = F and strain F/EA, and at small
strain the shell's hyperelastic law reduces to plane-stress membrane stiffness
E t / (1 - nu^2), plate bending stiffness E t^3 / (12 (1 - nu^2)) and the drilling
stiffness E t^3."""
import numpy as np
import pytest

from giraffe_amd import abi
from giraffe_amd import model as M


def test_beam_follower_moment_results(oracle_lib):
    m, p, info = M.cantilever_rollup(n_elem=20)
    p.end_time = 0.5
    om = oracle_lib.OracleModel(m)
    rep, _ = om.static_solve(p)
    assert rep.aborted == 0 and rep.diverged_steps == 0
    b, s = om.element_results()
    assert b.shape == (20, 12) and s.shape == (0, 24)
    Mt, EI, L = 0.5 * info["M"], info["EI"], info["L"]
    # local E2 is global z: the moment sits in m_r[1], the curvature in kappa_r[1]
    np.testing.assert_allclose(b[:, 4], Mt, rtol=1e-8)
    np.testing.assert_allclose(b[:, 10], Mt / EI, rtol=1e-8)
    assert np.abs(b[:, [3, 5]]).max() <= 1e-10 * Mt
    assert np.abs(b[:, :3]).max() <= 1e-10 * Mt / L
    assert np.abs(b[:, [6, 7, 8, 9, 11]]).max() <= 1e-12


def test_beam_axial_force_results(oracle_lib):
    m, p, info = M.cantilever_rollup(n_elem=10)
    m.loads.clear()
    E = 210e9
    A = M.tube_section(0.2, 0.18)[0]
    F = 1e6
    m.add_load(abi.GF_LOAD_NODAL, [info["tip"]], [0.0, 1.0], [[0] * 6, [F, 0, 0, 0, 0, 0]])
    p.end_time, p.force_min, p.moment_min = 0.2, F, F * info["L"]
    om = oracle_lib.OracleModel(m)
    rep, _ = om.static_solve(p)
    assert rep.aborted == 0
    b, _ = om.element_results()
    Ft = 0.2 * F
    # local E3 is the bar axis: n_r[2] = F, eta_r[2] = F / EA
    np.testing.assert_allclose(b[:, 2], Ft, rtol=1e-9)
    np.testing.assert_allclose(b[:, 8], Ft / (E * A), rtol=1e-9)
    assert np.abs(b[:, [0, 1]]).max() <= 1e-9 * Ft
    assert np.abs(b[:, 3:6]).max() <= 1e-9 * Ft * info["L"]


E, NU, T = 70e9, 0.33, 0.005
PROP = [E, NU, 2700.0, T]
# corners 1-3 counter-clockwise in the x-y plane (normal +z, e1 = global X), mid-sides 4-6
X18 = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0.0]) * 0.1


def _shell_results(oracle_lib, d36):
    el = oracle_lib.ShellElement(X18, PROP)
    return el.mount(d36, results=True)[2].reshape(8, 3)  # n1 m1 n2 m2 eta1 k1 eta2 k2


def test_shell_membrane_results(oracle_lib):
    eps = 1e-6
    d = np.zeros(36)
    for a in range(6):
        d[6 * a] = eps * X18[3 * a]  # u_x = eps x
    n1, m1, n2, m2, e1, k1, e2, k2 = _shell_results(oracle_lib, d)
    Cm = E * T / (1 - NU ** 2)
    assert n1[0] == pytest.approx(Cm * eps, rel=1e-5)
    assert n2[1] == pytest.approx(NU * Cm * eps, rel=1e-5)
    assert abs(n1[1]) + abs(n2[0]) <= 1e-9 * Cm * eps
    assert e1[0] == pytest.approx(eps, rel=1e-12) and abs(e2[1]) <= 1e-18
    assert np.abs(np.concatenate([m1, m2, k1, k2])).max() <= 1e-9 * Cm * eps * T


def test_shell_bending_and_drilling_results(oracle_lib):
    kap, kd = 1e-6, 2e-6
    d = np.zeros(36)
    for a in range(3, 6):  # rotations live at the mid-side nodes
        d[6 * a + 4] = kap * X18[3 * a]  # alpha_y = kap x: curvature kappa_r1 = (0, kap, 0)
        d[6 * a + 5] = kd * X18[3 * a]   # alpha_z = kd x: drilling curvature kappa_r1,z = kd
    n1, m1, n2, m2, e1, k1, e2, k2 = _shell_results(oracle_lib, d)
    Db = E * T ** 3 / (12 * (1 - NU ** 2))
    assert k1[1] == pytest.approx(kap, rel=1e-10) and k1[2] == pytest.approx(kd, rel=1e-10)
    assert m1[1] == pytest.approx(Db * kap, rel=1e-5)
    assert m2[0] == pytest.approx(-NU * Db * kap, rel=1e-5)
    assert m1[2] == pytest.approx(E * T ** 3 * kd, rel=1e-12)
    assert abs(m1[0]) + abs(m2[1]) <= 1e-6 * Db * kap  # second order in the rotations
