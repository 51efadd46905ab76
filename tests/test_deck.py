"""GIRAFFE input decks (gf_deck_read / gf_deck_write, hot-path subset of the
reference grammar, src/IO.cpp:290-657): write -> read round trips reproduce the
description and the parameters, a deck solves exactly like the model it was
written from (oracle), a hand-written deck in the reference's own layout
(comments, Sequence sets, ignored output blocks) parses, and physics outside
the path is refused with a message."""
import ctypes as C
import os

import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    ct = {np.float64: C.c_double, np.int32: C.c_int, np.uint8: C.c_ubyte}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


def _desc_arrays(d):
    nn, nb, ns, nl = d.n_nodes, d.n_beams, d.n_shells, d.n_loads
    off = _arr(d.load_set_off, nl + 1, np.int32)
    toff = _arr(d.load_tab_off, nl + 1, np.int32)
    rows = int(toff[-1]) if nl else 0
    return {
        "xyz": _arr(d.node_xyz, 3 * nn, np.float64), "fix": _arr(d.node_fix, 6 * nn, np.uint8),
        "beam_conn": _arr(d.beam_conn, 3 * nb, np.int32), "beam_cs": _arr(d.beam_cs, 9 * nb, np.float64),
        "beam_prop": _arr(d.beam_prop, abi.GF_BEAM_NPROP * nb, np.float64),
        "shell_conn": _arr(d.shell_conn, 6 * ns, np.int32),
        "shell_prop": _arr(d.shell_prop, abi.GF_SHELL_NPROP * ns, np.float64),
        "load_kind": _arr(d.load_kind, nl, np.int32), "load_set_off": off,
        "load_set_items": _arr(d.load_set_items, int(off[-1]) if nl else 0, np.int32),
        "load_cs": _arr(d.load_cs, 9 * nl, np.float64), "load_tab_off": toff,
        "tab_time": _arr(d.tab_time, rows, np.float64), "tab_val": _arr(d.tab_val, 6 * rows, np.float64),
        "gravity": np.array(list(d.gravity)), "gravity_on": np.array([d.gravity_on]),
    }


def _models():
    return [M.catenary(n_elem=30)[:2], M.shell_plate(nq=4)[:2], M.cantilever_rollup(n_elem=10)[:2]]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_round_trip(tmp_path, which):
    m, p = _models()[which]
    path = str(tmp_path / "deck.inp")
    runtime.write_deck(path, m, p)
    deck = runtime.Deck(path)
    a, b = _desc_arrays(m.desc()), _desc_arrays(deck.desc())
    for k in a:
        if a[k].dtype.kind == "f":
            # Tube properties are recomputed from De / Di: a few ulps
            assert np.allclose(a[k], b[k], rtol=1e-14, atol=0), k
        else:
            assert np.array_equal(a[k], b[k]), k
    for f in p.__dataclass_fields__:
        assert getattr(deck.params, f) == pytest.approx(getattr(p, f), rel=1e-15), f


@pytest.mark.parametrize("which", [1, 2])
def test_deck_solves_like_its_model(tmp_path, oracle_lib, which):
    m, p = _models()[which]
    path = str(tmp_path / "deck.inp")
    runtime.write_deck(path, m, p)
    deck = runtime.Deck(path)
    ra, ia = oracle_lib.OracleModel(m).static_solve(p)
    od = oracle_lib.OracleModel(deck)
    rb, ib = od.static_solve(deck.params)
    assert list(ia) == list(ib)
    om = oracle_lib.OracleModel(m)
    om.static_solve(p)
    na, nb = om.nodes(), od.nodes()
    assert np.abs(na - nb).max() <= 1e-12 * np.abs(na).max()


def test_element_numbers_round_trip(tmp_path, oracle_lib):
    """Element numbers travel through the deck (elements written in number order, shell
    load sets by number); the re-read deck solves like the model."""
    m, p = M.shell_plate(nq=4)[:2]
    p.end_time = 2 * p.time_step
    ns = m.n_shells
    m.shell_number = np.arange(ns, 0, -1)
    path = str(tmp_path / "deck.inp")
    runtime.write_deck(path, m, p)
    lines = [ln.split("\t") for ln in open(path).read().split("\n") if ln.startswith("Shell_1\t")]
    assert [int(f[1]) for f in lines] == list(range(1, ns + 1))
    deck = runtime.Deck(path)
    d = deck.desc()
    assert list(_arr(d.shell_number, ns, np.int32)) == list(range(1, ns + 1))
    a, b = _desc_arrays(m.desc()), _desc_arrays(deck.desc())
    assert np.array_equal(b["shell_conn"].reshape(-1, 6), a["shell_conn"].reshape(-1, 6)[::-1])
    om, od = oracle_lib.OracleModel(m), oracle_lib.OracleModel(deck)
    ia, ib = om.static_solve(p)[1], od.static_solve(deck.params)[1]
    assert list(ia) == list(ib)
    assert np.abs(om.nodes() - od.nodes()).max() <= 1e-12 * np.abs(om.nodes()).max()


PIPE_DECK = """\
Nodes\t5
Node\t1\t0\t0\t0
Node\t2\t0\t0\t-5
Node\t3\t0\t0\t-10
Node\t4\t0\t0\t-15
Node\t5\t0\t0\t-20
NodeSets\t1
NodeSet\t1\tNodes\t1\tList\t1
Elements\t2
Pipe_1\t1\tPipeSec\t1\tCS\t1\tNodes\t1\t2\t3
Pipe_1\t2\tPipeSec\t1\tCS\t1\tNodes\t3\t4\t5
PipeSections\t1
PS\t1\tEA\t2e9\tEI\t3e6\tGJ\t2.4e6\tGA\t8e8\tRho\t120\tCDt\t0.1\tCDn\t1.2\tCAt\t0\tCAn\t1\tDe\t0.3\tDi\t0.2
CoordinateSystems\t1
CS\t1\tE1\t1\t0\t0\tE3\t0\t0\t-1
Constraints\t1
NodalConstraint\t1\tNodeSet\t1
\tUX\t\tBoolTable\t1
\tUY\t\tBoolTable\t1
\tUZ\t\tBoolTable\t1
\tROTX\tBoolTable\t1
\tROTY\tBoolTable\t1
\tROTZ\tBoolTable\t1
Environment
GravityData\tG\t0\t0\t-9.81\tBoolTable\t1
OceanData
RhoFluid\t1025
SurfacePosition\t0\t0\t-10
SeaCurrent\tN\t2\tBoolTable\t0
Depth\t0\tSpeed\t1.5\tAngle\t0
Depth\t100\tSpeed\t0.5\tAngle\t0
SolutionSteps\t1
Static\t1
EndTime\t1
TimeStep\t0.25
MaxTimeStep\t0.25
MinTimeStep\t0.01
MaxIt\t20
MinIt\t3
ConvIncrease\t4
IncFactor\t1.0
Sample\t1
"""


def test_pipe_deck(tmp_path, oracle_lib):
    """Pipe_1 / PipeSections / OceanData (an inactive sea current): the description
    carries the pipe section's stiffness and the buoyant mass per length; the deck
    writes back and solves alike."""
    path = tmp_path / "pipe.inp"
    path.write_text(PIPE_DECK)
    deck = runtime.Deck(str(path))
    d = deck.desc()
    assert deck.n_beams == 2 and d.ocean_on == 1 and list(d.ocean_surface) == [0, 0, -10]
    assert list(_arr(d.beam_kind, 2, np.int32)) == [abi.GF_BEAM_KIND_PIPE1] * 2
    np.testing.assert_allclose(_arr(d.beam_buoy, 2, np.float64), 1025 * np.pi * 0.09 / 4, rtol=1e-15)
    a = _desc_arrays(d)
    assert list(a["beam_prop"][:5]) == [2e9, 3e6, 2.4e6, 8e8, 120.0]
    p2 = str(tmp_path / "again.inp")
    runtime.write_deck(p2, deck, deck.params)
    again = runtime.Deck(p2)
    b = _desc_arrays(again.desc())
    for k in a:
        assert np.allclose(a[k], b[k], rtol=1e-14, atol=0), k
    ra, ia = oracle_lib.OracleModel(deck).static_solve(deck.params)
    rb, ib = oracle_lib.OracleModel(again).static_solve(again.params)
    assert ra.aborted == 0 and list(ia) == list(ib)
    moving = PIPE_DECK.replace("SeaCurrent\tN\t2\tBoolTable\t0", "SeaCurrent\tN\t2\tBoolTable\t1")
    path.write_text(moving)
    with pytest.raises(RuntimeError, match="Morison"):
        runtime.Deck(str(path))


TUTORIAL_STYLE = """\
/////////////////////////////////
// cantilever, reference deck layout
/////////////////////////////////
Nodes\t5
//Number\tX\tY\tZ
Node\t1\t0\t0\t0.0
Node\t2\t0\t0\t0.25
Node\t3\t0\t0\t0.5
Node\t4\t0\t0\t0.75
Node\t5\t0\t0\t1.0
NodeSets\t2
NodeSet\t1\tNodes\t1\tList\t1
NodeSet\t2\tNodes\t2\tSequence\tInitial\t4\tIncrement\t1
Elements\t2
Beam_1\t1\tMat\t1\tSec\t1\tCS\t1\tNodes\t1\t2\t3
Beam_1\t2\tMat\t1\tSec\t1\tCS\t1\tNodes\t3\t4\t5\tPreTension\t10.0
Materials\t1
Hooke\t1\tE\t2.1e11\tNu\t0.3\tRho\t7850
Sections\t1
Tube\t1\tDe\t0.1\tDi\t0.08
CoordinateSystems\t1
CS\t1\tE1\t1\t0\t0\tE3\t0\t0\t1
SolutionSteps\t1
Static\t1
EndTime\t1
TimeStep\t0.1
MaxTimeStep\t0.1
MinTimeStep\t0.01
MaxIt\t20
MinIt\t3
ConvIncrease\t4
IncFactor\t1.0
Sample\t2
Loads\t1
NodalLoad\t1\tNodeSet\t2\tCS\t1\tNTimes\t2
//Time FX FY FZ MX MY MZ
0\t0\t0\t0\t0\t0\t0
1\t1000\t0\t0\t0\t0\t0
Constraints\t1
NodalConstraint\t1\tNodeSet\t1
\tUX\t\tBoolTable\t1
\tUY\t\tBoolTable\t1
\tUZ\t\tBoolTable\t1
\tROTX\tBoolTable\t1
\tROTY\tBoolTable\t1
\tROTZ\tBoolTable\t0 1
Environment
GravityData\tG\t0\t0\t-9.81\tBoolTable\t1
SolverOptions
Processors\t1\tLinSys\tDirect
Monitor\tSample\t1
MonitorNodes\t1\t5
PostFiles
MagFactor\t1
WriteMesh\t1
"""


def test_reference_layout_deck(tmp_path):
    path = tmp_path / "t.inp"
    path.write_text(TUTORIAL_STYLE)
    deck = runtime.Deck(str(path))
    a = _desc_arrays(deck.desc())
    assert deck.n_nodes == 5 and deck.n_beams == 2
    assert list(a["fix"][:6]) == [1, 1, 1, 1, 1, 0]   # ROTZ's table is 0 for the first step
    assert a["fix"][6:].sum() == 0
    assert list(a["load_set_items"]) == [3, 4]         # Sequence Initial 4 Increment 1, 0-based
    assert a["beam_prop"][abi.GF_BEAM_NPROP + 8] == 10.0
    A = np.pi / 4 * (0.1 ** 2 - 0.08 ** 2)
    assert a["beam_prop"][3] == pytest.approx(A, rel=1e-15)
    assert list(a["gravity"]) == [0, 0, -9.81] and a["gravity_on"][0] == 1
    assert deck.params.max_it == 20 and deck.params.min_it == 3 and deck.params.time_step == 0.1
    assert deck.params.force_tol == 1e-4  # ConvergenceCriteria defaults


@pytest.mark.parametrize("edit,msg", [
    (("Tube\t1\tDe\t0.1\tDi\t0.08", "SuperEllipse\t1\tA\t0.1\tB\t0.1\tN\t2"), "SuperEllipse"),
    (("SolverOptions", "Contacts\t0\nSolverOptions"), "Contacts"),
    (("SolutionSteps\t1\nStatic\t1", "SolutionSteps\t1\nDynamic\t1"), "Dynamic"),
    (("NodalLoad\t1\tNodeSet\t2\tCS\t1\tNTimes\t2", "NodalLoad\t1\tNodeSet\t2\tCS\t1\tMathCode"), "MathCode"),
])
def test_refusals(tmp_path, edit, msg):
    path = tmp_path / "t.inp"
    path.write_text(TUTORIAL_STYLE.replace(*edit))
    with pytest.raises(runtime.GiraffeError, match=msg):
        runtime.Deck(str(path))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reissner_cantilever(E, nu, b, h, L, P):
    """Tip (x, z, rotation) of a geometrically exact planar beam (extension and
    shear with shear factor 1, as Beam_1's D), clamped at s = 0 along z, under a
    dead tip force P along x: theta'' = -(P / EI) z', integrated by shooting on
    theta'(0) so that the tip moment vanishes."""
    from scipy.integrate import solve_ivp
    from scipy.optimize import brentq

    A, I, G = b * h, b * h ** 3 / 12.0, E / (2.0 * (1.0 + nu))

    def rhs(s, y):
        th, k = y[0], y[1]
        eps, gam = P * np.sin(th) / (E * A), P * np.cos(th) / (G * A)
        rx = (1 + eps) * np.sin(th) + gam * np.cos(th)
        rz = (1 + eps) * np.cos(th) - gam * np.sin(th)
        return [k, -P * rz / (E * I), rx, rz]

    def shoot(k0):
        return solve_ivp(rhs, [0, L], [0, k0, 0, 0], rtol=1e-12, atol=1e-14).y[:, -1]

    k0 = brentq(lambda k: shoot(k)[1], 1e-3, 50.0)
    th, _, x, z = shoot(k0)
    return x, z, th


def test_tutorial01_deck_matches_the_exact_beam(oracle_lib):
    """The reference's inputs/tutorial01 read unchanged (a fixture copied by
    tests/golden/copy_tutorial_decks.py): five Beam_1 on a 1 m cantilever with a
    0.1 x 0.1 Rectangle section, E = 1e7, a dead tip load of 1,000 N over ten
    steps (PL^2/EI = 12, tip rotation ~84 degrees).  The section's properties
    match the closed forms (square torsion constant 0.1406 a^4), and the oracle's
    tip matches an independent integration of the exact beam to the five
    elements' discretisation error (< 1e-3)."""
    deck = runtime.Deck(os.path.join(GOLDEN, "tutorial01.inp"))
    a = _desc_arrays(deck.desc())
    assert deck.n_beams == 5 and deck.n_nodes == 11
    pr = a["beam_prop"][:abi.GF_BEAM_NPROP]
    assert pr[3] == pytest.approx(0.01, rel=1e-15)
    assert pr[4] == pytest.approx(0.1 ** 4 / 12, rel=1e-14) and pr[5] == pr[4] and pr[6] == 0.0
    assert pr[7] == pytest.approx(0.1406 * 0.1 ** 4, rel=1e-3)
    assert list(a["fix"][:6]) == [1] * 6 and a["fix"][6:].sum() == 0
    om = oracle_lib.OracleModel(deck, oracle_lib.SOLVER_BAND)
    ro, io = om.static_solve(deck.params)
    assert ro.aborted == 0 and ro.converged_steps == 10
    tip = om.nodes()[-1]
    x, z, th = reissner_cantilever(1e7, 0.3, 0.1, 0.1, 1.0, 1000.0)
    assert abs(tip[0] - x) < 1e-3 * x and abs(tip[2] - z) < 1e-3 * z
    assert abs(tip[4] - th) < 1e-3 * th and abs(tip[1]) < 1e-12


def test_rectangle_section_round_trip(tmp_path):
    """A B x H Rectangle (B != H) written back from its properties is the same section."""
    text = open(os.path.join(GOLDEN, "tutorial01.inp")).read().replace("B\t0.1\tH\t0.1", "B\t0.05\tH\t0.2")
    path = tmp_path / "r.inp"
    path.write_text(text)
    deck = runtime.Deck(str(path))
    pr = _desc_arrays(deck.desc())["beam_prop"][:abi.GF_BEAM_NPROP]
    assert pr[4] == pytest.approx(0.05 * 0.2 ** 3 / 12, rel=1e-14)
    assert pr[5] == pytest.approx(0.2 * 0.05 ** 3 / 12, rel=1e-14)
    p2 = str(tmp_path / "r2.inp")
    runtime.write_deck(p2, deck, deck.params)
    assert "Rectangle" in open(p2).read()
    again = runtime.Deck(p2)
    np.testing.assert_allclose(_desc_arrays(again.desc())["beam_prop"], _desc_arrays(deck.desc())["beam_prop"],
                               rtol=1e-12)


def test_multi_step_decks_parse_with_their_gravity_modes():
    """Several Static steps run in sequence (src/main.cpp:53-62): start times follow
    the previous end times (Database::EvaluateStartEndTimes); the gravity factor of
    a step follows the BoolTable's entries for it and the step before
    (src/BoolTable.cpp:84-109): tutorial07's table 0 0 1 1 1 0 gives off, off,
    ramp up, on, on, ramp down."""
    d7 = runtime.Deck(os.path.join(GOLDEN, "tutorial07.inp"))
    assert [m for _, m in d7.steps] == [abi.GF_GRAVITY_OFF, abi.GF_GRAVITY_OFF, abi.GF_GRAVITY_RAMP_UP,
                                        abi.GF_GRAVITY_ON, abi.GF_GRAVITY_ON, abi.GF_GRAVITY_RAMP_DOWN]
    assert [(q.start_time, q.end_time) for q, _ in d7.steps] == [(i, i + 1.0) for i in range(6)]
    assert d7.desc().gravity_on == 1 and list(d7.desc().gravity) == [9.81, 0.0, 0.0]
    d3 = runtime.Deck(os.path.join(GOLDEN, "tutorial03.inp"))
    assert [(q.time_step, q.min_time_step) for q, _ in d3.steps] == [(1.0, 0.01), (0.005, 0.002)]
    assert d3.steps[0][0].force_tol == d3.steps[1][0].force_tol  # one ConvergenceCriteria for all steps
    for name in ("tutorial08", "tutorial09"):
        assert len(runtime.Deck(os.path.join(GOLDEN, name + ".inp")).steps) in (2, 3)


def test_changing_constraints_are_refused(tmp_path):
    text = open(os.path.join(GOLDEN, "tutorial07.inp")).read().replace("UX\t\tBoolTable\t1", "UX\t\tBoolTable\t1 0")
    path = tmp_path / "c.inp"
    path.write_text(text)
    with pytest.raises(runtime.GiraffeError, match="change between solution steps"):
        runtime.Deck(str(path))


@pytest.mark.parametrize("name", ["tutorial07", "tutorial08"])
def test_unloaded_end_state(oracle_lib, name):
    """tutorial07 (tip force, tip moment, then gravity switched on, the loads taken
    off, gravity ramped off) and tutorial08 (loads on, then off) end with no load
    at all: the oracle's cantilever returns to its undeformed configuration, to the
    Newton tolerance (1e-3 of the 1 m length; the loads deflect the tip by about
    half a metre on the way)."""
    deck = runtime.Deck(os.path.join(GOLDEN, name + ".inp"))
    om = oracle_lib.OracleModel(deck, oracle_lib.SOLVER_BAND)
    peak = []
    for q, mode in deck.steps:
        om.set_gravity_mode(mode)
        rep, _ = om.static_solve(q)
        assert rep.aborted == 0
        peak.append(np.abs(om.nodes()[:, :3] - np.array(deck.desc().node_xyz[:3 * deck.n_nodes]).reshape(-1, 3)).max())
    assert max(peak) > 0.1
    assert peak[-1] < 1e-3
