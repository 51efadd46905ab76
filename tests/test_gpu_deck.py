"""The reference's tutorial decks that the static subset reads, unchanged, solved
on the device against the oracle: the same steps and Newton iterations, nodes
within 1e-8 (of the displacement's size, or of the model's size when the model
ends unloaded)."""
import os

import numpy as np
import pytest

from giraffe_amd import runtime

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_tutorial01_on_device(oracle_lib):
    print("tutorial01 deck on the device", flush=True)
    deck = runtime.Deck(os.path.join(GOLDEN, "tutorial01.inp"))
    om = oracle_lib.OracleModel(deck, oracle_lib.SOLVER_BAND)
    ro, io = om.static_solve(deck.params)
    dm = runtime.DeviceModel(deck)
    rep, it = dm.static_solve(deck.params)
    assert rep.aborted == ro.aborted == 0 and rep.converged_steps == ro.converged_steps == 10
    assert list(it) == list(io)
    no, nd = om.nodes(), dm.nodes()
    x0 = np.array(deck.desc().node_xyz[:3 * deck.n_nodes]).reshape(-1, 3)
    scale_u = np.abs(no[:, :3] - x0).max()
    assert np.abs(nd[:, :3] - no[:, :3]).max() <= 1e-8 * scale_u
    assert np.abs(nd[:, 3:] - no[:, 3:]).max() <= 1e-8 * np.abs(no[:, 3:]).max()


def _compare(deck, out_o, out_d):
    assert len(out_d) == len(out_o)
    for (ro, io), (rd, idd) in zip(out_o, out_d):
        assert rd.aborted == ro.aborted == 0 and rd.converged_steps == ro.converged_steps
        io, idd = np.asarray(io), np.asarray(idd)
        assert len(io) == len(idd) and np.array_equal(np.sign(io), np.sign(idd))
        conv = io > 0
        assert np.abs(io[conv] - idd[conv]).max(initial=0) <= 1


@pytest.mark.parametrize("name", ["tutorial03", "tutorial07", "tutorial08", "tutorial09"])
def test_multi_step_deck_on_device(oracle_lib, name):
    """The reference's multi-step static decks, read unchanged, through
    gf_static_solve_steps (gravity switched on, kept and ramped off across
    tutorial07's steps): per step the same converged steps and Newton iterations
    within one as the oracle, nodes within 1e-8 at the end."""
    print(f"{name} deck on the device", flush=True)
    deck = runtime.Deck(os.path.join(GOLDEN, name + ".inp"))
    om = oracle_lib.OracleModel(deck, oracle_lib.SOLVER_BAND)
    out_o = om.static_solve_steps(deck.steps)
    dm = runtime.DeviceModel(deck)
    out_d = dm.static_solve_steps(deck.steps)
    _compare(deck, out_o, out_d)
    no, nd = om.nodes(), dm.nodes()
    x0 = np.array(deck.desc().node_xyz[:3 * deck.n_nodes]).reshape(-1, 3)
    scale_u = max(np.abs(no[:, :3] - x0).max(), np.ptp(x0, axis=0).max())
    scale_r = max(np.abs(no[:, 3:]).max(), 1.0)
    assert np.abs(nd[:, :3] - no[:, :3]).max() <= 1e-8 * scale_u
    assert np.abs(nd[:, 3:] - no[:, 3:]).max() <= 1e-8 * scale_r
