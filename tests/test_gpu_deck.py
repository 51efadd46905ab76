"""The reference's tutorial01 deck, read unchanged, solved on the device: the
same steps and Newton iterations as the oracle, nodes within 1e-8."""
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
