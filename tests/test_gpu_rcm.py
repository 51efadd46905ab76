"""Reverse Cuthill-McKee ordering on the device path: a clamped Shell_1 plate
numbered like inputs/tutorial05 (mid-side nodes after every corner) solves to
the same nodes as the oracle on the generator's numbering, with a half-band
within 5 % of the generator's."""
import numpy as np
import pytest

from giraffe_amd import model as M, runtime
from renumber import corners_first, renumber

pytestmark = pytest.mark.gpu


def test_tutorial05_numbering_solves_like_the_generator(oracle_lib):
    nq = 12
    m, p, _ = M.shell_plate(nq=nq, n_steps=6, p_max=6e6)
    new_of_old = corners_first(nq)
    r = renumber(m, new_of_old)
    solver = oracle_lib.SOLVER_PARDISO if oracle_lib.pardiso_available() else oracle_lib.SOLVER_BAND
    om = oracle_lib.OracleModel(m, solver)
    ro, io = om.static_solve(p)
    dm = runtime.DeviceModel(r)
    info = dm.info()
    assert info["half_band"] <= 1.05 * runtime.plan(m)["half_band"]
    rg, ig = dm.static_solve(p)
    assert rg.aborted == ro.aborted == 0 and rg.converged_steps == ro.converged_steps
    assert np.abs(np.asarray(ig) - np.asarray(io)).max() <= 1
    no = om.nodes()
    ng = dm.nodes()[new_of_old]  # back to the generator's numbering
    x0 = m.freeze()["xyz"]
    scale = max(np.abs(no[:, :3] - x0).max(), 1e-12)
    assert np.abs(ng[:, :3] - no[:, :3]).max() <= 1e-8 * scale
    assert np.abs(ng[:, 3:] - no[:, 3:]).max() <= 1e-8 * max(np.abs(no[:, 3:]).max(), 1e-12)
