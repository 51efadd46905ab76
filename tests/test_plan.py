"""Host-side planning (gf_model_plan, no device call): DOF count, half-bandwidth,
storage layout and the reverse Cuthill-McKee node order."""
import os

import numpy as np
import pytest

from giraffe_amd import model as M, runtime
from renumber import corners_first, renumber


def test_generator_orders_kept():
    """The generators' own numberings are already narrow: RCM does not beat them."""
    m, _, _ = M.shell_plate(nq=16)
    assert runtime.plan(m) == {"n_free": 5091, "half_band": 338, "block": 0, "rcm": False}
    m, _, _ = M.catenary(n_elem=300)
    assert runtime.plan(m)["block"] == 16


@pytest.mark.parametrize("nq", [8, 16, 64])
def test_rcm_recovers_a_narrow_band(nq):
    """Mid-side nodes numbered after all corners (as inputs/tutorial05 does) make
    the natural band nearly dense; RCM brings it back within 5 % of the
    generator's row-by-row order.  A random numbering gives the same."""
    m, _, _ = M.shell_plate(nq=nq)
    ref = runtime.plan(m)["half_band"]
    r = renumber(m, corners_first(nq))
    os.environ["GF_NO_RCM"] = "1"
    try:
        natural = runtime.plan(r)["half_band"]
    finally:
        del os.environ["GF_NO_RCM"]
    got = runtime.plan(r)
    assert natural > 3 * ref
    assert got["rcm"] and got["half_band"] <= 1.05 * ref
    rng = np.random.default_rng(nq)
    assert runtime.plan(renumber(m, rng.permutation(m.n_nodes)))["half_band"] <= 1.05 * ref
