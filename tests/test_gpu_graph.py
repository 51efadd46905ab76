"""The FP64 Newton iteration replayed from a captured graph (factorisation,
solve, UpdateDisps, the next assembly and residual scan, the copy of the scalar
block; load time and gravity factor read from device memory) gives
bit-identical Newton histories and results to direct launches, across load
steps, bisections and a reset."""
import os

import numpy as np
import pytest

from giraffe_amd import model as M, runtime

pytestmark = pytest.mark.gpu


def _solve(m, p, graph):
    if graph:
        os.environ.pop("GF_NO_GRAPH", None)
    else:
        os.environ["GF_NO_GRAPH"] = "1"
    try:
        dm = runtime.DeviceModel(m)
    finally:
        os.environ.pop("GF_NO_GRAPH", None)
    rep, iters = dm.static_solve(p)
    nodes = dm.nodes()
    dm.reset()
    rep2, iters2 = dm.static_solve(p)  # the same graph replayed on a reset state
    assert list(iters2) == list(iters)
    np.testing.assert_array_equal(dm.nodes(), nodes)
    return rep, iters, nodes


@pytest.mark.parametrize("which", ["catenary", "plate"])
def test_graph_replay_matches_direct_launches(which):
    if which == "catenary":
        m, p, _ = M.catenary(n_elem=400)
    else:
        m, p, _ = M.shell_plate(nq=8)
    r1, i1, n1 = _solve(m, p, graph=False)
    r2, i2, n2 = _solve(m, p, graph=True)
    assert list(i1) == list(i2)
    assert r1.total_solves == r2.total_solves
    np.testing.assert_array_equal(n1, n2)
