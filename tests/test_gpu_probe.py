"""Per-kernel event timing (gf_kernel_probe): the launch and work-unit counts
that bench.py's roofline line is computed from."""
import numpy as np
import pytest

from giraffe_amd import model as M, runtime

pytestmark = pytest.mark.gpu


def test_probe_counts_match_the_work():
    m, p, _ = M.catenary(n_elem=500)
    dm = runtime.DeviceModel(m)
    info = dm.info()
    assert info["block"] > 0  # block-tridiagonal path
    dm.timing(True)
    dm.kernel_probe(dm.KERNEL_CLASSES)
    rep, iters = dm.static_solve(p)
    kp = dm.kernel_probe_read(reset=True)
    ph = dm.timing_read(reset=True)
    # beam assembly: every element once per Newton iteration
    ms, launches, units = kp["beam_assemble"]
    assert units == m.n_beams * rep.total_iterations
    assert launches == ph["assemble"][1]
    assert ms > 0
    # cyclic reduction: the level launches eliminate nblk/2 + nblk/4 + ... block rows per solve
    nblk = info["n_blocks"]
    n, per_solve, lv = nblk, 0, 0
    while n > 16:
        per_solve += n // 2
        n //= 2
        lv += 1
    ms, launches, units = kp["btd_level"]
    assert launches == lv * rep.total_solves
    assert units == per_solve * rep.total_solves
    assert kp["btd_tail"][1] == rep.total_solves
    # switched off: nothing more is recorded
    dm.kernel_probe([])
    dm.reset()
    dm.static_solve(p)
    assert dm.kernel_probe_read(reset=True) == {}
