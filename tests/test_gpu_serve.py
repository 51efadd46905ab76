"""Request server (gf_server_*): independent static solves resident together on
one GPU.  Every request's result must be the same alone (gf_static_solve on its
own model) and in the mix (admitted while others are resident, interleaved at
Newton-step granularity on pooled streams), and admission must respect the
slot and memory limits."""
import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime

pytestmark = pytest.mark.gpu


def _requests(n, seed, lo=30, hi=400):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ne = int(rng.integers(lo, hi))
        m, p, _ = M.riser_request(rng, n_elem=ne)
        out.append((m, p))
    # one roll-up (follower moment) in the mix
    m, p, _ = M.cantilever_rollup(n_elem=40)
    out.append((m, p))
    return out


def _alone(m, p):
    dm = runtime.DeviceModel(m)
    rep, iters = dm.static_solve(p)
    return rep, iters, dm.nodes()


def _drain(srv, ids, watch=None, limit_s=100.0):
    import time
    results = {}
    t0 = time.monotonic()
    while len(results) < len(ids):
        assert time.monotonic() - t0 < limit_s, f"requests not finished: {sorted(set(ids) - set(results))}"
        srv.pump(2000)
        c = srv.counts()
        if watch:
            watch(c)
        for rid in ids:
            if rid not in results:
                try:
                    results[rid] = srv.take(rid)
                except runtime.GiraffeError:
                    pass
    return results


def test_results_same_alone_and_in_the_mix():
    reqs = _requests(7, 11)
    alone = [_alone(m, p) for m, p in reqs]
    srv = runtime.Server(max_inflight=3)
    ids = [srv.submit(m, p) for m, p in reqs]
    res = _drain(srv, ids)
    for (rep_a, it_a, nodes_a), rid in zip(alone, ids):
        rep, it, nodes, times = res[rid]
        assert rep.aborted == rep_a.aborted == 0
        assert list(it) == list(it_a)
        # same kernels, same order of operations on another stream: identical
        assert np.array_equal(nodes, nodes_a)
        assert times[1] >= times[2] >= 0.0
    srv.close()


def test_admission_by_memory_and_slots():
    reqs = _requests(5, 12, lo=200, hi=300)
    sizes = [runtime.footprint(m) for m, _ in reqs]
    cap = int(max(sizes) * 1.5)  # at most one of these fits beside another only if both are small
    srv = runtime.Server(max_inflight=4, mem_cap=cap)
    ids = [srv.submit(m, p) for m, p in reqs]
    seen = []

    def watch(c):
        assert c["bytes"] <= cap
        assert c["resident"] <= 4
        seen.append(c["resident"])

    res = _drain(srv, ids, watch)
    assert all(res[i][0].aborted == 0 for i in ids)
    assert max(seen) <= 2
    srv.close()


def test_request_larger_than_cap_is_refused():
    m, p, _ = M.catenary(n_elem=200)
    srv = runtime.Server(max_inflight=2, mem_cap=runtime.footprint(m) // 2)
    with pytest.raises(runtime.GiraffeError):
        srv.submit(m, p)
    srv.close()


def test_submit_copies_every_array():
    """gf_server_submit copies the whole description: Pipe_1 kinds, buoyancy and
    deck element numbers included.  The requests wait in the queue behind a
    blocker while the caller's arrays are overwritten; their results must still
    equal the standalone solves."""
    from test_gpu_pipe import _hanging
    pipe, pp = _hanging()
    pipe.beam_number = list(range(pipe.n_beams, 0, -1))
    plate, ps, _ = M.shell_plate(nq=4, n_steps=3, p_max=2e5)
    plate.shell_number = list(range(plate.n_shells, 0, -1))
    alone = [_alone(pipe, pp), _alone(plate, ps)]
    blocker, pb, _ = M.catenary(n_elem=300)
    srv = runtime.Server(max_inflight=1)
    ids = [srv.submit(blocker, pb), srv.submit(pipe, pp), srv.submit(plate, ps)]
    f = pipe.freeze()
    f["beam_kind"][:] = 7
    f["beam_buoy"][:] = 1e6
    f["beam_number"][:] = 0
    plate.freeze()["shell_number"][:] = 0
    res = _drain(srv, ids)
    for (rep_a, it_a, nodes_a), rid in zip(alone, ids[1:]):
        rep, it, nodes, _ = res[rid]
        assert rep.aborted == rep_a.aborted == 0
        assert list(it) == list(it_a)
        assert np.array_equal(nodes, nodes_a)
    srv.close()
