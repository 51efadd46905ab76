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


def _drain(srv, ids, watch=None):
    results = {}
    while len(results) < len(ids):
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
