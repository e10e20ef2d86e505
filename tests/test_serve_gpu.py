"""Continuous-batching serving: every result equals the same query's result
alone (the batch search and the CPU oracle), in any mix of arrivals, K and L,
across ring wrap-around and across serving activations that end and restart.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(d):
    from scattersearch_amd import Index
    ix = Index(d["base"])
    ix.set_graph(d["adj"], d["ep"])
    return ix


def test_serving_results_equal_alone(small):
    import oracle
    from scattersearch_amd import serve
    ix = _index(small)
    rng = np.random.default_rng(5)
    q = np.concatenate([small["queries"]] * 4)  # repeats mix with other requests
    nq = q.shape[0]
    K = rng.choice([1, 10, 32], size=nq).astype(np.uint32)
    L = np.array([rng.integers(max(k, 16), 129) for k in K], dtype=np.uint32)
    arr = serve.poisson_arrivals(nq, 20000, seed=1)
    ids, ds, lat, st = serve.replay(ix, q, K, L, arr, k_max=32, l_max=128, visited_slots=8192)
    assert (lat > 0).all()
    for i in range(nq):
        k = int(K[i])
        o_ids, o_d, o_ndc, o_hops = oracle.search(small["base"], small["adj"], small["ep"], q[i], k, int(L[i]))
        assert np.array_equal(ids[i, :k], o_ids), f"request {i} K={k} L={L[i]}"
        assert np.array_equal(ds[i, :k].view(np.uint32), o_d.view(np.uint32))
        assert (ids[i, k:] == 0xFFFFFFFF).all()
        assert st["hops"][i] == o_hops and st["ndc"][i] == o_ndc


def test_serving_results_equal_alone_at_the_serving_mix(small):
    """The serving workload's request mix (K = 10 for 80% of the requests and
    100 for the rest, L uniform in max(64, K)..512) through activations sized
    for k_max = 100 and l_max = 512 (pools of up to eight 64-entry chunks):
    every result equals the oracle's search of that request alone, ids,
    distance bits and hops, and the distance count wherever the visited hash
    forgot nothing."""
    import oracle
    from scattersearch_amd import serve
    ix = _index(small)
    rng = np.random.default_rng(17)
    q = np.concatenate([small["queries"]] * 6)
    nq = q.shape[0]
    K = np.where(rng.random(nq) < 0.8, 10, 100).astype(np.uint32)
    L = rng.integers(np.maximum(64, K), 513).astype(np.uint32)
    arr = serve.poisson_arrivals(nq, 50000, seed=3)
    print(f"serving mix: {nq} requests, K in {{10, 100}}, L {L.min()}..{L.max()}", flush=True)
    ids, ds, lat, st, tr = serve.replay(ix, q, K, L, arr, k_max=100, l_max=512, with_trace=True)
    assert (lat > 0).all()
    assert (st["status"] == 0).all()
    # the timeline of every request is in order (device times are mapped to
    # the host clock through one reading, good to some tens of microseconds)
    submit, claim, finish, seen = tr["submit_us"], tr["claim_us"], tr["finish_us"], tr["seen_us"]
    assert (submit >= arr).all() and (seen > submit).all()
    np.testing.assert_allclose(seen - arr, lat, atol=1e-6)
    assert (finish >= claim).all()
    assert (claim > submit - 100.0).all() and (finish < seen + 100.0).all()
    assert (tr["waves_alive"] >= tr["waves_standby"]).all() and tr["waves_alive"].max() > 0
    o_ids, o_d, o_ndc, o_hops = oracle.search_batch(small["base"], small["adj"], small["ep"], q, K, L, k_stride=100)
    for i in range(nq):
        k = int(K[i])
        assert np.array_equal(ids[i, :k], o_ids[i, :k]), f"request {i} K={k} L={L[i]}"
        assert np.array_equal(ds[i, :k].view(np.uint32), o_d[i, :k].view(np.uint32)), f"request {i}"
        assert (ids[i, k:] == 0xFFFFFFFF).all()
        assert st["hops"][i] == o_hops[i], f"request {i}"
        if st["lossy"][i] == 0:
            assert st["ndc"][i] == o_ndc[i], f"request {i}"


def test_submit_and_poll(small):
    from scattersearch_amd import serve
    ix = _index(small)
    alone, _, _ = ix.search(small["queries"][:8], 10, 48, visited_slots=8192)
    with serve.Server(ix, 10, 48, visited_slots=8192, streams=2) as srv:
        got = {}
        for rnd in range(2):
            tickets = [srv.submit(small["queries"][i], 10, 48) for i in range(8)]
            t0 = time.time()
            while len(got) < 8 * (rnd + 1) and time.time() - t0 < 20:
                tk, ids, ds, st = srv.poll()
                for t, r in zip(tk, ids):
                    got[int(t)] = r.copy()
            assert len(got) == 8 * (rnd + 1)
            for j, t in enumerate(tickets):
                assert np.array_equal(got[t], alone[j])
            time.sleep(0.05)


def test_admission_limits(small):
    from scattersearch_amd import SSError, serve
    from scattersearch_amd._lib import SS_EBUSY, SS_EINVAL
    ix = _index(small)
    with serve.Server(ix, 10, 48, ring_slots=64) as srv:
        with pytest.raises(SSError) as e:
            srv.submit(small["queries"][0], 20, 48)  # K above k_max
        assert e.value.code == SS_EINVAL
        with pytest.raises(SSError):
            srv.submit(small["queries"][0], 10, 64)  # L above l_max
        for i in range(64):
            srv.submit(small["queries"][i % 64], 10, 48)
        with pytest.raises(SSError) as e:
            srv.submit(small["queries"][0], 10, 48)  # 64 unpolled requests fill the ring
        assert e.value.code == SS_EBUSY
        n, t0 = 0, time.time()
        while n < 64 and time.time() - t0 < 20:
            n += len(srv.poll()[0])
        assert n == 64
        srv.submit(small["queries"][0], 10, 48)  # room again


def test_stop_with_requests_in_flight(small):
    from scattersearch_amd import serve
    ix = _index(small)
    srv = serve.Server(ix, 10, 48)
    for i in range(32):
        srv.submit(small["queries"][i], 10, 48)
    srv.stop()


def test_ring_wraps_and_activations_restart(small):
    """300 requests through a 64-slot ring, in bursts separated by pauses
    longer than an idle wavefront's linger, so slots are reused and new
    activations start; each result equals the query searched alone."""
    from scattersearch_amd import serve
    ix = _index(small)
    q = small["queries"]
    alone, alone_d, _ = ix.search(q, 10, 48, visited_slots=4096)
    got = {}
    with serve.Server(ix, 10, 48, ring_slots=64, visited_slots=4096, linger_us=500) as srv:
        for burst in range(6):
            for i in range(50):
                t = 50 * burst + i
                assert srv.submit(q[t % q.shape[0]], 10, 48) == t
            t0 = time.time()
            while len(got) < 50 * (burst + 1) and time.time() - t0 < 20:
                tk, ids, ds, st = srv.poll()
                for t, r, d in zip(tk, ids, ds):
                    got[int(t)] = (r.copy(), d.copy())
            assert len(got) == 50 * (burst + 1), f"burst {burst}"
            time.sleep(0.005)  # every wavefront has left; the next burst starts a new activation
    assert len(got) == 300
    for t, (r, d) in got.items():
        j = t % q.shape[0]
        assert np.array_equal(r, alone[j]), f"ticket {t}"
        assert np.array_equal(d.view(np.uint32), alone_d[j].view(np.uint32))
