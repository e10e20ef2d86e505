"""ss_search_device_noalloc, the search entry point that allocates nothing
(no counterpart in the reference; for callers that cannot allocate): called
through the C ABI on device buffers, it must return exactly what
ss_search_device returns for the same arguments (ids, distance bits and
every per-query counter), on two calls in a row, the second with a smaller
batch, so that state the scratch array keeps from the first call cannot leak
into the second; and it refuses batches above its maximum."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(fn, ix, dq, nq, dK, dL, k_max, l_max, params, stream):
    import torch
    ids = torch.full((nq, k_max), -1, dtype=torch.int32, device="cuda")
    ds = torch.full((nq, k_max), float("nan"), dtype=torch.float32, device="cuda")
    st = torch.zeros((nq, 8), dtype=torch.int32, device="cuda")
    vp = ctypes.c_void_p
    rc = fn(ix.handle, vp(dq.data_ptr()), nq, vp(dK.data_ptr() if dK is not None else None),
            vp(dL.data_ptr() if dL is not None else None), k_max, l_max, k_max, ctypes.byref(params),
            vp(ids.data_ptr()), vp(ds.data_ptr()), vp(st.data_ptr()), vp(stream.cuda_stream))
    torch.cuda.synchronize()
    return rc, ids.cpu().numpy().view(np.uint32), ds.cpu().numpy(), st.cpu().numpy().view(np.uint32)


def test_noalloc_search_equals_search_device(c0):
    import torch
    from scattersearch_amd import Index, dim_padded
    from scattersearch_amd._lib import SearchParams, lib, SS_ORDER_MIN_BATCH
    base = c0["base"]
    ix = Index(base)
    ix.set_graph(c0["adj"], c0["ep"])
    rng = np.random.default_rng(23)
    nq1 = SS_ORDER_MIN_BATCH + 904  # the first call takes the locality order
    q = base[rng.choice(base.shape[0], nq1, replace=False)] + np.float32(0.05) * rng.standard_normal(
        (nq1, base.shape[1]), dtype=np.float32)
    dp = dim_padded(base.shape[1])
    dq = torch.zeros((nq1, dp), dtype=torch.float32, device="cuda")
    dq[:, :base.shape[1]] = torch.from_numpy(q).cuda()
    K = rng.choice([1, 10, 32], size=nq1).astype(np.int32)
    L = np.array([rng.integers(max(int(k), 16), 129) for k in K], dtype=np.int32)
    dK = torch.from_numpy(K).cuda()
    dL = torch.from_numpy(L).cuda()
    stream = torch.cuda.Stream()
    for call, (nq, per_query, wq) in enumerate(((nq1, True, 1), (300, False, 1), (200, False, 2))):
        print(f"no-alloc call {call}: {nq} queries, per-query K/L {per_query}, waves_per_query {wq}", flush=True)
        p = SearchParams()
        p.visited_slots = 4096
        p.waves_per_query = wq
        args = (ix, dq, nq, dK if per_query else None, dL if per_query else None, 32 if per_query else 10,
                128 if per_query else 48, p, stream)
        rc_ref, ids_ref, ds_ref, st_ref = _run(lib.ss_search_device, *args)
        rc, ids, ds, st = _run(lib.ss_search_device_noalloc, *args)
        assert rc_ref == 0 and rc == 0, (rc_ref, rc, lib.ss_last_error())
        assert (st[:, 3] == 0).all(), f"call {call}: a query did not complete"
        if wq == 1:
            assert np.array_equal(ids, ids_ref), f"call {call}: ids differ"
            assert np.array_equal(ds.view(np.uint32), ds_ref.view(np.uint32)), f"call {call}: distances differ"
            assert np.array_equal(st[:, :3], st_ref[:, :3]), f"call {call}: ndc, hops or lossy counts differ"
        else:  # several wavefronts per query depend on timing: entries checked, not equality
            assert (ids[:, :10] < base.shape[0]).all()
            assert all(len(set(r.tolist())) == 10 for r in ids[:, :10])


def test_noalloc_search_refuses_what_it_cannot_do(c0):
    import torch
    from scattersearch_amd import Index, dim_padded
    from scattersearch_amd._lib import SearchParams, lib, SS_EINVAL, SS_NOALLOC_MAX_QUERIES
    ix = Index(c0["base"])
    ix.set_graph(c0["adj"], c0["ep"])
    dq = torch.zeros((4, dim_padded(c0["base"].shape[1])), dtype=torch.float32, device="cuda")
    out = torch.zeros((4, 10), dtype=torch.int32, device="cuda")
    vp = ctypes.c_void_p
    p = SearchParams()
    rc = lib.ss_search_device_noalloc(ix.handle, vp(dq.data_ptr()), SS_NOALLOC_MAX_QUERIES + 1, None, None, 10, 48,
                                      10, ctypes.byref(p), vp(out.data_ptr()), None, None, None)
    assert rc == SS_EINVAL and b"SS_NOALLOC_MAX_QUERIES" in lib.ss_last_error()
    p.use_graph = 1
    rc = lib.ss_search_device_noalloc(ix.handle, vp(dq.data_ptr()), 4, None, None, 10, 48, 10, ctypes.byref(p),
                                      vp(out.data_ptr()), None, None, None)
    assert rc == SS_EINVAL and b"use_graph" in lib.ss_last_error()
