"""The C-ABI library loads and exports every function include/scattersearch.h
declares; entry points that need no GPU work without one."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "scattersearch.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ss_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("ss_create", "ss_destroy", "ss_load_graph", "ss_set_graph", "ss_search",
                     "ss_search_device", "ss_bruteforce", "ss_last_error", "ss_fbin_read",
                     "ss_gt_read", "ss_nsg_read", "ss_build_graph"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from scattersearch_amd import _lib
    missing = [n for n in declared_functions() if not hasattr(_lib.lib, n)]
    assert not missing, f"not exported: {missing}"


def test_dim_padding():
    from scattersearch_amd import dim_padded
    assert [dim_padded(d) for d in (1, 8, 96, 100, 128, 200)] == [8, 8, 96, 104, 128, 200]


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_create_without_gpu_reports_error():
    from scattersearch_amd import Index, SSError
    from scattersearch_amd._lib import SS_EDEVICE
    with pytest.raises(SSError) as e:
        Index(np.zeros((4, 8), dtype=np.float32))
    assert e.value.code == SS_EDEVICE
    assert "device" in str(e.value)


def test_invalid_create_arguments():
    from scattersearch_amd._lib import lib, SS_EINVAL
    h = ctypes.c_void_p()
    base = np.zeros((4, 8), dtype=np.float32)
    p = base.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.ss_create(8, 4, p, 7, ctypes.byref(h)) == SS_EINVAL  # unknown metric
    assert lib.ss_create(0, 4, p, 0, ctypes.byref(h)) == SS_EINVAL
    assert lib.ss_create(8, 0, p, 0, ctypes.byref(h)) == SS_EINVAL
    assert lib.ss_create(8, 4, None, 0, ctypes.byref(h)) == SS_EINVAL
    assert b"ss_create" in lib.ss_last_error()


def test_null_index_calls_fail_cleanly():
    from scattersearch_amd._lib import lib, SS_EINVAL
    assert lib.ss_destroy(None) == 0
    assert lib.ss_set_graph(None, None, 4, 0) == SS_EINVAL
    assert lib.ss_search(None, None, 1, None, None, 1, 1, 1, None, None, None, None) == SS_EINVAL


def test_status_codes_match_the_header():
    """The Python constants are the header's: error codes, per-query status
    words, and the stats layout the kernels write (status is its 4th word)."""
    import ctypes
    import re
    from scattersearch_amd import _lib
    from scattersearch_amd.index import STATS_DTYPE
    hdr = open(os.path.join(ROOT, "include", "scattersearch.h")).read()
    defs = {m.group(1): int(m.group(2), 0) for m in re.finditer(r"#define (SS_\w+) \(?(-?[0-9xA-Fa-f]+)u?\)?", hdr)}
    for name in ("SS_OK", "SS_EINVAL", "SS_ENOMEM", "SS_EIO", "SS_EDEVICE", "SS_ESTATE", "SS_EBUSY",
                 "SS_EINTERNAL", "SS_QUERY_OK", "SS_QUERY_REJECTED", "SS_QUERY_FAULT", "SS_ORDER_AUTO",
                 "SS_ORDER_INDEX", "SS_ORDER_LOCALITY", "SS_ORDER_MIN_BATCH"):
        assert name in defs, name
        assert getattr(_lib, name) == defs[name], name
    assert ctypes.sizeof(_lib.QueryStats) == 32 == STATS_DTYPE.itemsize
    assert _lib.QueryStats.status.offset == 12 == STATS_DTYPE.fields["status"][1]
    assert ctypes.sizeof(_lib.SearchParams) == 48 and _lib.SearchParams.query_order.offset == 28
    assert _lib.SearchParams.team_size.offset == 32 and _lib.SearchParams.pool_layout.offset == 36
    assert _lib.SearchParams.pool_classes.offset == 40 and _lib.SearchParams.visited_bucket.offset == 44
    for f in ("query_order", "team_size", "pool_layout", "pool_classes", "visited_bucket"):
        assert f"uint32_t {f};" in hdr, f
    # ss_serve_trace: four doubles, then the two wavefront counts
    from scattersearch_amd.serve import TRACE_DTYPE
    assert TRACE_DTYPE.itemsize == 40 and TRACE_DTYPE.fields["waves_alive"][1] == 32
    for f in ("double submit_us;", "double claim_us;", "double finish_us;", "double seen_us;",
              "uint32_t waves_alive;", "uint32_t waves_standby;"):
        assert f in hdr, f
    for name in ("SS_POOL_AUTO", "SS_POOL_REGISTERS", "SS_POOL_LDS", "SS_POOL_SHIFT", "SS_NOALLOC_MAX_QUERIES",
                 "SS_CLASSES_AUTO",
                 "SS_CLASSES_OFF", "SS_CLASSES_SERIAL", "SS_VISITED_TWO_CHOICE"):
        assert getattr(_lib, name) == defs[name], name


def test_non_finite_base_is_refused():
    """NaN and infinite coordinates have no distance order; ss_create refuses
    them before it touches the device."""
    from scattersearch_amd._lib import lib, SS_EINVAL
    h = ctypes.c_void_p()
    for bad in (np.nan, np.inf, -np.inf):
        base = np.zeros((4, 8), dtype=np.float32)
        base[2, 5] = bad
        p = base.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        assert lib.ss_create(8, 4, p, 0, ctypes.byref(h)) == SS_EINVAL
        assert b"base vector 2" in lib.ss_last_error()
