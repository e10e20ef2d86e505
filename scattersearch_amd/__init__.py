"""MI355X-native graph-ANN query path (ScatterSearch's annsearch.h interface).

The compute path is libscattersearch.so: HIP kernels for gfx950 behind the C
ABI in include/scattersearch.h.  This package is the Python view of that ABI
plus harness pieces (synthetic workloads, graph construction, metrics).
"""
from ._lib import SSError, SS_NONE, SS_MODE_BATCHED, SS_MODE_PER_QUERY  # noqa: F401
from .index import (ANNSearch, Index, Metric, Neighbor, METRIC_L2,  # noqa: F401
                    METRIC_INNER_PRODUCT, STATS_DTYPE, dim_padded)

__all__ = ["ANNSearch", "Index", "Metric", "Neighbor", "SSError", "SS_NONE",
           "SS_MODE_BATCHED", "SS_MODE_PER_QUERY", "METRIC_L2", "METRIC_INNER_PRODUCT",
           "STATS_DTYPE", "dim_padded"]
