"""dygrass_amd -- MI355X-native dyGRASS random-walk sparsifier updaters.

The compute path is libdygrass.so (hand-written HIP for gfx950 behind a C
ABI, include/dygrass.h).  This package is its Python host interface; it has
no CPU implementation of any walk.
"""
from ._lib import DygError, WalkParams, WalkStats, RunStats, LIB_PATH, BIN_DIR  # noqa: F401
from .core import (  # noqa: F401
    Context,
    DeviceArray,
    DeviceGraph,
    DualGraph,
    Server,
    Stream,
    WalkResult,
    csr_from_edges,
    incremental_update,
    incremental_update_dev,
    params,
    read_mtx_auto,
    stream_params,
    walk_decremental,
    walk_heuristic,
    walk_incremental,
)

__all__ = [
    "Context", "DeviceArray", "DeviceGraph", "DualGraph", "Server", "Stream", "stream_params", "read_mtx_auto", "WalkResult", "csr_from_edges", "incremental_update", "incremental_update_dev",
    "params", "walk_decremental", "walk_heuristic", "walk_incremental", "DygError", "WalkParams",
    "WalkStats", "RunStats", "LIB_PATH", "BIN_DIR",
]
