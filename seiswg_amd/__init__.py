"""seiswg_amd — MI355X-native staggered-grid time loop of seiswg.

The product is libseiswg_hip.so (hand-written gfx950 kernels, the C ABI of
include/seiswg_hip.h and the C++ host of include/seiswg_host.h) plus the
`seiswg` CLI in seiswg_amd/bin.  This Python module is a ctypes binding of
that library for tests and benchmarks.
"""
from ._lib import (Context, Multi, SeiswgError, load, ricker_table, infer_grid, grid_index, run_config, bench_copy,  # noqa: F401
                   release_cached_memory, Vfs, run_simulation,
                   LIB_PATH, CLI_PATH, EXPORTS)
from . import _lib as abi  # noqa: F401
