"""mettle_amd — MI355X (gfx950) execution backend for the mettle DuckDB
extension.  The compute lives in libmettle_hip.so (HIP kernels behind the C
ABI of include/mettle_hip.h); this package holds the host-side bindings.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
