"""phoebus_amd: MI355X-native (gfx950) implementation of phoebus's per-RK-stage fluid path.

The product is libphx.so (phoebus_amd/lib), HIP kernels behind the C ABI of include/phx.h.
This package only binds it (ctypes) for tests and the benchmark.
"""
from . import _abi  # noqa: F401

__all__ = ["_abi"]
