"""MI355X-native Eos render path: hand-written HIP kernels behind a C ABI.

``eos_amd._lib`` binds include/eos_c.h; ``eos_amd.scenes`` makes the seeded
workloads.  The HIP library is built in-tree by ``eos_amd/build.py``.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
