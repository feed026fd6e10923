"""MI355X-native overlapInCore: canu's k-mer seed / banded prefix-edit-distance overlapper on gfx950.

Native pieces: liboic_hip.so (HIP kernels + C ABI, include/oic.h) and libreadsim.so (workload
simulator).  Build them with ``python -c "import __graft_entry__; __graft_entry__.build()"``.
"""
__all__ = ["oic", "readsim", "build"]
