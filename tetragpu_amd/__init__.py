"""tetragpu_amd: TetraGPU's relation extraction and critical-point path on
AMD MI355X (gfx950), behind the reference's relation API.

Native code: tetragpu_amd/lib/libtetra_hip.so (HIP kernels + C ABI,
include/tetra_c.h) and libtetra_host.so (mesh generator, C++ drop-in layer).
"""
from .batch import Batch
from .mesh import (ALL_FLAGS, CLASS_NAMES, LABEL6_NAMES, REL, Mesh, device_count, flags, init,
                   partitioned_classes, split_classes, vv_footprint)
from .meshgen import TetMesh, kuhn, kuhn_counts

__all__ = ["Batch", "Mesh", "REL", "ALL_FLAGS", "CLASS_NAMES", "LABEL6_NAMES", "flags", "init",
           "device_count", "partitioned_classes", "split_classes", "vv_footprint", "TetMesh",
           "kuhn", "kuhn_counts"]
