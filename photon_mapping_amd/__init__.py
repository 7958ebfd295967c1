"""MI355X-native kNN photon gather (gfx950 HIP kernels behind a C ABI).

The compute path is libpm_gather.so; this package holds its Python mirror
of the reference's photon-map interface, the procedural benchmark scenes and
the torch autograd wrapper of the gather.
"""

from ._lib import PmError  # noqa: F401
from .photon_map import (BatchedPlan, Context, Photon, PhotonMap, autotune,  # noqa: F401
                         gather_batched, make_materials, make_queries)

__all__ = ["BatchedPlan", "Context", "Photon", "PhotonMap", "PmError", "autotune",
           "gather_batched", "make_materials", "make_queries"]
