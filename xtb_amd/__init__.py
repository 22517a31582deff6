"""MI355X-native GFN2-xTB single point with analytic gradient.

The compute path is the HIP library ``libxtb_amd.so`` built in-tree by
``xtb_amd.build``; there is no CPU fallback.
"""
from .engine import Batch, Server, SolverFailure, singlepoint, singlepoint_noalloc, eigh, library, device_count  # noqa: F401
