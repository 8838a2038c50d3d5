"""MI355X (gfx950) implementation of FFEA's massless fluctuating-FEM time step.

The compute path is libffea_hip.so (ffea_amd/csrc, C-ABI in include/ffea_hip.h); this
package holds its ctypes binding, the host-side mirror of the reference's Blob/Solver
interface, the FFEA unit system, mesh generation and file formats.
"""
from . import units, mesh  # noqa: F401

__all__ = ["units", "mesh", "blob"]
