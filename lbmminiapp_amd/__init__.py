"""lbmminiapp_amd -- MI355X-native (gfx950 HIP) lattice Boltzmann collide-stream engine.

The compute path is the C ABI in include/lbm_capi.h (liblbm_hip.so); this
package exposes it to Python through ctypes (``capi``) and holds the C++ host
layer that mirrors the reference's Parameters / Lattice / Geometry / Kernel /
Propagation classes (``host/``).
"""
from .capi import (DenseLattice, DeviceBuffer, LbmError, SparseKernels, copy_bandwidth,
                   device_count, load)

__all__ = ["DenseLattice", "DeviceBuffer", "LbmError", "SparseKernels", "copy_bandwidth",
           "device_count", "load"]
