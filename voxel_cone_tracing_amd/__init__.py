"""MI355X (gfx950) voxel cone tracing core.

The product is libvct.so (HIP kernels + the C ABI of include/vct.h).  This
package holds its build recipe, a ctypes binding and the procedural
workloads; nothing here computes on the CPU in place of the GPU.
"""
from . import _lib
from ._lib import load, device_count, VctError

__all__ = ["_lib", "load", "device_count", "VctError"]
