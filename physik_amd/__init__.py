"""MI355X-native implicit-Euler step of PhysiK's corotational soft bodies.

The product is libphysik_amd.so (built by physik_amd.build): HIP kernels for
gfx950 behind two C ABIs, the drop-in PHYSIK_* subset (include/PhysiKAPI.h)
and the device-level pk_* layer (include/physik_amd.h). This package holds the
ctypes bindings (native, system, world), the seeded mesh generator and the
benchmark workloads.
"""

from .native import LIB_PATH, lib  # noqa: F401

__all__ = ["LIB_PATH", "lib"]
