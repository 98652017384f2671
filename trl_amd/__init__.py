"""trl_amd: MI355X-native HIP backend for trl's thick-restart block Lanczos.

Native pieces (built in-tree by trl_amd.build):
  libtrl_hip.so         gfx950 kernels behind the C-ABI of include/trl_hip.h
  libtrl_hip_solver.so  the C++20 trl::hip backend + driver, C-ABI include/trl_hip_solver.h

Python access is through ctypes (trl_amd._lib) and the thin wrappers in
trl_amd.solver; every call runs on the GPU, there is no CPU fallback.
"""

from . import problems  # noqa: F401

__all__ = ["problems"]
