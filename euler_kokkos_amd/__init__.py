"""euler_kokkos_amd -- MI355X-native MUSCL-Hancock Godunov step (hydro) behind the
euler_kokkos Solver/ConfigMap interface.

Native parts (built in-tree by ``python -m euler_kokkos_amd.build``):
  libek_hip.so      gfx950 HIP kernels, the C ABI of include/ek_hip.h and the C++ host layer
                    (include/ek_host.h)
  bin/euler_kokkos_amd   the .ini-driven executable

Python modules: ``native`` (ctypes binding of the C ABI), ``params`` (HydroParams mirror),
``problems`` (initial conditions in the device layout).
"""
from . import native, params, problems  # noqa: F401

__all__ = ["native", "params", "problems"]
