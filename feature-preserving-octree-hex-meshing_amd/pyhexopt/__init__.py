"""Python side of the MI355X hex-optimisation path: ctypes bindings of libhexopt.so
(capi; meshio for the mesh and feature-graph files), synthetic workloads (meshgen) and the
torch autograd wrapper of the adjoint (pyhexopt.autograd, imported on demand because it
needs torch)."""
from . import capi, meshgen, meshio  # noqa: F401
