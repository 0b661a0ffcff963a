"""Python side of the MI355X hex-optimisation path: ctypes binding of libhexopt.so
(capi), synthetic workloads (meshgen) and the torch autograd wrapper of the adjoint
(autograd, imported on demand)."""
from . import capi, meshgen  # noqa: F401
