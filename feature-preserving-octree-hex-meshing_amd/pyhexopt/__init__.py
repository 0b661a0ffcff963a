"""Python side of the MI355X hex-optimisation path: ctypes binding of libhexopt.so
(capi), synthetic workloads (meshgen) and the torch autograd wrapper of the adjoint
(pyhexopt.autograd, imported on demand because it needs torch)."""
from . import capi, meshgen  # noqa: F401
