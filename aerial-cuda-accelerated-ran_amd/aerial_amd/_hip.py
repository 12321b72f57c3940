"""Minimal ctypes access to the HIP runtime for launching kernel nodes that
the C ABI hands back (hipKernelNodeParams)."""
import ctypes as C

from . import _lib as L

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipLaunchKernel.argtypes = [C.c_void_p, L.Dim3, L.Dim3, C.c_void_p, C.c_size_t, C.c_void_p]
        h.hipLaunchKernel.restype = C.c_int
        _hip = h
    return _hip


def launch_node(node, stream_handle):
    """Launch a hipKernelNodeParams on a stream (like a graph kernel node)."""
    rc = hip().hipLaunchKernel(node.func, node.gridDim, node.blockDim, node.kernelParams, node.sharedMemBytes,
                               stream_handle)
    if rc != 0:
        raise RuntimeError("hipLaunchKernel failed: %d" % rc)
