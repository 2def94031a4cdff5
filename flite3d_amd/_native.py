"""ctypes loaders for the in-tree native libraries.

The libraries are built by ``flite3d_amd.build`` (or ``__graft_entry__.build``)
next to this file.  A missing library raises: there is no Python fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))

_prepro = None
_device = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -m flite3d_amd.build` first")
    return ctypes.CDLL(path)


def prepro():
    global _prepro
    if _prepro is None:
        lib = _load("libflite3d_prepro.so")
        P = ctypes.POINTER
        i, d = ctypes.c_int, ctypes.c_double
        lib.fp_mesh_build.restype = ctypes.c_void_p
        lib.fp_mesh_build.argtypes = [i, P(d), i, P(i), i, P(i), P(i), i]
        lib.fp_mesh_free.argtypes = [ctypes.c_void_p]
        lib.fp_last_error.restype = ctypes.c_char_p
        lib.fp_mesh_nlevels.argtypes = [ctypes.c_void_p]
        lib.fp_level_sizes.argtypes = [ctypes.c_void_p, i, P(i), P(i), P(i),
                                       P(i)]
        lib.fp_level_copy.argtypes = [ctypes.c_void_p, i, P(i), P(d), P(i),
                                      P(d), P(d), P(d), P(d), P(i), P(i),
                                      P(i), P(i), P(i), P(i), P(d)]
        lib.fp_fine_old_index.argtypes = [ctypes.c_void_p, P(i)]
        lib.fp_metric_vjp.argtypes = [i, P(d), i, P(i), i, P(i), P(i), P(d), P(d), P(d)]
        lib.fp_mesh_vjp.argtypes = [i, P(d), i, P(i), i, P(i), P(i), P(d), P(d), P(d), P(d), P(d)]
        lib.fp_set_progress.argtypes = [ctypes.c_void_p]
        lib.fp_set_progress.restype = None
        _prepro = lib
    return _prepro


def device_path():
    return os.path.join(HERE, "libflite3d_amd.so")


def device():
    """The HIP library; its C-ABI is declared in include/flite3d_amd.h."""
    global _device
    if _device is None:
        _device = _load("libflite3d_amd.so")
    return _device
