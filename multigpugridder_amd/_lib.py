"""ctypes binding of ``libMultiGPUGridder.so`` (the C ABI in ``include/mgg.h``).

The library is built in-tree by ``make -C multigpugridder_amd/csrc`` (or
``__graft_entry__.build()``).  There is no fallback: if the library is missing,
every call raises :class:`LibraryNotBuilt`.

torch (when importable) is imported before the library is loaded.  torch ships
its own ``libamdhip64.so.7`` / ``librocfft.so.0``; loading torch first lets the
gridder bind to those same copies through their SONAMEs, so one process never
holds two HIP runtimes.
"""
from __future__ import annotations

import ctypes
import os
import threading

try:  # keep one HIP runtime per process (see module docstring)
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for the C ABI
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libMultiGPUGridder.so")

MGG_OK = 0
MGG_ERR_INVALID_ARGUMENT = 1
MGG_ERR_NOT_SET = 2
MGG_ERR_DEVICE = 3
MGG_ERR_OUT_OF_MEMORY = 4
MGG_ERR_HIP = 5
MGG_ERR_FFT = 6
MGG_ERR_UNSUPPORTED = 7
MGG_ERR_TIMEOUT = 8
MGG_ERR_BUSY = 9

DTYPE_F32, DTYPE_BF16, DTYPE_F16 = 0, 1, 2

_ERR_NAMES = {
    1: "invalid argument", 2: "not set", 3: "device", 4: "out of memory",
    5: "HIP", 6: "FFT", 7: "unsupported", 8: "timeout", 9: "busy",
}


class LibraryNotBuilt(RuntimeError):
    pass


class MggError(RuntimeError):
    """A non-zero status from the C ABI; ``code`` is the MGG_ERR_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{_ERR_NAMES.get(code, code)}] {message}")
        self.code = code


_lock = threading.Lock()
_lib = None

_vp, _ci, _cf, _cd, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "mgg_last_error": [],
    "mgg_version": [],
    "mgg_device_count": [_ip],
    "mgg_create": [_ci, _ci, _cf, _ci, _ci, _ip, _ci, _ci, ctypes.POINTER(_vp)],
    "mgg_destroy": [_vp],
    "mgg_set_volume": [_vp, _vp, _ip],
    "mgg_set_coord_axes": [_vp, _vp, _ip],
    "mgg_set_images": [_vp, _vp, _ip],
    "mgg_set_cas_images": [_vp, _vp, _ip],
    "mgg_set_cas_volume": [_vp, _vp, _ip],
    "mgg_set_kb_table": [_vp, _vp, _ip],
    "mgg_set_mask_radius": [_vp, _cf],
    "mgg_set_num_axes": [_vp, _ci],
    "mgg_set_max_axes_to_allocate": [_vp, _ci],
    "mgg_set_num_streams_fp": [_vp, _ci],
    "mgg_set_num_streams_bp": [_vp, _ci],
    "mgg_set_gpus": [_vp, _ip, _ci],
    "mgg_set_ker_hwidth": [_vp, _cf],
    "mgg_get_sizes": [_vp, _ip, _ip, _ip, _ip],
    "mgg_get_mask_radius": [_vp],
    "mgg_forward_project": [_vp],
    "mgg_back_project": [_vp],
    "mgg_cas_volume_to_host": [_vp],
    "mgg_volume_to_cas_volume": [_vp],
    "mgg_get_volume": [_vp],
    "mgg_reconstruct_volume": [_vp],
    "mgg_reset_volume": [_vp],
    "mgg_op_forward_project": [_vp, _ci, _ci, _vp, _ci, _vp, _ci, _cf, _vp, _ci, _cf, _vp],
    "mgg_op_forward_project_checked": [_vp, _ci, _ci, _vp, _ci, _vp, _ci, _cf, _vp, _ci, _cf, _vp, _vp],
    "mgg_op_check_volume": [_vp, _ci, _ci, _vp, _ci, _vp, _vp],
    "mgg_op_back_project": [_vp, _vp, _ci, _vp, _ci, _vp, _ci, _cf, _vp, _ci, _cf, _vp],
    "mgg_op_back_project_adjoint": [_vp, _ci, _vp, _ci, _vp, _ci, _cf, _vp, _ci, _cf, _vp],
    "mgg_op_back_project_adjoint_checked": [_vp, _ci, _vp, _ci, _vp, _ci, _cf, _vp, _ci, _cf, _vp, _vp],
    "mgg_op_check_images": [_vp, _ci, _ci, _vp, _ci, _vp, _vp],
    "mgg_op_cas_images_to_images": [_vp, _vp, _ci, _ci, _ci, _vp],
    "mgg_op_cas_images_to_images_noalloc": [_vp, _vp, _ci, _ci, _ci, _vp],
    "mgg_op_images_to_cas_images": [_vp, _vp, _ci, _ci, _ci, _vp],
    "mgg_op_volume_to_cas_volume": [_vp, _ci, _vp, _ci, _ci, _cf, _vp],
    "mgg_op_cas_volume_to_volume": [_vp, _ci, _ci, _vp, _ci, _cf, _ci, _vp],
    "mgg_op_cas_images_to_images_transpose": [_vp, _vp, _ci, _ci, _ci, _vp],
    "mgg_op_volume_to_cas_volume_transpose": [_vp, _ci, _ci, _vp, _ci, _cf, _vp],
    "mgg_op_divide_density": [_vp, _vp, _vp, _sz, _cd, _vp],
    "mgg_op_reduce_partials": [_vp, _vp, _ci, _sz, _sz, _ci, _vp],
    "mgg_op_convert_volume": [_vp, _ci, _vp, _sz, _vp],
    "mgg_filter_cas_to_complex": [_vp, _vp, _ci, _ci, _vp],
    "mgg_filter_complex_to_cas": [_vp, _vp, _ci, _ci, _vp],
    "mgg_filter_complex_to_real": [_vp, _vp, _ci, _ci, _vp],
    "mgg_filter_real_to_complex": [_vp, _vp, _ci, _ci, _vp],
    "mgg_filter_fftshift2d": [_vp, _ci, _ci, _ci, _vp],
    "mgg_filter_fftshift3d": [_vp, _ci, _ci, _vp],
    "mgg_filter_pad": [_vp, _vp, _ci, _ci, _ci, _vp],
    "mgg_filter_crop": [_vp, _vp, _ci, _ci, _ci, _ci, _vp],
    "mgg_filter_divide_scalar": [_vp, _cf, _sz, _vp],
    "mgg_filter_divide_volume": [_vp, _vp, _sz, _vp],
    "mgg_filter_multiply_volume": [_vp, _vp, _sz, _vp],
    "mgg_filter_multiply_complex_volume": [_vp, _vp, _sz, _vp],
    "mgg_filter_add_volume": [_vp, _vp, _sz, _vp],
    "mgg_kb_table": [_vp, _ci, _ci],
    "mgg_precomp": [_vp, _ci, _ci],
    "mgg_uniform_axes": [_vp, _ci, _ci],
    # MEX dispatch layer (include/mgg_mex.h)
    "mgg_mex_call": [ctypes.c_char_p, _ci, _ci, _vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, _sz],
    # request server
    "mgg_server_create": [_vp, ctypes.POINTER(_vp)],
    "mgg_server_destroy": [_vp],
    "mgg_server_add_volume": [_vp, _vp, _ci, _cf, _ci, _ip],
    "mgg_server_remove_volume": [_vp, _ci],
    "mgg_server_submit": [_vp, _vp, ctypes.POINTER(ctypes.c_ulonglong)],
    "mgg_server_wait": [_vp, ctypes.c_ulonglong, _cd, _vp],
    "mgg_server_get_stats": [_vp, _vp],
    "mgg_server_hold": [_vp, _ci],
    "mgg_server_clock": [_vp],
    "mgg_host_alloc": [_sz, ctypes.POINTER(_vp)],
    "mgg_host_free": [_vp],
    # legacy ctypes symbols (Python/MultiGPUGridder_Python.py:11-48)
    "Gridder_new": [],
    "SetNumberGPUs": [_vp, _ci],
    "SetNumberStreams": [_vp, _ci],
    "SetVolume": [_vp, _fp, _ip],
    "GetVolume": [_vp],
    "ResetVolume": [_vp],
    "SetImages": [_vp, _fp],
    "GetImages": [_vp],
    "SetAxes": [_vp, _fp, _ip],
    "SetImgSize": [_vp, _ip],
    "SetMaskRadius": [_vp, _fp],
    "Forward_Project": [_vp],
    "Back_Project": [_vp],
}
_RESTYPES = {
    "mgg_last_error": ctypes.c_char_p,
    "mgg_version": ctypes.c_char_p,
    "mgg_get_mask_radius": _cf,
    "mgg_server_clock": _cd,
    "Gridder_new": _vp, "SetNumberGPUs": _vp, "SetNumberStreams": _vp, "SetVolume": _vp,
    "GetVolume": _fp, "ResetVolume": _vp, "SetImages": _vp, "GetImages": _fp, "SetAxes": _vp,
    "SetImgSize": _vp, "SetMaskRadius": _vp, "Forward_Project": _vp, "Back_Project": _vp,
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load (once) and return the ctypes handle of libMultiGPUGridder.so."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryNotBuilt(
                    f"{LIB_PATH} is missing: build it with `make -C multigpugridder_amd/csrc` "
                    "(or __graft_entry__.build()); there is no CPU fallback")
            h = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = _RESTYPES.get(name, _ci)
            _lib = h
    return _lib


def last_error() -> str:
    return lib().mgg_last_error().decode(errors="replace")


def check(status: int) -> None:
    if status != MGG_OK:
        raise MggError(status, last_error())


def call(name: str, *args) -> None:
    """Call an int-returning entry point and raise MggError on failure."""
    check(getattr(lib(), name)(*args))


def device_count() -> int:
    n = ctypes.c_int(0)
    call("mgg_device_count", ctypes.byref(n))
    return n.value


def int_array(values):
    arr = (ctypes.c_int * len(values))(*[int(v) for v in values])
    return arr
