"""Request server over the C ABI ``mgg_server_*``.

The server keeps CAS volumes resident in HBM and runs forward / back projection
requests of mixed size with continuous batching: forward requests on one volume
share launches, several batches run at once on their own streams, and device
memory is admitted against a budget.  A forward request's result is
``forwardProject``'s images of its directions; a back request's result is
``getVol`` of the back projection of its images into a zeroed volume
(``MultiGPUGridder_Matlab_Class.m:198``, ``:245``, ``:314``).

Arrays are numpy arrays or torch tensors (float32, C-contiguous).  Host arrays
from :func:`pinned_empty` transfer asynchronously; device tensors are used in
place.  Every array of a request is kept alive until :meth:`Server.wait`
returns for it.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib

FORWARD, BACK = 0, 1


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("budget_bytes", ctypes.c_size_t),
                ("max_batch_directions", ctypes.c_int), ("num_pipelines", ctypes.c_int),
                ("max_queued", ctypes.c_int)]


class _Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("volume_id", ctypes.c_int), ("n_axes", ctypes.c_int),
                ("axes", ctypes.c_void_p), ("images", ctypes.c_void_p), ("volume", ctypes.c_void_p)]


class RequestInfo(ctypes.Structure):
    """Status and server-clock times (seconds) of a finished request."""
    _fields_ = [("status", ctypes.c_int), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]


class _Stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_ulonglong), ("completed", ctypes.c_ulonglong),
                ("failed", ctypes.c_ulonglong), ("batches", ctypes.c_ulonglong),
                ("batched_directions", ctypes.c_ulonglong), ("merged_batches", ctypes.c_ulonglong),
                ("budget_bytes", ctypes.c_size_t), ("resident_bytes", ctypes.c_size_t),
                ("in_use_bytes", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t),
                ("queued", ctypes.c_int), ("in_flight", ctypes.c_int), ("volumes", ctypes.c_int)]


def pinned_empty(shape, dtype=np.float32) -> np.ndarray:
    """Uninitialised numpy array in pinned host memory (freed with the array)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    ptr = ctypes.c_void_p()
    _lib.call("mgg_host_alloc", max(n, 1), ctypes.byref(ptr))
    buf = (ctypes.c_char * max(n, 1)).from_address(ptr.value)
    weakref.finalize(buf, _lib.lib().mgg_host_free, ptr)
    arr = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
    return arr


def _ptr(a, what, count):
    """Address of a float32 C-contiguous numpy array / torch tensor with count values."""
    if hasattr(a, "data_ptr"):  # torch tensor
        import torch
        if a.dtype != torch.float32 or not a.is_contiguous():
            raise ValueError(f"{what} must be a contiguous float32 tensor")
        if a.numel() < count:
            raise ValueError(f"{what} holds {a.numel()} values, needs {count}")
        return a.data_ptr()
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous:
        raise ValueError(f"{what} must be a C-contiguous float32 numpy array")
    if a.size < count:
        raise ValueError(f"{what} holds {a.size} values, needs {count}")
    return a.ctypes.data


class Server:
    """Resident volumes plus a continuously batched request queue on one GPU."""

    def __init__(self, device: int = 0, budget_bytes: int = 0, max_batch_directions: int = 0,
                 num_pipelines: int = 0, max_queued: int = 0):
        cfg = _Config(device, budget_bytes, max_batch_directions, num_pipelines, max_queued)
        h = ctypes.c_void_p()
        _lib.call("mgg_server_create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h
        self._sizes = {}      # volume id -> N
        self._keep = {}       # ticket -> arrays kept alive

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if self._h:
            _lib.call("mgg_server_destroy", self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- volumes --------------------------------------------------------------
    def add_volume(self, volume, interp_factor: float = 2.0, extra_padding: int = 3) -> int:
        """Makes ``volume`` (N^3 float32, x fastest) resident; returns its id."""
        vol = np.ascontiguousarray(volume, dtype=np.float32)
        n = vol.shape[0]
        if vol.shape != (n, n, n):
            raise ValueError("volume must be N x N x N")
        vid = ctypes.c_int()
        _lib.call("mgg_server_add_volume", self._h, vol.ctypes.data, n, float(interp_factor),
                  int(extra_padding), ctypes.byref(vid))
        self._sizes[vid.value] = n
        return vid.value

    def remove_volume(self, volume_id: int) -> None:
        _lib.call("mgg_server_remove_volume", self._h, int(volume_id))
        self._sizes.pop(volume_id, None)

    def volume_size(self, volume_id: int) -> int:
        return self._sizes[volume_id]

    # -- requests -------------------------------------------------------------
    def _submit(self, kind, volume_id, axes, images, volume):
        n_axes = int(axes.shape[0]) if axes.ndim == 2 else int(axes.size) // 9
        nn = self._sizes[volume_id]
        req = _Request(kind, int(volume_id), n_axes, _ptr(axes, "axes", 9 * n_axes),
                       _ptr(images, "images", n_axes * nn * nn),
                       _ptr(volume, "volume", nn ** 3) if volume is not None else None)
        t = ctypes.c_ulonglong()
        _lib.call("mgg_server_submit", self._h, ctypes.byref(req), ctypes.byref(t))
        self._keep[t.value] = (axes, images, volume)
        return t.value

    def submit_forward(self, volume_id: int, axes, images_out) -> int:
        """Queues the forward projection of ``axes`` (n x 9) into ``images_out`` (n x N x N)."""
        return self._submit(FORWARD, volume_id, axes, images_out, None)

    def submit_back(self, volume_id: int, axes, images, volume_out) -> int:
        """Queues getVol of the back projection of ``images`` (n x N x N) into ``volume_out`` (N^3)."""
        return self._submit(BACK, volume_id, axes, images, volume_out)

    def wait(self, ticket: int, timeout: float | None = None) -> RequestInfo:
        """Waits for a request; raises MggError on its failure (or a timeout)."""
        info = RequestInfo()
        st = _lib.lib().mgg_server_wait(self._h, ctypes.c_ulonglong(ticket),
                                        -1.0 if timeout is None else float(timeout), ctypes.byref(info))
        if st != _lib.MGG_ERR_TIMEOUT:
            self._keep.pop(ticket, None)
        _lib.check(st)
        return info

    def hold(self, on: bool = True) -> None:
        """While held, requests queue up and no batch starts; release to batch them together."""
        _lib.call("mgg_server_hold", self._h, 1 if on else 0)

    def clock(self) -> float:
        """Seconds on the server clock (the clock of RequestInfo)."""
        return _lib.lib().mgg_server_clock(self._h)

    def stats(self) -> dict:
        st = _Stats()
        _lib.call("mgg_server_get_stats", self._h, ctypes.byref(st))
        return {name: getattr(st, name) for name, _ in _Stats._fields_}
