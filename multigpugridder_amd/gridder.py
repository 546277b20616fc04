"""Python mirror of the reference's user-facing gridder class
(src/Matlab/MultiGPUGridder_Matlab_Class.m), driving libMultiGPUGridder.so.

Same names, argument meaning and defaults as the Matlab class:
``setVolume``, ``forwardProject``, ``backProject``, ``resetVolume``,
``getVol``, ``reconstructVol``.  Arrays are numpy float32 with the reference's
memory layout (column-major with x fastest), exposed C-ordered:

* ``Volume``     shape (N, N, N), indexed [z, y, x]
* ``Images``     shape (P, N, N), indexed [a, y, x]
* ``CASVolume``  shape (V, V, V);  ``CASImages`` shape (P, M, M)
* ``coordAxes``  shape (P, 9): row a holds [u; v; n] of direction a (the
  memory of the Matlab 9 x P matrix).

Differences from the Matlab class (all deliberate):

* Vol -> CAS and CAS -> Vol run on the GPU (the reference runs them in Matlab
  on every forwardProject / getVol, MultiGPUGridder_Matlab_Class.m:226, :323).
* ``reconstructVol`` works with RunFFTOnGPU on or off (the reference errors
  when it is on, .m:286-288) and computes the plane density in the same
  back-projection pass.
* ``interpFactor`` is kept as given; the reference silently rescales it for
  non-power-of-two N (.m:69-83).  Pass ``pow2_interp=True`` for that behaviour.
* The class's own arrays (Volume, Images, CASImages, CASVolume) live in pinned
  host memory, so the device copies run at full PCIe rate; the reference pins
  the Matlab arrays in place instead (cudaHostRegister, AbstractGridder.cpp:
  121-157), which this platform does not offer for existing memory.  Falls
  back to ordinary numpy arrays when pinned memory is not available
  (``pinned=False`` asks for them).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hostmath import cas_imgs_from_imgs, imgs_from_cas_imgs, kb_table
from .server import pinned_empty


def _copy_into(dst: np.ndarray, src) -> None:
    """``dst[...] = src`` over the host's threads.  An image stack is 655 MB at
    10,000 directions and N=128; numpy copies it on one thread (50 ms on an
    MI355X host, a third of forwardProject), torch's CPU copy on all of them
    (13 ms).  Other shapes, dtypes and layouts go through numpy."""
    src = np.asarray(src)
    t = _lib.torch
    if (t is not None and src.shape == dst.shape and src.dtype == dst.dtype == np.float32
            and src.flags.c_contiguous and src.flags.writeable and dst.flags.c_contiguous
            and dst.flags.writeable and dst.size >= (1 << 20)):
        t.from_numpy(dst).copy_(t.from_numpy(src))
    else:
        np.copyto(dst, src, casting="same_kind")


def _host_zeros(shape, pinned: bool) -> np.ndarray:
    """Zeroed float32 host array, pinned when asked and available."""
    if pinned:
        try:
            a = pinned_empty(shape)
            a[...] = 0.0
            return a
        except _lib.MggError:
            pass
    return np.zeros(shape, dtype=np.float32)


class MultiGPUGridder:
    kerHWidth = 2
    kerTblSize = 501
    extraPadding = 3

    def __init__(self, VolumeSize, NumAxes, interpFactor=2, GPUs=None, RunFFTOnGPU=1, verbose=0,
                 nStreamsFP=10, nStreamsBP=4, MaxAxesToAllocate=1000, pow2_interp=False, pinned=True):
        if not VolumeSize:
            raise ValueError("VolumeSize is a required input.")
        if not NumAxes:
            raise ValueError("NumAxes is a required input.")
        self.VolumeSize = int(VolumeSize)
        self.NumAxes = int(NumAxes)
        f = float(interpFactor)
        if pow2_interp:  # MultiGPUGridder_Matlab_Class.m:69-83
            n = self.VolumeSize
            if n < 64:
                f = 128.0 / n
            elif 64 < n < 128:
                f = 256.0 / n
            elif 128 < n < 256:
                f = 512.0 / n
        self.interpFactor = np.float32(f)
        self.GPUs = [0] if GPUs is None else [int(g) for g in GPUs]
        self.RunFFTOnGPU = int(RunFFTOnGPU)
        self.verbose = int(verbose)
        self.nStreamsFP = int(nStreamsFP)
        self.nStreamsBP = int(nStreamsBP)
        self.MaxAxesToAllocate = int(MaxAxesToAllocate)
        n = self.VolumeSize
        self.M = int(np.float32(n) * self.interpFactor)
        self.V = self.M + 2 * self.extraPadding
        self.MaskRadius = np.float32((np.float32(n) * self.interpFactor) / 2 - 1)

        self.Volume = _host_zeros((n, n, n), pinned)
        self.Images = _host_zeros((self.NumAxes, n, n), pinned)
        self.KBTable = kb_table(self.kerHWidth, self.kerTblSize)
        self.CASImages = None
        if not self.RunFFTOnGPU or self.verbose:
            self.CASImages = _host_zeros((self.NumAxes, self.M, self.M), pinned)
        self.CASVolume = _host_zeros((self.V, self.V, self.V), pinned)
        self.coordAxes = np.zeros((self.NumAxes, 9), dtype=np.float32)
        self._numAxes = self.NumAxes

        h = ctypes.c_void_p()
        gpus = _lib.int_array(self.GPUs)
        _lib.call("mgg_create", n, self.NumAxes, float(self.interpFactor), self.extraPadding,
                  len(self.GPUs), gpus, self.RunFFTOnGPU, self.verbose, ctypes.byref(h))
        self.objectHandle = h
        self.Set()

    # -- lifetime ---------------------------------------------------------
    def delete(self):
        if getattr(self, "objectHandle", None):
            _lib.lib().mgg_destroy(self.objectHandle)
            self.objectHandle = None

    def __del__(self):
        try:
            self.delete()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.delete()

    # -- Set(): push pointers and parameters (.m:155-195) -------------------
    def Set(self):
        h = self.objectHandle
        if self.Volume.shape != (self.VolumeSize,) * 3:
            raise ValueError("Volume must be a cube of VolumeSize")
        if self.Images.shape[1:] != (self.VolumeSize, self.VolumeSize):
            raise ValueError("Images must be the same size as Volume")
        for name in ("Volume", "Images", "coordAxes", "CASVolume", "KBTable"):
            a = getattr(self, name)
            if a.dtype != np.float32 or not a.flags.c_contiguous:
                setattr(self, name, np.ascontiguousarray(a, dtype=np.float32))
        _lib.call("mgg_set_coord_axes", h, self.coordAxes.ctypes.data,
                  _lib.int_array([self.coordAxes.size, 1]))
        _lib.call("mgg_set_volume", h, self.Volume.ctypes.data, _lib.int_array(self.Volume.shape))
        _lib.call("mgg_set_images", h, self.Images.ctypes.data,
                  _lib.int_array([self.VolumeSize, self.VolumeSize, self.Images.shape[0]]))
        _lib.call("mgg_set_kb_table", h, self.KBTable.ctypes.data, _lib.int_array([self.KBTable.size]))
        _lib.call("mgg_set_num_streams_fp", h, self.nStreamsFP)
        _lib.call("mgg_set_num_streams_bp", h, self.nStreamsBP)
        _lib.call("mgg_set_mask_radius", h, float(self.MaskRadius))
        _lib.call("mgg_set_num_axes", h, self._numAxes)
        _lib.call("mgg_set_cas_volume", h, self.CASVolume.ctypes.data, _lib.int_array(self.CASVolume.shape))
        if self.CASImages is not None:
            _lib.call("mgg_set_cas_images", h, self.CASImages.ctypes.data,
                      _lib.int_array([self.M, self.M, self.CASImages.shape[0]]))
        _lib.call("mgg_set_max_axes_to_allocate", h, self.MaxAxesToAllocate)

    def _set_axes(self, axes):
        axes = np.asarray(axes, dtype=np.float32).reshape(-1, 9)
        p = axes.shape[0]
        if p > self.coordAxes.shape[0]:
            raise ValueError(f"{p} directions exceed NumAxes={self.coordAxes.shape[0]}")
        self.coordAxes[:p] = axes
        self._numAxes = p
        _lib.call("mgg_set_num_axes", self.objectHandle, p)
        return p

    # -- forwardProject (.m:198-243) ------------------------------------------
    def forwardProject(self, *args):
        """forwardProject() | forwardProject(axes) | forwardProject(vol, axes, maskRadius).

        Returns the (P, N, N) projection images of the current volume."""
        if len(args) == 1:
            self._set_axes(args[0])
        elif len(args) == 3:
            self.Volume[...] = np.asarray(args[0], dtype=np.float32)
            self._set_axes(args[1])
            self.MaskRadius = np.float32(args[2])
            _lib.call("mgg_set_mask_radius", self.objectHandle, float(self.MaskRadius))
        elif args:
            raise TypeError("forwardProject(): Unknown input")
        if self._numAxes < 1:
            raise ValueError("no projection directions set")
        _lib.call("mgg_volume_to_cas_volume", self.objectHandle)
        _lib.call("mgg_forward_project", self.objectHandle)
        p = self._numAxes
        if not self.RunFFTOnGPU:
            self.Images[:p] = imgs_from_cas_imgs(self.CASImages[:p], self.VolumeSize)
        out = np.empty((p,) + self.Images.shape[1:], dtype=np.float32)
        _copy_into(out, self.Images[:p])
        return out

    # -- backProject (.m:245-270) ---------------------------------------------
    def backProject(self, images=None, axes=None):
        if images is not None:
            images = np.asarray(images, dtype=np.float32)
            if images.ndim == 2:
                images = images[None]
            _copy_into(self.Images[:images.shape[0]], images)
            if axes is None:
                raise TypeError("backProject(images, axes) needs the directions")
            self._set_axes(axes)
            if not self.RunFFTOnGPU:
                p = self._numAxes
                self.CASImages[:p] = cas_imgs_from_imgs(self.Images[:p], self.M)
        _lib.call("mgg_back_project", self.objectHandle)

    # -- setVolume / resetVolume (.m:272-282) ---------------------------------
    def setVolume(self, vol):
        self.Volume[...] = np.asarray(vol, dtype=np.float32)

    def resetVolume(self):
        _lib.call("mgg_reset_volume", self.objectHandle)

    # -- getVol (.m:314-325) ----------------------------------------------------
    def getVol(self):
        _lib.call("mgg_get_volume", self.objectHandle)
        return self.Volume.copy()

    # -- reconstructVol (.m:284-312) ----------------------------------------
    def reconstructVol(self, images=None, axes=None):
        """Back project, divide by the plane density (+1e-6, FP64) and return the volume."""
        if images is not None:
            images = np.asarray(images, dtype=np.float32)
            _copy_into(self.Images[:images.shape[0]], images)
            if axes is None:
                raise TypeError("reconstructVol(images, axes) needs the directions")
            self._set_axes(axes)
            if not self.RunFFTOnGPU:
                p = self._numAxes
                self.CASImages[:p] = cas_imgs_from_imgs(self.Images[:p], self.M)
        _lib.call("mgg_reconstruct_volume", self.objectHandle)
        return self.Volume.copy()

    # -- CAS-level access (the MEX-level calls) ---------------------------------
    def casVolumeToHost(self):
        """MultiGPUGridder::CASVolumeToVolume: device CAS volume -> self.CASVolume."""
        _lib.call("mgg_cas_volume_to_host", self.objectHandle)
        return self.CASVolume.copy()
