"""CPU oracle for the Fourier-domain projection gridder.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker: the shipped library never routes through it.

Two halves:

* the projection kernels, restated in C (``gridder_oracle.c``, FP64
  accumulation, FP32 decision arithmetic pinned to the reference's op order)
  and reached through ctypes;
* the host math and FFT chains, restated here in numpy FP64:

  ============================  =============================================
  function                      reference it follows
  ============================  =============================================
  ``kb_table``                  Matlab/getKernelFiltTable.m:1-22,
                                utils/gridLibKaiser.m:1-15, utils/kernelAlpha.m:6
  ``precomp``                   Matlab/getPreComp.m:1-28
  ``vol_to_cas``                Matlab/utils/CASFromVol_Gridder.m:26-60, ToCAS.m
  ``cas_to_vol``                Matlab/utils/volFromCAS_Gridder.m:27-56, FromCAS.m
  ``cas_imgs_to_imgs``          Gridders/gpuGridder.cpp:432-518 (equal to
                                utils/imgsFromCASImgs.m:7-21)
  ``imgs_to_cas_imgs``          Gridders/gpuGridder.cpp:520-596 (equal to
                                utils/CASImgsFromImgs.m:3-10)
  ``uniform_axes``              Matlab/utils/create_uniform_axes.m:16-28,
                                utils/get_axes_from_n_rot.m:7-37
  ``reconstruct``               Matlab/MultiGPUGridder_Matlab_Class.m:284-312
  filters (``cas_to_complex``   Filters/*.cu kernels (CASToComplexFilter.cu:3-81,
  ... ``add_volume``)           ComplexToCASFilter.cu:3-26, ...)
  ============================  =============================================

Array conventions: a volume is a numpy array of shape (n, n, n) indexed
[z, y, x] in C order, so its memory is the reference's column-major
``vol[x + y*n + z*n*n]``; an image stack has shape (P, m, m) indexed
[a, j, i] (``img[i + j*m + a*m*m]``); directions are (P, 9) float32 rows
``[u; v; n]``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        lib = ctypes.CDLL(path)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ci, cf = ctypes.c_int, ctypes.c_float
        lib.orc_forward_project.argtypes = [f32p, ci, f32p, ci, f32p, ci, cf, ci, cf, f64p]
        lib.orc_back_project.argtypes = [f64p, ci, f32p, ci, f32p, ci, cf, f32p, ci, cf]
        lib.orc_back_project_adjoint.argtypes = [f64p, ci, f32p, ci, f32p, ci, cf, f32p, ci, cf]
        lib.cpu_forward_project_f32.argtypes = [f32p, ci, f32p, ci, f32p, ci, cf, ci, cf, f32p, ci]
        lib.cpu_back_project_f32.argtypes = [f32p, ci, f32p, ci, f32p, ci, cf, f32p, ci, cf, ci]
        lib.cpu_max_threads.restype = ci
        lib.orc_set_reading.argtypes = [ci]
        lib.orc_set_reading.restype = ci
        lib.orc_get_reading.restype = ci
        for fn in ("orc_forward_project", "orc_back_project", "orc_back_project_adjoint",
                   "cpu_forward_project_f32", "cpu_back_project_f32"):
            getattr(lib, fn).restype = None
        _LIB = lib
    return _LIB


# Arithmetic readings of the reference kernels' decision expressions (the
# reference is built by nvcc with no flags, src/CMakeLists.txt:80-83, and
# nvcc's default --fmad=true contracts multiply-adds).  gridder_oracle.c
# documents each; READING_FIRST is the default and the one the HIP kernels
# follow.
READING_SOURCE, READING_FIRST, READING_SECOND, READING_REASSOC, READING_INDEX = range(5)
READINGS = {
    "source": READING_SOURCE,
    "first": READING_FIRST,
    "second": READING_SECOND,
    "reassoc": READING_REASSOC,
    "index": READING_INDEX,
}


class reading:
    """Context manager selecting the oracle's arithmetic reading:
    ``with reading("source"): ...``.  Process-wide (not thread-safe)."""

    def __init__(self, name):
        self.value = READINGS[name] if isinstance(name, str) else int(name)

    def __enter__(self):
        prev = _lib().orc_set_reading(self.value)
        if prev < 0:
            raise ValueError(f"unknown reading {self.value}")
        self.prev = prev
        return self

    def __exit__(self, *exc):
        _lib().orc_set_reading(self.prev)
        return False


def current_reading() -> int:
    return int(_lib().orc_get_reading())


# ----------------------------------------------------------------------------
# Kaiser-Bessel table and precompensation (host math)
# ----------------------------------------------------------------------------
_ALPHA = {1: 5.5, 2: 10.0, 3: 10.2, 4: 13.0, 5: 17.0}  # utils/kernelAlpha.m:6


def _kaiser(hw: int, alpha: float, x: np.ndarray) -> np.ndarray:
    """Kaiser-Bessel window I0(alpha*sqrt(1-(x/hw)^2))/I0(alpha), 0 outside
    |x|<=hw (utils/gridLibKaiser.m:1-15)."""
    from scipy.special import i0

    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    inside = np.abs(x) <= hw
    r = np.sqrt(1.0 - (x[inside] / hw) ** 2)
    out[inside] = i0(alpha * r) / i0(alpha)
    return out


def kb_table(hw: int = 2, size: int = 501) -> np.ndarray:
    """Kernel lookup table (Matlab/getKernelFiltTable.m:1-22), FP64.

    Sampled at x = -hw:delta:hw, normalised to unit integral, shifted so both
    ends are 0, then scaled so that the (2hw+1)^3 integer-offset taps of the
    separable 3-D kernel sum to 1."""
    alpha = _ALPHA[hw]
    delta = 2.0 * hw / (size - 1)
    x = (np.arange(size) - (size - 1) / 2.0) * delta
    w = _kaiser(hw, alpha, x)
    w = w / np.sum(w * delta)
    w = w - w[0]
    w[-1] = 0.0
    center = (size + 1) // 2  # 1-based
    scale = (center - 1) / hw
    idx = (np.arange(-hw, hw + 1) * scale + center).astype(np.int64) - 1
    ker = w[idx]
    denom = np.sum(np.kron(np.kron(ker, ker[:, None]), ker))
    return w / denom ** (1.0 / 3.0)


def precomp(m: int, hw: int = 2) -> np.ndarray:
    """Per-axis precompensation vector of length m (Matlab/getPreComp.m:1-28):
    the reciprocal of the window's transform, normalised to 1 at the centre."""
    ovr = 4
    alpha = _ALPHA[hw]
    x = np.arange(-hw * ovr, hw * ovr + 1) / ovr
    w = _kaiser(hw, alpha, x)
    w = w - w[0]
    w[-1] = 0.0
    half = (len(x) - 1) // 2
    line = np.zeros(m * ovr)
    c = len(line) // 2  # 0-based centre (Matlab numel/2+1)
    line[c - half:c + half + 1] = w
    p = np.real(np.fft.fftshift(np.fft.fft(np.fft.fftshift(line))))
    pc = m * ovr // 2
    return p[pc] / p[pc - m // 2:pc + m // 2]


# ----------------------------------------------------------------------------
# Direction generators
# ----------------------------------------------------------------------------
def axes_from_normal(n, phi_r: float) -> np.ndarray:
    """utils/get_axes_from_n_rot.m:7-37: rotation whose third column is n,
    followed by an in-plane rotation by phi_r; returns the 9 column values."""
    n = np.asarray(n, dtype=np.float64)
    a = np.cross([0.0, 0.0, 1.0], n)
    if np.sqrt(np.sum(a * a)) < 1e-4:
        t1 = np.eye(3)
        t2 = np.eye(3)
    else:
        phi = np.arctan2(n[1], n[0])
        t1 = np.array([[np.cos(phi), -np.sin(phi), 0], [np.sin(phi), np.cos(phi), 0], [0, 0, 1.0]])
        th = np.arctan2(np.sqrt(n[0] ** 2 + n[1] ** 2), n[2])
        t2 = np.array([[np.cos(th), 0, np.sin(th)], [0, 1.0, 0], [-np.sin(th), 0, np.cos(th)]])
    t3 = np.array([[np.cos(phi_r), -np.sin(phi_r), 0], [np.sin(phi_r), np.cos(phi_r), 0], [0, 0, 1.0]])
    t = (t1 @ t2) @ t3
    return t.T.reshape(9)  # [T(:,1); T(:,2); T(:,3)]


def uniform_axes(ntheta: int, nphi: int) -> np.ndarray:
    """utils/create_uniform_axes.m:16-28 -> (ntheta*nphi, 9) float32."""
    al = np.pi / 180.0
    out = []
    for kt in range(1, ntheta + 1):
        theta = kt * (90.0 / ntheta)
        for kp in range(nphi):
            phi = kp * (360.0 / nphi)
            z = [np.sin(theta * al) * np.cos(phi * al), np.sin(theta * al) * np.sin(phi * al),
                 np.cos(theta * al)]
            out.append(axes_from_normal(z, 30 * al))
    return np.asarray(out, dtype=np.float32)


# ----------------------------------------------------------------------------
# Projection kernels (C restatement)
# ----------------------------------------------------------------------------
def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def forward_project(cas_vol, axes, ker, m, mask_radius=None, hw=2.0):
    """Reference-semantics forward projection -> (P, m, m) float64 CAS images."""
    cas_vol = _c32(cas_vol)
    axes = _c32(axes).reshape(-1, 9)
    ker = _c32(ker)
    v = cas_vol.shape[0]
    if mask_radius is None:
        mask_radius = float(m // 2 - 1)
    out = np.zeros((axes.shape[0], m, m), dtype=np.float64)
    _lib().orc_forward_project(cas_vol.reshape(-1), v, axes.reshape(-1), axes.shape[0], ker,
                               ker.size, hw, m, mask_radius, out.reshape(-1))
    return out


def back_project(cas_imgs, axes, ker, v, mask_radius=None, hw=2.0, out=None):
    """Reference-semantics back projection -> (v, v, v) float64 CAS volume."""
    cas_imgs = _c32(cas_imgs)
    axes = _c32(axes).reshape(-1, 9)
    ker = _c32(ker)
    m = cas_imgs.shape[-1]
    if mask_radius is None:
        mask_radius = float(m // 2 - 1)
    if out is None:
        out = np.zeros((v, v, v), dtype=np.float64)
    _lib().orc_back_project(out.reshape(-1), v, axes.reshape(-1), axes.shape[0],
                            cas_imgs.reshape(-1), m, mask_radius, ker, ker.size, hw)
    return out


def back_project_adjoint(cas_imgs, axes, ker, v, mask_radius=None, hw=2.0, out=None):
    """Exact transpose of :func:`forward_project` -> (v, v, v) float64."""
    cas_imgs = _c32(cas_imgs)
    axes = _c32(axes).reshape(-1, 9)
    ker = _c32(ker)
    m = cas_imgs.shape[-1]
    if mask_radius is None:
        mask_radius = float(m // 2 - 1)
    if out is None:
        out = np.zeros((v, v, v), dtype=np.float64)
    _lib().orc_back_project_adjoint(out.reshape(-1), v, axes.reshape(-1), axes.shape[0],
                                    cas_imgs.reshape(-1), m, mask_radius, ker, ker.size, hw)
    return out


def cpu_forward_project_f32(cas_vol, axes, ker, m, mask_radius=None, hw=2.0, nthreads=0):
    """The timed CPU path (FP32, OpenMP) -> (P, m, m) float32."""
    cas_vol = _c32(cas_vol)
    axes = _c32(axes).reshape(-1, 9)
    ker = _c32(ker)
    if mask_radius is None:
        mask_radius = float(m // 2 - 1)
    out = np.zeros((axes.shape[0], m, m), dtype=np.float32)
    _lib().cpu_forward_project_f32(cas_vol.reshape(-1), cas_vol.shape[0], axes.reshape(-1),
                                   axes.shape[0], ker, ker.size, hw, m, mask_radius,
                                   out.reshape(-1), nthreads)
    return out


def cpu_back_project_f32(cas_imgs, axes, ker, v, mask_radius=None, hw=2.0, nthreads=0):
    cas_imgs = _c32(cas_imgs)
    axes = _c32(axes).reshape(-1, 9)
    ker = _c32(ker)
    m = cas_imgs.shape[-1]
    if mask_radius is None:
        mask_radius = float(m // 2 - 1)
    out = np.zeros((v, v, v), dtype=np.float32)
    _lib().cpu_back_project_f32(out.reshape(-1), v, axes.reshape(-1), axes.shape[0],
                                cas_imgs.reshape(-1), m, mask_radius, ker, ker.size, hw, nthreads)
    return out


def cpu_threads() -> int:
    return int(_lib().cpu_max_threads())


# ----------------------------------------------------------------------------
# CAS / FFT chains (numpy FP64)
# ----------------------------------------------------------------------------
def to_cas(c):
    """utils/ToCAS.m:5 and ComplexToCASFilter.cu:21: re + im."""
    return np.real(c) + np.imag(c)


def from_cas(r, axes):
    """utils/FromCAS.m:7-13 / CASToComplexFilter.cu:24-75 over the given axes:
    Z = 0.5(r + r_mirror) + 0.5i(r - r_mirror) with mirror index n-i, and the
    index-0 row/column/slice set to 0."""
    r = np.asarray(r, dtype=np.float64)
    mir = r
    for ax in axes:
        n = r.shape[ax]
        idx = (n - np.arange(n)) % n
        mir = np.take(mir, idx, axis=ax)
    z = 0.5 * (r + mir) + 0.5j * (r - mir)
    for ax in axes:
        sl = [slice(None)] * r.ndim
        sl[ax] = 0
        z[tuple(sl)] = 0
    return z


def vol_to_cas(vol, interp_factor=2, extra_padding=3, hw=2):
    """Spatial volume (n^3) -> padded CAS volume (V^3), utils/CASFromVol_Gridder.m."""
    vol = np.asarray(vol, dtype=np.float64)
    n = vol.shape[0]
    m = int(round(n * interp_factor))
    off = (m - n) // 2
    x = np.zeros((m, m, m))
    x[off:off + n, off:off + n, off:off + n] = vol
    p = precomp(m, hw)
    x *= p[:, None, None] * p[None, :, None] * p[None, None, :]
    f = np.fft.fftshift(np.fft.fftn(np.fft.fftshift(x)))
    cas = to_cas(f)
    v = m + 2 * extra_padding
    out = np.zeros((v, v, v))
    e = extra_padding
    out[e:e + m, e:e + m, e:e + m] = cas
    return out


def cas_to_vol(cas_vol, n, interp_factor=2, extra_padding=3, hw=2, zero_nyquist=True):
    """Padded CAS volume -> spatial volume (n^3), utils/volFromCAS_Gridder.m.

    ``zero_nyquist=False`` keeps the index-0 planes (used to build the exact
    transpose of :func:`vol_to_cas`)."""
    cas_vol = np.asarray(cas_vol, dtype=np.float64)
    m = int(round(n * interp_factor))
    e = extra_padding
    c = cas_vol[e:e + m, e:e + m, e:e + m]
    if zero_nyquist:
        z = from_cas(c, (0, 1, 2))
    else:
        mir = c[::-1, ::-1, ::-1]
        mir = np.roll(mir, 1, axis=(0, 1, 2))
        z = 0.5 * (c + mir) + 0.5j * (c - mir)
    x = np.real(np.fft.fftshift(np.fft.ifftn(np.fft.fftshift(z))))
    p = precomp(m, hw)
    x *= p[:, None, None] * p[None, :, None] * p[None, None, :]
    off = (m - n) // 2
    return x[off:off + n, off:off + n, off:off + n]


def cas_imgs_to_imgs(cas_imgs, n):
    """CAS images (P, m, m) -> images (P, n, n): gpuGridder.cpp:432-518."""
    cas_imgs = np.asarray(cas_imgs, dtype=np.float64)
    m = cas_imgs.shape[-1]
    z = from_cas(cas_imgs, (1, 2))
    z = np.fft.fftshift(z, axes=(1, 2))
    x = np.fft.ifft2(z, axes=(1, 2))  # includes the 1/m^2 of DivideScalarFilter
    x = np.real(np.fft.fftshift(x, axes=(1, 2)))
    off = (m - n) // 2
    return x[:, off:off + n, off:off + n]


def imgs_to_cas_imgs(imgs, m):
    """Images (P, n, n) -> CAS images (P, m, m): gpuGridder.cpp:520-596."""
    imgs = np.asarray(imgs, dtype=np.float64)
    p, n, _ = imgs.shape
    off = (m - n) // 2
    x = np.zeros((p, m, m))
    x[:, off:off + n, off:off + n] = imgs
    f = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(1, 2)), axes=(1, 2)), axes=(1, 2))
    return to_cas(f)


def cas_imgs_to_imgs_transpose(imgs, m):
    """Transpose of :func:`cas_imgs_to_imgs`: (1/m^2) * imgs_to_cas_imgs with
    the index-0 row and column of each CAS image zeroed (FromCAS discards
    them).  Pinned against the explicit matrix in the tests."""
    c = imgs_to_cas_imgs(imgs, m) / float(m * m)
    c[:, 0, :] = 0.0
    c[:, :, 0] = 0.0
    return c


def vol_to_cas_transpose(cas_vol, n, interp_factor=2, extra_padding=3, hw=2):
    """Transpose of :func:`vol_to_cas`: m^3 * cas_to_vol keeping the index-0
    planes.  Pinned against the explicit matrix in the tests."""
    m = int(round(n * interp_factor))
    return float(m) ** 3 * cas_to_vol(cas_vol, n, interp_factor, extra_padding, hw, zero_nyquist=False)


def projector(vol, axes, ker, interp_factor=2, extra_padding=3):
    """The full spatial projector images = CASImgsToImgs(Gather(CASFromVol(vol))),
    FP64 except the FP32 CAS volume the gather reads (as the GPU path)."""
    n = np.asarray(vol).shape[0]
    m = int(round(n * interp_factor))
    cas = vol_to_cas(vol, interp_factor, extra_padding).astype(np.float32)
    return cas_imgs_to_imgs(forward_project(cas, axes, ker, m), n)


def projector_transpose(imgs, axes, ker, n, interp_factor=2, extra_padding=3):
    """Exact transpose of :func:`projector` (FP64), composed stage by stage."""
    m = int(round(n * interp_factor))
    v = m + 2 * extra_padding
    c = cas_imgs_to_imgs_transpose(imgs, m)
    vol_c = back_project_adjoint(c.astype(np.float32), axes, ker, v)
    return vol_to_cas_transpose(vol_c, n, interp_factor, extra_padding)


def reconstruct(cas_vol_bp, cas_vol_density, n, interp_factor=2, extra_padding=3, hw=2):
    """MultiGPUGridder_Matlab_Class.m:305-309: CAS(BP) / (CAS(density)+1e-6)
    then CAS -> volume."""
    ratio = np.asarray(cas_vol_bp, np.float64) / (np.asarray(cas_vol_density, np.float64) + 1e-6)
    return cas_to_vol(ratio, n, interp_factor, extra_padding, hw)


# ----------------------------------------------------------------------------
# Elementwise filters (Filters/*.cu), restated for the bit-exact filter KATs
# ----------------------------------------------------------------------------
def fftshift2d(x):
    """FFTShift2DFilter.cu:3-62 on a (P, m, m) stack (even m)."""
    return np.fft.fftshift(x, axes=(-2, -1))


def fftshift3d(x):
    """FFTShift3DFilter.cu:3-98 (even n)."""
    return np.fft.fftshift(x)


def pad_images(x, m):
    """PadVolumeFilter.cu:4-36 with padding (m-n)/2 in x and y, 0 in z."""
    p, n, _ = x.shape
    off = (m - n) // 2
    out = np.zeros((p, m, m), dtype=x.dtype)
    out[:, off:off + n, off:off + n] = x
    return out


def crop_images(x, n):
    """CropVolumeFilter.cu:3-39 with crop (m-n)/2 in x and y, 0 in z."""
    m = x.shape[-1]
    off = (m - n) // 2
    return x[:, off:off + n, off:off + n].copy()


def divide_volume(a, b):
    """DivideVolumeFilter.cu:24-31: a/b where both are non-zero, else 0."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    out = np.zeros_like(a)
    ok = (a != 0) & (b != 0)
    out[ok] = a[ok] / b[ok]
    return out
