"""Host math of the reference's Matlab layer, computed by the native library.

* :func:`kb_table`    - Matlab/getKernelFiltTable.m:1-22
* :func:`precomp`     - Matlab/getPreComp.m:1-28
* :func:`uniform_axes` - Matlab/utils/create_uniform_axes.m:16-28
* :func:`imgs_from_cas_imgs`, :func:`cas_imgs_from_imgs` - the host FFT path the
  reference runs in Matlab when RunFFTOnGPU is false
  (utils/imgsFromCASImgs.m:7-21, utils/CASImgsFromImgs.m:3-10).
"""
from __future__ import annotations

import numpy as np

from . import _lib


def kb_table(hwidth: int = 2, size: int = 501) -> np.ndarray:
    out = np.empty(size, dtype=np.float32)
    _lib.call("mgg_kb_table", out.ctypes.data, size, hwidth)
    return out


def precomp(m: int, hwidth: int = 2) -> np.ndarray:
    out = np.empty(m, dtype=np.float32)
    _lib.call("mgg_precomp", out.ctypes.data, m, hwidth)
    return out


def uniform_axes(ntheta: int, nphi: int) -> np.ndarray:
    """(ntheta*nphi, 9) float32 directions, row a = [u; v; n] of direction a."""
    out = np.empty((ntheta * nphi, 9), dtype=np.float32)
    _lib.call("mgg_uniform_axes", out.ctypes.data, ntheta, nphi)
    return out


def _from_cas_2d(r: np.ndarray) -> np.ndarray:
    m = r.shape[-1]
    idx = (m - np.arange(m)) % m
    mir = r[:, idx][:, :, idx]
    z = (0.5 * (r + mir) + 0.5j * (r - mir)).astype(np.complex64)
    z[:, 0, :] = 0
    z[:, :, 0] = 0
    return z


def imgs_from_cas_imgs(cas_imgs: np.ndarray, n: int) -> np.ndarray:
    """(P, m, m) CAS images -> (P, n, n) images on the host (single precision)."""
    cas_imgs = np.asarray(cas_imgs, dtype=np.float32)
    m = cas_imgs.shape[-1]
    z = np.fft.fftshift(_from_cas_2d(cas_imgs), axes=(1, 2))
    x = np.real(np.fft.fftshift(np.fft.ifft2(z, axes=(1, 2)), axes=(1, 2)))
    off = (m - n) // 2
    return np.ascontiguousarray(x[:, off:off + n, off:off + n], dtype=np.float32)


def cas_imgs_from_imgs(imgs: np.ndarray, m: int) -> np.ndarray:
    """(P, n, n) images -> (P, m, m) CAS images on the host (single precision)."""
    imgs = np.asarray(imgs, dtype=np.float32)
    p, n, _ = imgs.shape
    off = (m - n) // 2
    x = np.zeros((p, m, m), dtype=np.float32)
    x[:, off:off + n, off:off + n] = imgs
    f = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(1, 2)), axes=(1, 2)), axes=(1, 2))
    return np.ascontiguousarray(np.real(f) + np.imag(f), dtype=np.float32)
