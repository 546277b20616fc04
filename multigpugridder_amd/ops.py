"""Device-pointer operators on torch tensors (layer 2 of include/mgg.h).

torch supplies device memory and streams only; every computation is a kernel
of libMultiGPUGridder.so launched on torch's current stream.  Tensors must be
contiguous and on the current CUDA (HIP) device.

Shapes (C order, x fastest in memory as in the reference):
  CAS volume (V, V, V); CAS images (P, M, M); images (P, N, N);
  volume (N, N, N); directions (P, 9).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .hostmath import kb_table as _kb_table_host

_DTYPES = {torch.float32: _lib.DTYPE_F32, torch.bfloat16: _lib.DTYPE_BF16, torch.float16: _lib.DTYPE_F16}
_ker_cache: dict = {}


def _stream():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(x):
    import ctypes
    return ctypes.c_void_p(int(x))


def _ptr(t: torch.Tensor):
    return ctypes_void(t.data_ptr())


def _check(t: torch.Tensor, name: str, dtype=torch.float32):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


# Shape checks run before the kernels see a pointer: a kernel trusts the sizes
# it is given, so a tensor smaller than what it reads or writes would be an
# out-of-bounds device access.  They come before the device checks so that
# they can be exercised without a GPU.
def _need_numel(t: torch.Tensor, n: int, name: str, exact: bool = True):
    if (t.numel() != n) if exact else (t.numel() < n):
        what = "exactly" if exact else "at least"
        raise ValueError(f"{name} must hold {what} {n} elements, got {t.numel()} (shape {tuple(t.shape)})")


def _cube_edge(t: torch.Tensor, name: str) -> int:
    if t.dim() != 3 or not (t.shape[0] == t.shape[1] == t.shape[2]) or t.shape[0] == 0:
        raise ValueError(f"{name} must be a non-empty cube (V, V, V), got shape {tuple(t.shape)}")
    return int(t.shape[0])


def _stack_edge(t: torch.Tensor, name: str) -> tuple:
    """(P, S, S) stack of square images -> (P, S)."""
    if t.dim() != 3 or t.shape[1] != t.shape[2] or t.shape[1] == 0:
        raise ValueError(f"{name} must be a stack of square images (P, S, S), got shape {tuple(t.shape)}")
    return int(t.shape[0]), int(t.shape[1])


def _n_axes(axes: torch.Tensor) -> int:
    if axes.numel() % 9:
        raise ValueError(f"axes must hold 9 floats per direction, got {axes.numel()} elements")
    return axes.numel() // 9


def _same_device(*named):
    devs = {(n, t.device) for n, t in named if t is not None}
    if len({d for _, d in devs}) > 1:
        raise ValueError("tensors on different devices: " + ", ".join(f"{n} on {d}" for n, d in sorted(devs, key=str)))


def kb_table(device=None, hwidth: int = 2, size: int = 501) -> torch.Tensor:
    """Device copy of the Kaiser-Bessel table (cached per device)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    key = (dev, hwidth, size)
    t = _ker_cache.get(key)
    if t is None:
        t = torch.from_numpy(_kb_table_host(hwidth, size)).to(dev)
        _ker_cache[key] = t
    return t


def default_mask_radius(m: int) -> float:
    return float(np.float32(m) / 2 - 1)


def check_volume(cas_volume: torch.Tensor, ker: torch.Tensor | None = None,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """One-element int32 device flag: 1 when ker is the 501-entry table with
    finite entries and zero ends and every value of cas_volume is finite.

    With it forward_project skips each stencil axis' end tap whose weight is
    table entry 0 or 500 (both 0): same results, smaller staged voxel boxes.
    The flag is valid for this volume and table until either changes."""
    if cas_volume.dtype not in _DTYPES:
        raise ValueError(f"unsupported CAS volume dtype {cas_volume.dtype}")
    v = _cube_edge(cas_volume, "cas_volume")
    if out is not None:
        _need_numel(out, 1, "out")
    _same_device(("cas_volume", cas_volume), ("ker", ker), ("out", out))
    _check(cas_volume, "cas_volume", None)
    ker = kb_table(cas_volume.device) if ker is None else ker
    _check(ker, "ker")
    if out is None:
        out = torch.empty(1, device=cas_volume.device, dtype=torch.int32)
    _check(out, "out", torch.int32)
    _lib.call("mgg_op_check_volume", _ptr(cas_volume), _DTYPES[cas_volume.dtype], v, _ptr(ker), ker.numel(),
              _ptr(out), _stream())
    return out


def forward_project(cas_volume: torch.Tensor, axes: torch.Tensor, m: int, mask_radius=None,
                    ker: torch.Tensor | None = None, hwidth: float = 2.0,
                    out: torch.Tensor | None = None, check=None) -> torch.Tensor:
    """Reference-semantics forward projection -> (P, m, m) FP32 CAS images.

    cas_volume may be float32, bfloat16 or float16 (gather in FP32).  check:
    None checks the volume and table on this call (check_volume) and uses the
    flag; a flag from check_volume for this volume and table is used as it is;
    False gathers every tap of every stencil.  The results are the same."""
    if cas_volume.dtype not in _DTYPES:
        raise ValueError(f"unsupported CAS volume dtype {cas_volume.dtype}")
    v = _cube_edge(cas_volume, "cas_volume")
    p = _n_axes(axes)
    if int(m) <= 0:
        raise ValueError(f"image size m must be positive, got {m}")
    if out is not None:
        _need_numel(out, p * m * m, "out")
    _same_device(("cas_volume", cas_volume), ("axes", axes), ("out", out), ("ker", ker))
    _check(cas_volume, "cas_volume", None)
    _check(axes, "axes")
    if out is None:
        out = torch.empty((p, m, m), device=cas_volume.device, dtype=torch.float32)
    _check(out, "out")
    ker = kb_table(cas_volume.device) if ker is None else ker
    _check(ker, "ker")
    mr = default_mask_radius(m) if mask_radius is None else float(mask_radius)
    if check is False:
        _lib.call("mgg_op_forward_project", _ptr(cas_volume), _DTYPES[cas_volume.dtype], v, _ptr(axes), p,
                  _ptr(out), m, mr, _ptr(ker), ker.numel(), float(hwidth), _stream())
        return out
    if check is None or check is True:
        check = check_volume(cas_volume, ker)
    else:
        _need_numel(check, 1, "check")
        _same_device(("cas_volume", cas_volume), ("check", check))
        _check(check, "check", torch.int32)
    _lib.call("mgg_op_forward_project_checked", _ptr(cas_volume), _DTYPES[cas_volume.dtype], v, _ptr(axes), p,
              _ptr(out), m, mr, _ptr(ker), ker.numel(), float(hwidth), _ptr(check), _stream())
    return out


def back_project(cas_images: torch.Tensor, axes: torch.Tensor, v: int, mask_radius=None,
                 ker: torch.Tensor | None = None, hwidth: float = 2.0, out: torch.Tensor | None = None,
                 density: torch.Tensor | None = None) -> torch.Tensor:
    """Reference-semantics back projection, accumulated into ``out`` (V^3, zeroed if new)."""
    pi, m = _stack_edge(cas_images, "cas_images")
    p = _n_axes(axes)
    if pi < p:
        raise ValueError(f"cas_images holds {pi} images for {p} directions")
    if int(v) <= 0:
        raise ValueError(f"volume size v must be positive, got {v}")
    if out is not None:
        _need_numel(out, v * v * v, "out")
    if density is not None:
        _need_numel(density, v * v * v, "density")
    _same_device(("cas_images", cas_images), ("axes", axes), ("out", out), ("density", density), ("ker", ker))
    _check(cas_images, "cas_images")
    _check(axes, "axes")
    if out is None:
        out = torch.zeros((v, v, v), device=cas_images.device, dtype=torch.float32)
    _check(out, "out")
    if density is not None:
        _check(density, "density")
    ker = kb_table(cas_images.device) if ker is None else ker
    _check(ker, "ker")
    mr = default_mask_radius(m) if mask_radius is None else float(mask_radius)
    _lib.call("mgg_op_back_project", _ptr(out), _ptr(density) if density is not None else None, v,
              _ptr(axes), p, _ptr(cas_images), m, mr, _ptr(ker), ker.numel(), float(hwidth), _stream())
    return out


def check_images(cas_images: torch.Tensor, ker: torch.Tensor | None = None,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """check_volume for a (P, M, M) stack of FP32 CAS images: one-element int32
    device flag, 1 when ker is the 501-entry table with finite entries and
    zero ends and every image value is finite."""
    pi, m = _stack_edge(cas_images, "cas_images")
    if out is not None:
        _need_numel(out, 1, "out")
    _same_device(("cas_images", cas_images), ("ker", ker), ("out", out))
    _check(cas_images, "cas_images")
    ker = kb_table(cas_images.device) if ker is None else ker
    _check(ker, "ker")
    if out is None:
        out = torch.empty(1, device=cas_images.device, dtype=torch.int32)
    _check(out, "out", torch.int32)
    _lib.call("mgg_op_check_images", _ptr(cas_images), pi, m, _ptr(ker), ker.numel(), _ptr(out), _stream())
    return out


def back_project_adjoint(cas_images: torch.Tensor, axes: torch.Tensor, v: int, mask_radius=None,
                         ker: torch.Tensor | None = None, hwidth: float = 2.0,
                         out: torch.Tensor | None = None, check=None) -> torch.Tensor:
    """Exact transpose of :func:`forward_project`, accumulated into ``out``.

    check: None checks the images and table on this call (check_images) and
    lets the scatter skip the taps of weight 0; a flag from check_images for
    these images and table is used as it is; False scatters every tap.  The
    results are the same."""
    pi, m = _stack_edge(cas_images, "cas_images")
    p = _n_axes(axes)
    if pi < p:
        raise ValueError(f"cas_images holds {pi} images for {p} directions")
    if int(v) <= 0:
        raise ValueError(f"volume size v must be positive, got {v}")
    if out is not None:
        _need_numel(out, v * v * v, "out")
    _same_device(("cas_images", cas_images), ("axes", axes), ("out", out), ("ker", ker))
    _check(cas_images, "cas_images")
    _check(axes, "axes")
    if out is None:
        out = torch.zeros((v, v, v), device=cas_images.device, dtype=torch.float32)
    _check(out, "out")
    ker = kb_table(cas_images.device) if ker is None else ker
    _check(ker, "ker")
    mr = default_mask_radius(m) if mask_radius is None else float(mask_radius)
    if p == 0:
        return out
    if check is False:
        _lib.call("mgg_op_back_project_adjoint", _ptr(out), v, _ptr(axes), p, _ptr(cas_images), m, mr,
                  _ptr(ker), ker.numel(), float(hwidth), _stream())
        return out
    if check is None or check is True:
        check = check_images(cas_images[:p], ker)
    else:
        _need_numel(check, 1, "check")
        _same_device(("cas_images", cas_images), ("check", check))
        _check(check, "check", torch.int32)
    _lib.call("mgg_op_back_project_adjoint_checked", _ptr(out), v, _ptr(axes), p, _ptr(cas_images), m, mr,
              _ptr(ker), ker.numel(), float(hwidth), _ptr(check), _stream())
    return out


def cas_images_to_images(cas_images: torch.Tensor, n: int, out: torch.Tensor | None = None,
                         clobber: bool = False) -> torch.Tensor:
    """(P, m, m) CAS images -> (P, n, n) images.  The kernel uses its input as
    scratch; unless ``clobber`` it works on a copy."""
    p, m = _stack_edge(cas_images, "cas_images")
    if not 0 < n <= m:
        raise ValueError(f"image size n must be in [1, {m}], got {n}")
    if out is not None:
        _need_numel(out, p * n * n, "out")
    _same_device(("cas_images", cas_images), ("out", out))
    _check(cas_images, "cas_images")
    src = cas_images if clobber else cas_images.clone()
    if out is None:
        out = torch.empty((p, n, n), device=cas_images.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_cas_images_to_images", _ptr(src), _ptr(out), p, m, n, _stream())
    return out


def images_to_cas_images(images: torch.Tensor, m: int, out: torch.Tensor | None = None) -> torch.Tensor:
    p, n = _stack_edge(images, "images")
    if m < n:
        raise ValueError(f"CAS image size m must be >= {n}, got {m}")
    if out is not None:
        _need_numel(out, p * m * m, "out")
    _same_device(("images", images), ("out", out))
    _check(images, "images")
    if out is None:
        out = torch.empty((p, m, m), device=images.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_images_to_cas_images", _ptr(images), _ptr(out), p, m, n, _stream())
    return out


def volume_to_cas_volume(volume: torch.Tensor, interp_factor: float = 2.0, extra_padding: int = 3,
                         hwidth: float = 2.0, out: torch.Tensor | None = None) -> torch.Tensor:
    n = _cube_edge(volume, "volume")
    m = int(np.float32(n) * np.float32(interp_factor))
    v = m + 2 * extra_padding
    if out is not None:
        _need_numel(out, v * v * v, "out")
    _same_device(("volume", volume), ("out", out))
    _check(volume, "volume")
    if out is None:
        out = torch.empty((v, v, v), device=volume.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_volume_to_cas_volume", _ptr(volume), n, _ptr(out), m, extra_padding, float(hwidth),
              _stream())
    return out


def cas_volume_to_volume(cas_volume: torch.Tensor, n: int, interp_factor: float = 2.0,
                         extra_padding: int = 3, hwidth: float = 2.0, zero_nyquist: bool = True,
                         out: torch.Tensor | None = None) -> torch.Tensor:
    m = int(np.float32(n) * np.float32(interp_factor))
    _need_numel(cas_volume, (m + 2 * extra_padding) ** 3, "cas_volume")
    if out is not None:
        _need_numel(out, n * n * n, "out")
    _same_device(("cas_volume", cas_volume), ("out", out))
    _check(cas_volume, "cas_volume")
    if out is None:
        out = torch.empty((n, n, n), device=cas_volume.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_cas_volume_to_volume", _ptr(cas_volume), m, extra_padding, _ptr(out), n,
              float(hwidth), int(bool(zero_nyquist)), _stream())
    return out


def cas_images_to_images_transpose(images: torch.Tensor, m: int,
                                   out: torch.Tensor | None = None) -> torch.Tensor:
    """Transpose of :func:`cas_images_to_images`: (P, n, n) -> (P, m, m),
    (1/m^2) ImgsToCAS with the index-0 row / column of each CAS image zeroed."""
    p, n = _stack_edge(images, "images")
    if m < n:
        raise ValueError(f"CAS image size m must be >= {n}, got {m}")
    if out is not None:
        _need_numel(out, p * m * m, "out")
    _same_device(("images", images), ("out", out))
    _check(images, "images")
    if out is None:
        out = torch.empty((p, m, m), device=images.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_cas_images_to_images_transpose", _ptr(images), _ptr(out), p, m, n, _stream())
    return out


def volume_to_cas_volume_transpose(cas_volume: torch.Tensor, n: int, interp_factor: float = 2.0,
                                   extra_padding: int = 3, hwidth: float = 2.0,
                                   out: torch.Tensor | None = None) -> torch.Tensor:
    """Transpose of :func:`volume_to_cas_volume`: (V^3) -> (n^3), m^3 CASToVol
    keeping the index-0 planes."""
    m = int(np.float32(n) * np.float32(interp_factor))
    _need_numel(cas_volume, (m + 2 * extra_padding) ** 3, "cas_volume")
    if out is not None:
        _need_numel(out, n * n * n, "out")
    _same_device(("cas_volume", cas_volume), ("out", out))
    _check(cas_volume, "cas_volume")
    if out is None:
        out = torch.empty((n, n, n), device=cas_volume.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_volume_to_cas_volume_transpose", _ptr(cas_volume), m, extra_padding, _ptr(out), n,
              float(hwidth), _stream())
    return out


def divide_density(num: torch.Tensor, den: torch.Tensor, eps: float = 1e-6,
                   out: torch.Tensor | None = None) -> torch.Tensor:
    _need_numel(den, num.numel(), "den")
    if out is not None:
        _need_numel(out, num.numel(), "out")
    _same_device(("num", num), ("den", den), ("out", out))
    _check(num, "num")
    _check(den, "den")
    out = torch.empty_like(num) if out is None else out
    _check(out, "out")
    _lib.call("mgg_op_divide_density", _ptr(out), _ptr(num), _ptr(den), num.numel(), float(eps), _stream())
    return out


def reduce_partials(parts: torch.Tensor, out: torch.Tensor | None = None, accumulate: bool = False):
    """Sum (K, ...) partial volumes along dim 0 in FP64 -> FP32."""
    if parts.dim() < 2 or parts.shape[0] == 0:
        raise ValueError(f"parts must be (K, ...) with K >= 1, got shape {tuple(parts.shape)}")
    k = parts.shape[0]
    n = parts[0].numel()
    if out is not None:
        _need_numel(out, n, "out")
    _same_device(("parts", parts), ("out", out))
    _check(parts, "parts")
    if out is None:
        out = torch.empty(parts.shape[1:], device=parts.device, dtype=torch.float32)
    _check(out, "out")
    _lib.call("mgg_op_reduce_partials", _ptr(out), _ptr(parts), k, n, n, int(accumulate), _stream())
    return out


def convert_volume(cas_volume: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    """FP32 CAS volume -> BF16 / FP16 storage copy (round to nearest even)."""
    _check(cas_volume, "cas_volume")
    if dtype not in (torch.bfloat16, torch.float16):
        raise ValueError("dtype must be bfloat16 or float16")
    out = torch.empty(cas_volume.shape, device=cas_volume.device, dtype=dtype)
    _lib.call("mgg_op_convert_volume", _ptr(out), _DTYPES[dtype], _ptr(cas_volume), cas_volume.numel(), _stream())
    return out


# -- standalone filters (Filters/*.cu semantics) ----------------------------------
def filter_call(name: str, *args):
    """Call an mgg_filter_* entry point; tensors are passed as device pointers."""
    conv = []
    for a in args:
        if isinstance(a, torch.Tensor):
            if not a.is_cuda or not a.is_contiguous():
                raise ValueError("filter tensors must be contiguous device tensors")
            conv.append(_ptr(a))
        else:
            conv.append(a)
    _lib.call(name, *conv, _stream())
