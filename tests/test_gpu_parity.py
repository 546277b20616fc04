"""GPU parity: every HIP kernel, through the C ABI, against the CPU oracle on
the same seeded inputs.

Tolerances (relative L2 against the FP64 oracle):
  FP32 storage, projectors and chains ......... 1e-5 (observed ~1e-7)
  BF16 CAS-volume storage (FP32 gather) ........ 2e-3
  FP16 CAS-volume storage (FP32 gather) ........ 5e-4
"""
import ctypes

import numpy as np
import pytest

from conftest import assert_parity, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _np(t):
    return t.detach().cpu().numpy()


def _setup(n, p, seed, oracle, grid=False):
    from multigpugridder_amd.synth import blob_volume, random_directions

    m = 2 * n
    vol = blob_volume(n, seed)
    cas = oracle.vol_to_cas(vol, 2, 3).astype(np.float32)
    axes = oracle.uniform_axes(*grid) if grid else random_directions(p, 100 + seed)
    return vol, cas, axes, m, m + 6


# ---------------------------------------------------------------------------
# Forward projection
# ---------------------------------------------------------------------------
# M = 60 and 100 are not multiples of the 8x8 tile: the last tile row and
# column hang over the image edge.
@pytest.mark.parametrize("n,p,seed,grid", [(8, 3, 1, False), (16, 8, 1, False), (32, 40, 2, False),
                                           (32, 0, 3, (4, 10)), (64, 64, 4, False), (30, 12, 5, False),
                                           (50, 20, 6, False)])
def test_forward_project_f32(cuda, oracle, ker32, n, p, seed, grid):
    from multigpugridder_amd import ops

    print(f"FP parity n={n} p={p} seed={seed} grid={grid}", flush=True)
    vol, cas, axes, m, v = _setup(n, p, seed, oracle, grid)
    got = _np(ops.forward_project(_t(cas, cuda), _t(axes, cuda), m, ker=_t(ker32, cuda)))
    ref = oracle.forward_project(cas, axes, ker32, m)
    assert_parity(got, ref)
    # out-of-mask pixels are exactly zero, as in the reference
    jj, ii = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    outside = (ii - m // 2) ** 2 + (jj - m // 2) ** 2 > (m // 2 - 1) ** 2
    assert np.all(got[:, outside] == 0)


def test_forward_project_non_finite_voxels(cuda, oracle, ker32):
    """A NaN and an Inf voxel reach exactly the pixels whose 5x5x5 stencil
    holds them (NaN where the Inf meets a zero table weight), as in the
    reference; every other pixel matches the oracle."""
    from multigpugridder_amd import ops

    n, p = 32, 24
    print(f"FP non-finite voxels n={n} p={p}", flush=True)
    vol, cas, axes, m, v = _setup(n, p, 5, oracle)
    c = v // 2
    cas[c + 1, c - 2, c + 3] = np.nan
    cas[c - 9, c + 7, c - 4] = np.inf
    got = _np(ops.forward_project(_t(cas, cuda), _t(axes, cuda), m, ker=_t(ker32, cuda)))
    ref = oracle.forward_project(cas, axes, ker32, m)
    assert np.isnan(ref).any() and np.isinf(ref).any()
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.array_equal(np.isinf(got), np.isinf(ref))
    fin = np.isfinite(ref)
    assert rel_l2(got[fin], ref[fin]) < TOL


def test_check_volume_flag(cuda, oracle, ker32):
    """ops.check_volume writes 1 only for a finite volume with the 501-entry
    table whose end entries are 0; each way of breaking that gives 0."""
    import torch

    from multigpugridder_amd import ops

    vol, cas, axes, m, v = _setup(16, 4, 3, oracle)
    k = _t(ker32, cuda)

    def flag(c, kk=k):
        return int(ops.check_volume(c, kk).item())

    base = _t(cas, cuda)
    assert flag(base) == 1
    for bad in (np.nan, np.inf, -np.inf):
        for at in ((0, 0, 0), (v - 1, v - 1, v - 1), (v // 2, 3, v - 2)):
            c2 = cas.copy()
            c2[at] = bad
            assert flag(_t(c2, cuda)) == 0, (bad, at)
    for dt in (torch.bfloat16, torch.float16):
        d = base.to(dt)
        assert flag(d) == 1
        for bad in (float("inf"), float("nan")):
            d2 = d.clone()
            d2[3, 4, 5] = bad
            assert flag(d2) == 0, (dt, bad)
    for idx, val in ((0, 1e-7), (-1, 1e-7), (250, np.inf), (17, np.nan)):
        kb = ker32.copy()
        kb[idx] = val
        assert flag(base, _t(kb, cuda)) == 0, (idx, val)
    assert flag(base, _t(oracle.kb_table(2, 301).astype(np.float32), cuda)) == 0
    # a volume that does not start on a 16-byte boundary (the scalar loop)
    flat = torch.zeros(v ** 3 + 1, device=cuda)
    flat[1:] = base.reshape(-1)
    assert flag(flat[1:].view(v, v, v)) == 1
    flat[1 + v ** 3 // 2] = float("nan")
    assert flag(flat[1:].view(v, v, v)) == 0


@pytest.mark.parametrize("n,p,grid", [(128, 16, False), (128, 0, (5, 6)), (64, 48, False), (30, 12, False)])
def test_forward_project_window_shortcut(cuda, oracle, ker32, n, p, grid):
    """With the check flag set the gather reads 4x4x4 windows from smaller
    boxes: it matches the full-stencil gather (only the order in which split
    tiles add their two halves can differ) and the oracle."""
    from multigpugridder_amd import ops

    print(f"FP window shortcut n={n} p={p} grid={grid}", flush=True)
    vol, cas, axes, m, v = _setup(n, p, 7, oracle, grid)
    c, a, k = _t(cas, cuda), _t(axes, cuda), _t(ker32, cuda)
    flag = ops.check_volume(c, k)
    assert int(flag.item()) == 1
    short = _np(ops.forward_project(c, a, m, ker=k, check=flag))
    full = _np(ops.forward_project(c, a, m, ker=k, check=False))
    assert_parity(short, full, tol=1e-6, peak=1e-6)
    assert_parity(short, oracle.forward_project(cas, axes, ker32, m))


def test_forward_project_golden(cuda):
    import os

    from multigpugridder_amd import ops

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "projector_n16.npz"))
    got = _np(ops.forward_project(_t(g["cas_vol"], cuda), _t(g["axes"], cuda), int(g["m"]), ker=_t(g["ker"], cuda)))
    assert_parity(got, g["fp"])


def test_forward_project_large_mask_takes_checked_path(cuda, oracle, ker32):
    """A mask radius beyond the default pushes stencils past the volume edge:
    the kernel's checked path must match the reference's linear-index test."""
    from multigpugridder_amd import ops

    vol, cas, axes, m, v = _setup(16, 6, 5, oracle)
    r = m / 2 + 3.5
    got = _np(ops.forward_project(_t(cas, cuda), _t(axes, cuda), m, mask_radius=r, ker=_t(ker32, cuda)))
    ref = oracle.forward_project(cas, axes, ker32, m, mask_radius=r)
    assert_parity(got, ref)


@pytest.mark.parametrize("dtype,tol", [("bfloat16", 2e-3), ("float16", 5e-4)])
def test_forward_project_reduced_storage(cuda, oracle, ker32, dtype, tol):
    import torch

    from multigpugridder_amd import ops

    vol, cas, axes, m, v = _setup(32, 24, 6, oracle)
    dv = ops.convert_volume(_t(cas, cuda), getattr(torch, dtype))
    # the conversion itself is round-to-nearest-even, as torch's
    assert torch.equal(dv, _t(cas, cuda).to(getattr(torch, dtype)))
    got = _np(ops.forward_project(dv, _t(axes, cuda), m, ker=_t(ker32, cuda)))
    ref = oracle.forward_project(cas, axes, ker32, m)
    err = rel_l2(got, ref)
    print(f"{dtype} storage: rel L2 {err:.3e}", flush=True)
    assert err < tol
    # exact against the oracle run on the rounded volume
    rounded = _np(dv.float())
    assert rel_l2(got, oracle.forward_project(rounded, axes, ker32, m)) < TOL


# ---------------------------------------------------------------------------
# Back projection (reference semantics), density, exact transpose
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,p,seed", [(8, 3, 1), (16, 8, 2), (32, 37, 3), (48, 20, 4)])
def test_back_project_reference(cuda, oracle, ker32, n, p, seed):
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    print(f"BP parity n={n} p={p}", flush=True)
    m, v = 2 * n, 2 * n + 6
    axes = random_directions(p, seed)
    y = np.random.default_rng(seed).standard_normal((p, m, m)).astype(np.float32)
    out = torch.zeros((v, v, v), device=cuda)
    dens = torch.zeros((v, v, v), device=cuda)
    ops.back_project(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker32, cuda), out=out, density=dens)
    assert_parity(_np(out), oracle.back_project(y, axes, ker32, v))
    ref_d = oracle.back_project(np.ones_like(y), axes, ker32, v)
    assert_parity(_np(dens), ref_d)


@pytest.mark.parametrize("scale", [1.05, 0.9])
def test_back_project_non_orthonormal_axes(cuda, oracle, ker32, scale):
    """Direction frames that are not orthonormal (scaled rotations, and one
    chunk mixing them with exact rotations) take the clamped table lookup;
    results still follow the reference arithmetic."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    n, p = 24, 150
    m, v = 2 * n, 2 * n + 6
    axes = random_directions(p, 31)
    axes[::3] *= np.float32(scale)
    print(f"BP non-orthonormal scale={scale}", flush=True)
    y = np.random.default_rng(5).standard_normal((p, m, m)).astype(np.float32)
    out = torch.zeros((v, v, v), device=cuda)
    ops.back_project(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker32, cuda), out=out)
    assert_parity(_np(out), oracle.back_project(y, axes, ker32, v))
    img = ops.forward_project(_t(oracle.vol_to_cas(np.ones((n, n, n)), 2, 3).astype(np.float32), cuda),
                              _t(axes, cuda), m, ker=_t(ker32, cuda))
    cas = oracle.vol_to_cas(np.ones((n, n, n)), 2, 3).astype(np.float32)
    assert_parity(_np(img), oracle.forward_project(cas, axes, ker32, m))


def test_zero_directions_are_no_ops(cuda):
    import torch

    from multigpugridder_amd import ops

    v, m = 38, 32
    cas = torch.ones((v, v, v), device=cuda)
    axes = torch.zeros((0, 9), device=cuda)
    imgs = ops.forward_project(cas, axes, m)
    assert tuple(imgs.shape) == (0, m, m)
    out = torch.full((v, v, v), 3.0, device=cuda)
    ops.back_project(torch.zeros((0, m, m), device=cuda), axes, v, out=out)
    ops.back_project_adjoint(torch.zeros((0, m, m), device=cuda), axes, v, out=out)
    torch.cuda.synchronize()
    assert bool((out == 3.0).all())


def test_back_project_repeatable(cuda, oracle, ker32):
    """Two identical back projections give bit-identical volumes."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    m, v, p = 64, 70, 300
    axes = _t(random_directions(p, 8), cuda)
    y = _t(np.random.default_rng(8).standard_normal((p, m, m)).astype(np.float32), cuda)
    a = ops.back_project(y, axes, v, ker=_t(ker32, cuda))
    b = ops.back_project(y, axes, v, ker=_t(ker32, cuda))
    assert torch.equal(a, b)


@pytest.mark.parametrize("ksize", [301, 1001])
def test_back_project_generic_table(cuda, oracle, ksize):
    """Tables other than 501 entries take the generic (voxel-per-lane) kernel."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    ker = oracle.kb_table(2, ksize).astype(np.float32)
    m, v, p = 32, 38, 9
    axes = random_directions(p, 4)
    y = np.random.default_rng(5).standard_normal((p, m, m)).astype(np.float32)
    out = torch.zeros((v, v, v), device=cuda)
    ops.back_project(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker, cuda), out=out)
    assert_parity(_np(out), oracle.back_project(y, axes, ker, v))
    got = _np(ops.forward_project(out, _t(axes, cuda), m, ker=_t(ker, cuda)))
    assert_parity(got, oracle.forward_project(_np(out), axes, ker, m))


@pytest.mark.parametrize("dtype", ["float32", "bfloat16"])
def test_forward_project_large_table(cuda, oracle, dtype):
    """A 1501-entry table does not fit the staged kernel's LDS table: the
    direct-gather kernel with 5-tap row loads runs instead."""
    import torch

    from multigpugridder_amd import ops

    ker = oracle.kb_table(2, 1501).astype(np.float32)
    vol, cas, axes, m, v = _setup(16, 7, 9, oracle)
    dv = _t(cas, cuda).to(getattr(torch, dtype))
    got = _np(ops.forward_project(dv, _t(axes, cuda), m, ker=_t(ker, cuda)))
    assert_parity(got, oracle.forward_project(_np(dv.float()), axes, ker, m))


def test_back_project_accumulates_and_wraps(cuda, oracle, ker32):
    """Two launches accumulate; the linear-index wrap quirk is reproduced."""
    import torch

    from multigpugridder_amd import ops

    m, v = 32, 38
    t = np.deg2rad(25.0)
    u = np.array([np.cos(t), np.sin(t), 0.0])
    w = np.array([-np.sin(t), np.cos(t), 0.0])
    axes = np.concatenate([u, w, np.cross(u, w)]).astype(np.float32)[None]
    y = np.zeros((1, m, m), np.float32)
    y[0, :, m - 1] = 1.0
    out = torch.zeros((v, v, v), device=cuda)
    ops.back_project(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker32, cuda), out=out)
    ops.back_project(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker32, cuda), out=out)
    ref = 2 * oracle.back_project(y, axes, ker32, v)
    assert_parity(_np(out), ref)
    assert np.count_nonzero(ref) > 0


@pytest.mark.parametrize("n,p,scale", [(16, 6, 1.0), (32, 20, 1.0), (48, 40, 1.0), (24, 30, 0.7), (24, 30, 1.3)])
def test_back_project_adjoint(cuda, oracle, ker32, n, p, scale):
    """The exact transpose against the oracle's.  n=48 splits most 8x8 tiles
    into 8x4 halves (their boxes exceed the LDS budget); scaled frames (not
    orthonormal: many pixels share a stencil centre) take the LDS-atomic
    branch."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    print(f"adjoint n={n} p={p} scale={scale}", flush=True)
    m, v = 2 * n, 2 * n + 6
    axes = random_directions(p, 9)
    axes[1::2] *= np.float32(scale)
    y = np.random.default_rng(3).standard_normal((p, m, m)).astype(np.float32)
    out = torch.zeros((v, v, v), device=cuda)
    ops.back_project_adjoint(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker32, cuda), out=out)
    assert_parity(_np(out), oracle.back_project_adjoint(y, axes, ker32, v))


@pytest.mark.parametrize("n,p,scale", [(64, 24, 1.0), (32, 20, 1.3)])
def test_back_project_adjoint_window_shortcut(cuda, oracle, ker32, n, p, scale):
    """With the image check flag set the transpose scatters 4x4x4 windows:
    the same volume as the full-stencil scatter (up to the order of the
    global atomic adds), and a non-finite image value clears the flag."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    print(f"adjoint window shortcut n={n} p={p} scale={scale}", flush=True)
    m, v = 2 * n, 2 * n + 6
    axes = random_directions(p, 19)
    axes[1::2] *= np.float32(scale)
    y = np.random.default_rng(4).standard_normal((p, m, m)).astype(np.float32)
    yt, at, k = _t(y, cuda), _t(axes, cuda), _t(ker32, cuda)
    flag = ops.check_images(yt, k)
    assert int(flag.item()) == 1
    short = torch.zeros((v, v, v), device=cuda)
    full = torch.zeros((v, v, v), device=cuda)
    ops.back_project_adjoint(yt, at, v, ker=k, out=short, check=flag)
    ops.back_project_adjoint(yt, at, v, ker=k, out=full, check=False)
    assert_parity(_np(short), _np(full), tol=1e-6, peak=1e-6)
    assert_parity(_np(short), oracle.back_project_adjoint(y, axes, ker32, v))
    y2 = y.copy()
    y2[p // 2, m // 2, m // 2 + 3] = np.inf
    assert int(ops.check_images(_t(y2, cuda), k).item()) == 0


def test_kernels_follow_the_contracted_reading(cuda, oracle, ker32):
    """The kernels take their rounding decisions as the reference compiled by
    nvcc with its default --fmad=true does (the reading the oracle pins,
    "first").  Readings that differ only in how multiply-adds are contracted
    give results the parity bar tells apart: the back projection of noise
    images moves by > 1e-5 under the unfused source order, and the forward
    projection by > 2e-6 of the peak under nested-FMA reassociation."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 32, 60
    m, v = 2 * n, 2 * n + 6
    print(f"reading discrimination n={n} p={p}", flush=True)
    cas = oracle.vol_to_cas(blob_volume(n, 1), 2, 3).astype(np.float32)
    axes = random_directions(p, 7)
    y = np.random.default_rng(3).standard_normal((p, m, m)).astype(np.float32)
    fp = _np(ops.forward_project(_t(cas, cuda), _t(axes, cuda), m, ker=_t(ker32, cuda)))
    out = torch.zeros((v, v, v), device=cuda)
    ops.back_project(_t(y, cuda), _t(axes, cuda), v, ker=_t(ker32, cuda), out=out)
    bp = _np(out)
    assert_parity(fp, oracle.forward_project(cas, axes, ker32, m))
    assert_parity(bp, oracle.back_project(y, axes, ker32, v))
    with oracle.reading("source"):
        bp_src = oracle.back_project(y, axes, ker32, v)
    with oracle.reading("reassoc"):
        fp_re = oracle.forward_project(cas, axes, ker32, m)
    print(f"  BP vs source reading {rel_l2(bp, bp_src):.2e}; FP vs reassoc max/peak "
          f"{np.abs(fp - fp_re).max() / np.abs(fp_re).max():.2e}", flush=True)
    assert rel_l2(bp, bp_src) > 1e-5
    assert np.abs(fp - fp_re).max() > 2e-6 * np.abs(fp_re).max()


def test_fractional_mask_radius_is_truncated(cuda, oracle, ker32):
    """RunKernel takes the mask radius as an int (gpuForwardProject.h:57,
    gpuBackProject.h:56): a radius of 19.5 behaves exactly as 19."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 24, 12
    m, v = 2 * n, 2 * n + 6
    print("fractional mask radius 19.5", flush=True)
    cas = oracle.vol_to_cas(blob_volume(n, 2), 2, 3).astype(np.float32)
    axes = random_directions(p, 12)
    y = np.random.default_rng(12).standard_normal((p, m, m)).astype(np.float32)
    dc, da, dy, dk = _t(cas, cuda), _t(axes, cuda), _t(y, cuda), _t(ker32, cuda)
    for r in (19.5, 19.0):
        fp = _np(ops.forward_project(dc, da, m, mask_radius=r, ker=dk))
        assert_parity(fp, oracle.forward_project(cas, axes, ker32, m, mask_radius=19.0))
        out = torch.zeros((v, v, v), device=cuda)
        ops.back_project(dy, da, v, mask_radius=r, ker=dk, out=out)
        assert_parity(_np(out), oracle.back_project(y, axes, ker32, v, mask_radius=19.0))
        adj = torch.zeros((v, v, v), device=cuda)
        ops.back_project_adjoint(dy, da, v, mask_radius=r, ker=dk, out=adj)
        assert_parity(_np(adj), oracle.back_project_adjoint(y, axes, ker32, v, mask_radius=19.0))
    f195 = ops.forward_project(dc, da, m, mask_radius=19.5, ker=dk)
    assert torch.equal(f195, ops.forward_project(dc, da, m, mask_radius=19.0, ker=dk))
    assert not torch.equal(f195, ops.forward_project(dc, da, m, mask_radius=20.0, ker=dk))


# ---------------------------------------------------------------------------
# Image and volume chains (rocFFT + fused pre/post kernels)
# ---------------------------------------------------------------------------
# m = 128, 192, 256, 384 and 512 run the fused two-pass chains (any n <= m with
# m - n even, an odd offset (m - n) / 2 included); other m run the rocFFT chains
# (300 images: full batches of 256 and a remainder split into power-of-two batches).
@pytest.mark.parametrize("n,m,p", [(8, 16, 3), (16, 32, 7), (100, 200, 3), (50, 100, 300),
                                   (64, 128, 5), (2, 128, 3), (30, 128, 7), (128, 128, 2), (100, 128, 260),
                                   (96, 192, 5), (2, 192, 3), (50, 192, 4), (192, 192, 2), (100, 192, 260),
                                   (128, 256, 300), (100, 256, 5), (6, 256, 3), (254, 256, 2), (256, 256, 2),
                                   (192, 384, 3), (98, 384, 2), (384, 384, 2),
                                   (256, 512, 3), (100, 512, 2)])
def test_image_chains(cuda, oracle, n, m, p):
    from multigpugridder_amd import ops

    print(f"image chains n={n} m={m} p={p}", flush=True)
    rng = np.random.default_rng(n)
    c = rng.standard_normal((p, m, m)).astype(np.float32)
    got = _np(ops.cas_images_to_images(_t(c, cuda), n))
    assert rel_l2(got, oracle.cas_imgs_to_imgs(c, n)) < TOL
    x = rng.standard_normal((p, n, n)).astype(np.float32)
    got = _np(ops.images_to_cas_images(_t(x, cuda), m))
    assert rel_l2(got, oracle.imgs_to_cas_imgs(x, m)) < TOL


@pytest.mark.parametrize("n,m,p", [(128, 256, 40), (98, 256, 3), (256, 512, 3), (64, 128, 9), (98, 128, 3),
                                   (96, 192, 7), (192, 384, 3), (24, 48, 5)])
def test_image_chain_transposes(cuda, oracle, n, m, p):
    """The differentiable projector's transposes of the image chains (scale and
    index-0 zeroing), on the fused (m = 128, 192, 256, 384, 512) and the rocFFT
    chains."""
    from multigpugridder_amd import ops

    print(f"image chain transposes n={n} m={m} p={p}", flush=True)
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal((p, n, n)).astype(np.float32)
    got = _np(ops.cas_images_to_images_transpose(_t(x, cuda), m))
    assert rel_l2(got, oracle.cas_imgs_to_imgs_transpose(x, m)) < TOL
    assert np.all(got[:, 0, :] == 0) and np.all(got[:, :, 0] == 0)


@pytest.mark.parametrize("n", [8, 16, 30, 64])
def test_volume_chains(cuda, oracle, n):
    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import blob_volume

    vol = blob_volume(n, 3)
    got = _np(ops.volume_to_cas_volume(_t(vol, cuda)))
    ref = oracle.vol_to_cas(vol, 2, 3)
    assert rel_l2(got, ref) < TOL
    back = _np(ops.cas_volume_to_volume(_t(ref.astype(np.float32), cuda), n))
    assert rel_l2(back, oracle.cas_to_vol(ref, n)) < TOL
    back2 = _np(ops.cas_volume_to_volume(_t(ref.astype(np.float32), cuda), n, zero_nyquist=False))
    assert rel_l2(back2, oracle.cas_to_vol(ref, n, zero_nyquist=False)) < TOL


def test_reduce_and_divide(cuda):
    import torch

    from multigpugridder_amd import ops

    rng = np.random.default_rng(0)
    parts = rng.standard_normal((5, 1000)).astype(np.float32)
    got = _np(ops.reduce_partials(_t(parts, cuda)))
    np.testing.assert_allclose(got, parts.astype(np.float64).sum(0).astype(np.float32), rtol=0, atol=0)
    num = rng.standard_normal(1000).astype(np.float32)
    den = rng.random(1000).astype(np.float32)
    got = _np(ops.divide_density(_t(num, cuda), _t(den, cuda)))
    ref = (num.astype(np.float64) / (den.astype(np.float64) + 1e-6)).astype(np.float32)
    np.testing.assert_array_equal(got, ref)
    assert torch.cuda.is_available()


# ---------------------------------------------------------------------------
# Standalone filters: bit-exact against the reference's Matlab KATs
# (Tests/*Test.m), restated with numpy.
# ---------------------------------------------------------------------------
def test_filters_bit_exact(cuda, oracle):
    import torch

    from multigpugridder_amd.ops import filter_call

    rng = np.random.default_rng(1)
    s, k = 16, 5
    a = rng.standard_normal((k, s, s)).astype(np.float32)
    b = (rng.random((k, s, s)).astype(np.float32) + 1e-2)
    ta, tb = _t(a, cuda), _t(b, cuda)

    out = torch.empty((k, s, s, 2), device=cuda)
    filter_call("mgg_filter_cas_to_complex", ta, out, s, k)
    z = oracle.from_cas(a, (1, 2))
    np.testing.assert_array_equal(_np(out)[..., 0], z.real.astype(np.float32))
    np.testing.assert_array_equal(_np(out)[..., 1], z.imag.astype(np.float32))

    vol = rng.standard_normal((s, s, s)).astype(np.float32)
    outv = torch.empty((s, s, s, 2), device=cuda)
    filter_call("mgg_filter_cas_to_complex", _t(vol, cuda), outv, s, 0)
    zv = oracle.from_cas(vol, (0, 1, 2))
    np.testing.assert_array_equal(_np(outv)[..., 0], zv.real.astype(np.float32))

    cplx = rng.standard_normal((k, s, s, 2)).astype(np.float32)
    r = torch.empty((k, s, s), device=cuda)
    filter_call("mgg_filter_complex_to_cas", _t(cplx, cuda), r, s, k)
    np.testing.assert_array_equal(_np(r), cplx[..., 0] + cplx[..., 1])
    filter_call("mgg_filter_complex_to_real", _t(cplx, cuda), r, s, k)
    np.testing.assert_array_equal(_np(r), cplx[..., 0])
    c2 = torch.empty((k, s, s, 2), device=cuda)
    filter_call("mgg_filter_real_to_complex", ta, c2, s, k)
    np.testing.assert_array_equal(_np(c2)[..., 0], a)
    assert not _np(c2)[..., 1].any()

    t2 = _t(a, cuda)
    filter_call("mgg_filter_fftshift2d", t2, 4, s, k)
    np.testing.assert_array_equal(_np(t2), oracle.fftshift2d(a))
    tc = _t(cplx, cuda)
    filter_call("mgg_filter_fftshift2d", tc, 8, s, k)
    np.testing.assert_array_equal(_np(tc), np.fft.fftshift(cplx, axes=(1, 2)))
    t3 = _t(vol, cuda)
    filter_call("mgg_filter_fftshift3d", t3, 4, s)
    np.testing.assert_array_equal(_np(t3), oracle.fftshift3d(vol))

    small = rng.standard_normal((k, 6, 6)).astype(np.float32)
    padded = torch.full((k, s, s), 7.0, device=cuda)
    filter_call("mgg_filter_pad", _t(small, cuda), padded, 6, s, k)
    np.testing.assert_array_equal(_np(padded), oracle.pad_images(small, s))
    cropped = torch.empty((k, 6, 6), device=cuda)
    filter_call("mgg_filter_crop", ta, cropped, 4, s, 6, k)
    np.testing.assert_array_equal(_np(cropped), oracle.crop_images(a, 6))

    t4 = _t(a, cuda)
    filter_call("mgg_filter_divide_scalar", t4, ctypes.c_float(3.0), a.size)
    np.testing.assert_array_equal(_np(t4), a / np.float32(3.0))
    az = a.copy()
    az[0, 0, :4] = 0
    t5 = _t(az, cuda)
    filter_call("mgg_filter_divide_volume", t5, tb, a.size)
    np.testing.assert_array_equal(_np(t5), oracle.divide_volume(az, b))
    t6 = _t(a, cuda)
    filter_call("mgg_filter_multiply_volume", t6, tb, a.size)
    np.testing.assert_array_equal(_np(t6), a * b)
    t7 = _t(cplx, cuda)
    filter_call("mgg_filter_multiply_complex_volume", t7, tb, a.size)
    np.testing.assert_array_equal(_np(t7), cplx * b[..., None])
    t8 = _t(a, cuda)
    filter_call("mgg_filter_add_volume", t8, tb, a.size)
    np.testing.assert_array_equal(_np(t8), a + b)
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,interp,pad", [(16, 3.0, 5), (20, 2.0, 0), (24, 2.0, 1), (12, 4.0, 3)])
def test_other_interp_factors_and_paddings(cuda, oracle, ker32, n, interp, pad):
    """The whole path with an interpolation factor or CAS-volume padding other
    than the defaults (the reference's kernels take any V = M + 2 pad): Vol ->
    CAS, forward projection, CAS images -> images, images -> CAS images, back
    projection and CAS -> Vol against the oracle.  Padding 0 puts the outer
    stencil taps outside the volume, where only the linear-index bound applies."""
    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import blob_volume, random_directions

    m = int(np.float32(n) * np.float32(interp))
    v = m + 2 * pad
    print(f"n={n} interp={interp} pad={pad}: m={m} v={v}", flush=True)
    vol = blob_volume(n, 3)
    axes = random_directions(9, 41)
    dk, da = _t(ker32, cuda), _t(axes, cuda)
    cas = ops.volume_to_cas_volume(_t(vol, cuda), interp, pad)
    cas_ref = oracle.vol_to_cas(vol, interp, pad)
    assert cas.shape == (v, v, v)
    assert rel_l2(cas.cpu().numpy(), cas_ref) < 1e-5
    cas32 = cas_ref.astype(np.float32)
    ci = ops.forward_project(_t(cas32, cuda), da, m, ker=dk)
    ci_ref = oracle.forward_project(cas32, axes, ker32, m)
    assert_parity(ci.cpu().numpy(), ci_ref)
    imgs = ops.cas_images_to_images(ci, n)
    assert rel_l2(imgs.cpu().numpy(), oracle.cas_imgs_to_imgs(ci_ref, n)) < 1e-5
    ci2 = ops.images_to_cas_images(imgs, m).cpu().numpy()
    assert rel_l2(ci2, oracle.imgs_to_cas_imgs(imgs.cpu().numpy(), m)) < 1e-5
    bp = ops.back_project(_t(ci2, cuda), da, v, ker=dk)
    bp_ref = oracle.back_project(ci2, axes, ker32, v)
    assert_parity(bp.cpu().numpy(), bp_ref)
    rec = ops.cas_volume_to_volume(bp, n, interp, pad)
    assert rel_l2(rec.cpu().numpy(), oracle.cas_to_vol(bp_ref, n, interp, pad)) < 1e-5


@pytest.mark.parametrize("r,r_int", [(0.0, 0.0), (0.7, 0.0), (1.0, 1.0), (-0.5, 0.0), (-1.0, -1.0)])
def test_tiny_and_negative_mask_radii(cuda, oracle, ker32, r, r_int):
    """Mask radii at the bottom of the range, through the int truncation of
    RunKernel: 0 and 0.7 keep the centre pixel and voxel only, 1 their
    neighbours too, -0.5 truncates to 0, and -1 masks everything out."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 16, 6
    m, v = 2 * n, 2 * n + 6
    print(f"mask radius {r} (as {r_int})", flush=True)
    cas = oracle.vol_to_cas(blob_volume(n, 5), 2, 3).astype(np.float32)
    axes = random_directions(p, 14)
    y = np.random.default_rng(14).standard_normal((p, m, m)).astype(np.float32)
    dc, da, dy, dk = _t(cas, cuda), _t(axes, cuda), _t(y, cuda), _t(ker32, cuda)
    fp = _np(ops.forward_project(dc, da, m, mask_radius=r, ker=dk))
    fp_ref = oracle.forward_project(cas, axes, ker32, m, mask_radius=r_int)
    assert np.count_nonzero(fp) == np.count_nonzero(fp_ref)
    if fp_ref.any():
        assert_parity(fp, fp_ref)
    out = torch.zeros((v, v, v), device=cuda)
    ops.back_project(dy, da, v, mask_radius=r, ker=dk, out=out)
    bp_ref = oracle.back_project(y, axes, ker32, v, mask_radius=r_int)
    assert np.count_nonzero(_np(out)) == np.count_nonzero(bp_ref)
    if bp_ref.any():
        assert_parity(_np(out), bp_ref)
    if r_int < 0:
        assert not fp.any() and not _np(out).any()


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_tiny_volumes(cuda, oracle, ker32, n):
    """CAS volumes far smaller than a pixel tile or a voxel brick (N = 1..5,
    M = 2..10, V = 8..16): forward projection and back projection against the
    oracle, and the image chain where M - N is even (the chains refuse an odd
    difference)."""
    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    m, v = 2 * n, 2 * n + 6
    print(f"tiny volume n={n} m={m} v={v}", flush=True)
    # (a random CAS volume: the reference's precompensation, getPreComp.m, needs
    # lines longer than the kernel's footprint)
    cas = np.random.default_rng(n).standard_normal((v, v, v)).astype(np.float32)
    axes = random_directions(5, 15)
    dk, da = _t(ker32, cuda), _t(axes, cuda)
    ci = ops.forward_project(_t(cas, cuda), da, m, ker=dk)
    ci_ref = oracle.forward_project(cas, axes, ker32, m)
    assert_parity(_np(ci), ci_ref)
    if n % 2 == 0:
        imgs = ops.cas_images_to_images(ci, n)
        assert rel_l2(_np(imgs), oracle.cas_imgs_to_imgs(ci_ref, n)) < 1e-5
    else:
        from multigpugridder_amd import _lib
        with pytest.raises(_lib.MggError):
            ops.cas_images_to_images(ci, n)
    y = np.random.default_rng(n + 1).standard_normal((5, m, m)).astype(np.float32)
    bp = ops.back_project(_t(y, cuda), da, v, ker=dk)
    assert_parity(_np(bp), oracle.back_project(y, axes, ker32, v))
