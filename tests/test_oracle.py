"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each test checks the oracle against something it does not share code with:
closed forms, invariants, independent formulations (explicit DFT matrices,
integer-geometry gathers, analytic Gaussian projections) or exact algebraic
identities, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails.
"""
import numpy as np
import pytest
from scipy.special import i0

from conftest import rel_l2

N, M, PAD = 16, 32, 3
V = M + 2 * PAD


# ---------------------------------------------------------------------------
# Kaiser-Bessel table (Matlab/getKernelFiltTable.m) and precompensation
# ---------------------------------------------------------------------------
def test_kb_table_invariants(oracle):
    t = oracle.kb_table(2, 501)
    assert t[0] == 0.0 and t[-1] == 0.0
    np.testing.assert_array_equal(t, t[::-1])
    # the 3-D kernel at integer offsets sums to 1 (getKernelFiltTable.m:15-22)
    assert abs(t[125] + t[250] + t[375] - 1.0) < 1e-14
    assert np.all(t[1:-1] > 0) and np.argmax(t) == 250


def test_kb_table_closed_form_shape(oracle):
    # Offset- and scale-free ratios against I0 directly.
    t = oracle.kb_table(2, 501)
    x = (np.arange(501) - 250) * 0.008
    w = i0(10.0 * np.sqrt(np.clip(1 - (x / 2) ** 2, 0, None)))
    for k in (1, 60, 125, 200, 249):
        r_or = (t[k] - t[0]) / (t[250] - t[0])
        r_cf = (w[k] - w[0]) / (w[250] - w[0])
        assert abs(r_or - r_cf) < 1e-12


def test_precomp_closed_form(oracle):
    # 1/precomp at frequency k is the window's cosine sum (independent of the
    # oracle's fftshift/fft route), normalised to 1 at the centre.
    for m in (16, 32, 64):
        p = oracle.precomp(m)
        xs = np.arange(-8, 9) / 4
        w = i0(10.0 * np.sqrt(np.clip(1 - (xs / 2) ** 2, 0, None))) / i0(10.0)
        w = w - w[0]
        w[-1] = 0
        ks = np.arange(m) - m // 2
        ft = np.array([np.sum(w * np.cos(2 * np.pi * k * np.arange(-8, 9) / (4 * m))) for k in ks])
        np.testing.assert_allclose(p, ft[m // 2] / ft, rtol=1e-12)
        assert p[m // 2] == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(p[1:], p[1:][::-1], rtol=1e-12)


# ---------------------------------------------------------------------------
# Forward projection (gpuForwardProject.cu)
# ---------------------------------------------------------------------------
def _signed_perm_axes():
    e = np.eye(3)
    out = []
    for (p, q), (su, sv) in [((0, 1), (1, 1)), ((1, 2), (1, 1)), ((2, 0), (1, -1)), ((1, 0), (-1, 1)),
                             ((0, 2), (-1, -1))]:
        u, v = su * e[p], sv * e[q]
        out.append(np.concatenate([u, v, np.cross(u, v)]))
    return np.asarray(out, dtype=np.float32)


def _integer_gather(vol, axes, ker, m):
    """Independent gather for directions whose u, v are signed unit axes: every
    pixel lands on a voxel centre, so the taps sit at integer offsets and use
    table entries 0, 125, 250, 375, 500."""
    v = vol.shape[0]
    vol = vol.astype(np.float64)
    cv, ci = v // 2, m // 2
    k5 = ker[[0, 125, 250, 375, 500]]
    wij = (k5[:, None] * k5[None, :]).astype(np.float32)
    w = (wij[:, :, None] * k5[None, None, :]).astype(np.float32)  # (k[a] * k[b]) * k[c] in FP32
    out = np.zeros((axes.shape[0], m, m))
    jj, ii = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    inside = (ii - ci) ** 2 + (jj - ci) ** 2 <= (m // 2 - 1) ** 2
    for a, ax in enumerate(axes):
        u, vv = ax[:3].astype(int), ax[3:6].astype(int)
        pos = cv + (ii - ci)[..., None] * u + (jj - ci)[..., None] * vv  # (m, m, 3) xyz
        acc = np.zeros((m, m))
        for da in range(5):
            for db in range(5):
                for dc in range(5):
                    x = pos[..., 0] + da - 2
                    y = pos[..., 1] + db - 2
                    z = pos[..., 2] + dc - 2
                    acc += np.float64(w[da, db, dc]) * vol[z, y, x]
        out[a] = np.where(inside, acc, 0.0)
    return out


def test_fp_constant_volume_is_one(oracle, ker32):
    vol = np.ones((V, V, V), np.float32)
    axes = np.concatenate([_signed_perm_axes(), np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1]], np.float32)])
    img = oracle.forward_project(vol, axes, ker32, M)
    jj, ii = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    inside = (ii - M // 2) ** 2 + (jj - M // 2) ** 2 <= (M // 2 - 1) ** 2
    assert np.all(np.abs(img[:, inside] - 1.0) < 1e-6)
    assert np.all(img[:, ~inside] == 0.0)


def test_fp_matches_integer_geometry_gather(oracle, ker32):
    rng = np.random.default_rng(11)
    vol = rng.standard_normal((V, V, V)).astype(np.float32)
    axes = _signed_perm_axes()
    got = oracle.forward_project(vol, axes, ker32, M)
    want = _integer_gather(vol, axes, ker32, M)
    assert rel_l2(got, want) < 1e-14


def _contained_blobs(n, seed, k=6):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0.35 * n, 0.65 * n, (k, 3))
    sig = rng.uniform(1.5, 2.5, k)
    amp = rng.uniform(0.5, 1.0, k)
    x = np.arange(n, dtype=float)
    vol = np.zeros((n, n, n))
    for (cx, cy, cz), s, a in zip(cen, sig, amp):
        g = lambda c: np.exp(-(x - c) ** 2 / (2 * s * s))  # noqa: E731
        vol += a * g(cz)[:, None, None] * g(cy)[None, :, None] * g(cx)[None, None, :]
    return vol, cen, sig, amp


def _analytic_projection(n, cen, sig, amp, u, v):
    c0 = n / 2
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    img = np.zeros((n, n))
    for c, s, a in zip(cen, sig, amp):
        d = np.asarray(c) - c0
        img += a * s * np.sqrt(2 * np.pi) * np.exp(-((ii - c0 - d @ u) ** 2 + (jj - c0 - d @ v) ** 2) / (2 * s * s))
    return img


def test_pipeline_matches_analytic_projection(oracle, ker32):
    """Fourier-slice theorem end to end: Vol->CAS, FP gather, CAS->image of a
    Gaussian-blob volume equals the analytic line integrals along n.  A
    transposed (u, v), a wrong sign or axis gives errors of tens of percent."""
    from multigpugridder_amd.synth import random_directions

    n, m = 32, 64
    vol, cen, sig, amp = _contained_blobs(n, 5)
    cas = oracle.vol_to_cas(vol, 2, 3).astype(np.float32)
    axes = random_directions(8, 3)
    imgs = oracle.cas_imgs_to_imgs(oracle.forward_project(cas, axes, ker32, m), n)
    for a in range(axes.shape[0]):
        u, v = axes[a, :3].astype(float), axes[a, 3:6].astype(float)
        ref = _analytic_projection(n, cen, sig, amp, u, v)
        assert rel_l2(imgs[a], ref) < 3e-3
        assert rel_l2(imgs[a], _analytic_projection(n, cen, sig, amp, v, u)) > 0.1


# ---------------------------------------------------------------------------
# Back projection: exact transpose and reference semantics
# ---------------------------------------------------------------------------
def test_adjoint_identity(oracle, ker32):
    from multigpugridder_amd.synth import random_directions

    rng = np.random.default_rng(2)
    vol = rng.standard_normal((V, V, V)).astype(np.float32)
    axes = random_directions(5, 7)
    y = rng.standard_normal((5, M, M)).astype(np.float32)
    lhs = np.sum(oracle.forward_project(vol, axes, ker32, M) * y)
    rhs = np.sum(vol.astype(np.float64) * oracle.back_project_adjoint(y, axes, ker32, V))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_bp_reference_equals_transpose_on_grid_directions(oracle, ker32):
    """For signed-permutation directions every tap is at an integer offset and
    the reference BP visits exactly the transpose's (pixel, voxel) pairs, as
    long as the image support stays clear of the mask edges."""
    rng = np.random.default_rng(3)
    axes = _signed_perm_axes()
    y = np.zeros((axes.shape[0], M, M), np.float32)
    c = M // 2
    y[:, c - 8:c + 8, c - 8:c + 8] = rng.standard_normal((axes.shape[0], 16, 16))
    ref = oracle.back_project(y, axes, ker32, V)
    adj = oracle.back_project_adjoint(y, axes, ker32, V)
    assert rel_l2(ref, adj) < 1e-14


def test_bp_reference_close_to_transpose_for_oblique(oracle, ker32):
    """Oblique directions: the reference BP drops the kernel tails that the 5x5
    window and the |fk| <= 2 slab gate exclude, so on smooth data it is close
    to (not equal to) the exact transpose; with u and v swapped it is ~20x
    further off."""
    from multigpugridder_amd.synth import blob_volume, random_directions

    axes = random_directions(6, 9)
    cas2 = oracle.vol_to_cas(blob_volume(N, 2), 2, PAD).astype(np.float32)
    y = oracle.forward_project(cas2, axes, ker32, M).astype(np.float32)
    adj = oracle.back_project_adjoint(y, axes, ker32, V)
    assert rel_l2(oracle.back_project(y, axes, ker32, V), adj) < 5e-3
    swapped = np.concatenate([axes[:, 3:6], axes[:, 0:3], -axes[:, 6:9]], axis=1)
    assert rel_l2(oracle.back_project(y, swapped, ker32, V), adj) > 5e-2


def test_bp_density_kat(oracle, ker32):
    """All-ones CAS images, identity direction: a voxel in the central plane
    gets (sum of integer taps)^2 * T(0) = T[250]; one plane off gets T[375];
    two planes off T[500] = 0; outside the sphere 0."""
    axes = np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1]], np.float32)
    ones = np.ones((1, M, M), np.float32)
    d = oracle.back_project(ones, axes, ker32, V)
    cv = V // 2
    s = float(ker32[125]) + float(ker32[250]) + float(ker32[375])
    assert d[cv, cv, cv] == pytest.approx(s * s * ker32[250], rel=1e-6)
    assert d[cv + 1, cv, cv] == pytest.approx(s * s * ker32[375], rel=1e-6)
    assert d[cv - 1, cv + 3, cv - 2] == pytest.approx(s * s * ker32[125], rel=1e-6)
    assert d[cv + 2, cv, cv] == 0.0 and d[cv - 2, cv, cv] == 0.0
    zz, yy, xx = np.meshgrid(np.arange(V), np.arange(V), np.arange(V), indexing="ij")
    outside = (xx - cv) ** 2 + (yy - cv) ** 2 + (zz - cv) ** 2 > (M // 2 - 1) ** 2
    assert np.all(d[outside] == 0.0)


def test_bp_reference_linear_index_wrap(oracle, ker32):
    """gpuBackProject.cu:99 bounds the image read by the linear index only, so
    a window column i1 = -1 reads the last pixel of the previous row.  With
    only the last image column (x = M-1) non-zero, voxels whose projection
    sits at the left edge (fi < 4, window columns <= 6) can receive a value
    only through that wrap."""
    t = np.deg2rad(25.0)
    u = np.array([np.cos(t), np.sin(t), 0.0])
    v = np.array([-np.sin(t), np.cos(t), 0.0])
    axes = np.concatenate([u, v, np.cross(u, v)]).astype(np.float32)[None]
    cv, ci = V // 2, M // 2
    y = np.zeros((1, M, M), np.float32)
    y[0, :, M - 1] = 1.0
    d = oracle.back_project(y, axes, ker32, V)
    wrapped = 0
    for z, yv, x in np.argwhere(d != 0):
        fi = (x - cv) * u[0] + (yv - cv) * u[1] + (z - cv) * u[2] + ci
        wrapped += fi < 4.0
    assert wrapped > 0


# ---------------------------------------------------------------------------
# CAS transforms and image chains
# ---------------------------------------------------------------------------
def _dft_matrix(m):
    """Centred DFT matrix E[k, x] = exp(-i 2 pi (k-c)(x-c)/m), built explicitly
    (no FFT, no fftshift)."""
    c = m // 2
    k = np.arange(m) - c
    return np.exp(-2j * np.pi * np.outer(k, k) / m)


def _cas2(e, x):
    f = np.einsum("ky,byx,lx->bkl", e, x, e)
    return f.real + f.imag


def test_image_chain_matches_explicit_hartley(oracle):
    """Forward image chain (gpuGridder.cpp:520-596) = ToCAS of the centred 2-D
    DFT of the zero-padded image."""
    rng = np.random.default_rng(6)
    n, m = 6, 12
    x = rng.standard_normal((2, n, n))
    pad = np.zeros((2, m, m))
    off = (m - n) // 2
    pad[:, off:off + n, off:off + n] = x
    assert rel_l2(oracle.imgs_to_cas_imgs(x, m), _cas2(_dft_matrix(m), pad)) < 1e-13


def test_inverse_image_chain_matches_explicit_hartley(oracle):
    """Inverse chain (gpuGridder.cpp:432-518): FromCAS drops the index-0 row and
    column, and the real part of the inverse DFT of the resulting Hermitian
    spectrum is the (self-inverse) Hartley transform divided by m^2."""
    rng = np.random.default_rng(7)
    n, m = 6, 12
    c = rng.standard_normal((2, m, m))
    cz = c.copy()
    cz[:, 0, :] = 0
    cz[:, :, 0] = 0
    off = (m - n) // 2
    want = _cas2(_dft_matrix(m), cz)[:, off:off + n, off:off + n] / (m * m)
    assert rel_l2(oracle.cas_imgs_to_imgs(c, n), want) < 1e-12


def test_cas_volume_involution(oracle):
    """The centred 3-D Hartley transform is its own inverse up to m^3."""
    rng = np.random.default_rng(8)
    m = 8
    x = rng.standard_normal((m, m, m))

    def h3(a):
        return oracle.to_cas(np.fft.fftshift(np.fft.fftn(np.fft.fftshift(a))))

    assert rel_l2(h3(h3(x)), m ** 3 * x) < 1e-13


def test_vol_to_cas_centre_delta_is_constant(oracle):
    vol = np.zeros((N, N, N))
    vol[N // 2, N // 2, N // 2] = 1.0
    cas = oracle.vol_to_cas(vol, 2, PAD)
    inner = cas[PAD:PAD + M, PAD:PAD + M, PAD:PAD + M]
    np.testing.assert_allclose(inner, 1.0, atol=1e-12)
    assert np.all(cas[:PAD] == 0) and np.all(cas[PAD + M:] == 0)


def test_vol_cas_round_trip(oracle):
    """volFromCAS(CASFromVol(x)) = x * precomp^2 (precompensation is applied on
    both sides), up to the Nyquist planes FromCAS drops."""
    vol, *_ = _contained_blobs(N, 9, k=4)
    rt = oracle.cas_to_vol(oracle.vol_to_cas(vol, 2, PAD), N, 2, PAD)
    p = oracle.precomp(M)[(M - N) // 2:(M - N) // 2 + N]
    w = p[:, None, None] * p[None, :, None] * p[None, None, :]
    assert rel_l2(rt, vol * w * w) < 1e-2
    # without the Nyquist zeroing the round trip is exact
    rt2 = oracle.cas_to_vol(oracle.vol_to_cas(vol, 2, PAD), N, 2, PAD, zero_nyquist=False)
    assert rel_l2(rt2, vol * w * w) < 1e-12


# ---------------------------------------------------------------------------
# Directions (utils/create_uniform_axes.m, get_axes_from_n_rot.m)
# ---------------------------------------------------------------------------
def test_uniform_axes_are_rotations(oracle):
    a = oracle.uniform_axes(10, 50).astype(np.float64)
    assert a.shape == (500, 9)
    r = a.reshape(-1, 3, 3)  # rows: u, v, n
    np.testing.assert_allclose(np.einsum("aij,akj->aik", r, r), np.broadcast_to(np.eye(3), r.shape), atol=1e-6)
    np.testing.assert_allclose(np.linalg.det(r), 1.0, atol=1e-6)
    # normal of direction (kt, kp) is (sin t cos p, sin t sin p, cos t)
    kt, kp = 3, 7
    t, p = np.deg2rad(kt * 9.0), np.deg2rad(kp * 7.2)
    np.testing.assert_allclose(r[(kt - 1) * 50 + kp, 2], [np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)],
                               atol=1e-6)


# ---------------------------------------------------------------------------
# Golden fixtures written by tests/golden/make_golden.py from this oracle
# ---------------------------------------------------------------------------
def test_golden_fixtures_reproduce(oracle, ker32):
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "projector_n16.npz")
    g = np.load(path)
    np.testing.assert_array_equal(ker32, g["ker"])
    fp = oracle.forward_project(g["cas_vol"], g["axes"], g["ker"], int(g["m"]))
    np.testing.assert_allclose(fp, g["fp"], rtol=0, atol=1e-12)
    bp = oracle.back_project(g["cas_imgs"], g["axes"], g["ker"], int(g["v"]))
    np.testing.assert_allclose(bp, g["bp"], rtol=0, atol=1e-12)
    adj = oracle.back_project_adjoint(g["cas_imgs"], g["axes"], g["ker"], int(g["v"]))
    np.testing.assert_allclose(adj, g["adj"], rtol=0, atol=1e-12)


# ---------------------------------------------------------------------------
# Transposes used by the differentiable projector, pinned by explicit matrices
# ---------------------------------------------------------------------------
def _matrix(fn, in_shape):
    cols = []
    for k in range(int(np.prod(in_shape))):
        e = np.zeros(int(np.prod(in_shape)))
        e[k] = 1.0
        cols.append(np.asarray(fn(e.reshape(in_shape)), dtype=np.float64).ravel())
    return np.stack(cols, axis=1)


def test_image_chain_transpose_is_exact(oracle):
    n, m = 4, 8
    A = _matrix(lambda c: oracle.cas_imgs_to_imgs(c, n), (1, m, m))        # (n^2, m^2)
    T = _matrix(lambda y: oracle.cas_imgs_to_imgs_transpose(y, m), (1, n, n))  # (m^2, n^2)
    assert np.max(np.abs(A.T - T)) < 1e-12
    # without clearing the index-0 row / column the candidate is not the transpose
    T0 = _matrix(lambda y: oracle.imgs_to_cas_imgs(y, m) / m ** 2, (1, n, n))
    assert np.max(np.abs(A.T - T0)) > 1e-3


def test_volume_chain_transpose_is_exact(oracle):
    n, m, pad = 4, 8, 1
    G = _matrix(lambda x: oracle.vol_to_cas(x, 2, pad), (n, n, n))               # (V^3, n^3)
    T = _matrix(lambda c: oracle.vol_to_cas_transpose(c, n, 2, pad), (m + 2 * pad,) * 3)
    assert np.max(np.abs(G.T - T)) < 1e-9 * np.max(np.abs(G))
    # FromCAS's zeroing of the index-0 planes would break it
    T0 = _matrix(lambda c: m ** 3 * oracle.cas_to_vol(c, n, 2, pad, zero_nyquist=True), (m + 2 * pad,) * 3)
    assert np.max(np.abs(G.T - T0)) > 1e-3 * np.max(np.abs(G))


def test_projector_adjoint_identity(oracle, ker32):
    n, p = 8, 9
    from multigpugridder_amd.synth import random_directions
    axes = random_directions(p, 21)
    rng = np.random.default_rng(4)
    v = rng.standard_normal((n, n, n))
    y = rng.standard_normal((p, n, n))
    lhs = np.sum(oracle.projector(v, axes, ker32) * y)
    rhs = np.sum(v * oracle.projector_transpose(y, axes, ker32, n))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 1e-6
