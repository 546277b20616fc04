"""GPU tests of the differentiable projector: the transpose operators against
the oracle, the dot-product adjoint test and FP64 central finite differences
at the differentiation configuration (blob volume 96^3, interp 2, random
directions), and a descent loop on ||P v - y||^2."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_transpose_operators_match_oracle(cuda, oracle):
    import torch
    from multigpugridder_amd import ops

    n, m, p = 24, 48, 5
    rng = np.random.default_rng(3)
    imgs = rng.standard_normal((p, n, n)).astype(np.float32)
    got = ops.cas_images_to_images_transpose(torch.from_numpy(imgs).cuda(), m).cpu().numpy()
    assert rel_l2(got, oracle.cas_imgs_to_imgs_transpose(imgs, m)) < 1e-5
    assert np.all(got[:, 0, :] == 0) and np.all(got[:, :, 0] == 0)
    cas = rng.standard_normal((m + 6,) * 3).astype(np.float32)
    got = ops.volume_to_cas_volume_transpose(torch.from_numpy(cas).cuda(), n).cpu().numpy()
    assert rel_l2(got, oracle.vol_to_cas_transpose(cas, n)) < 1e-5


def test_adjoint_dot_product(cuda):
    """|<Pv,y> - <v,P^T y>| / max(...) at N=96, interp 2, 2,048 random directions,
    random FP32 v and y (seed 4)."""
    import torch
    from multigpugridder_amd.autograd import project, project_transpose
    from multigpugridder_amd.synth import random_directions

    n, p = 96, 2048
    print(f"adjoint test n={n} p={p}", flush=True)
    axes = torch.from_numpy(random_directions(p, 104)).cuda()
    g = torch.Generator(device="cpu").manual_seed(4)
    v = torch.randn((n, n, n), generator=g).cuda()
    y = torch.randn((p, n, n), generator=g).cuda()
    pv = project(v, axes)
    pty = project_transpose(y, axes, n)
    lhs = float((pv.double() * y.double()).sum())
    rhs = float((v.double() * pty.double()).sum())
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"adjoint mismatch {err:.3e}", flush=True)
    assert err < 1e-5


def test_gradient_matches_fp64_finite_differences(cuda, oracle, ker32):
    """Gradient of ||P v - y||^2 (autograd, backward = adjoint kernels) against
    FP64 central differences from the oracle on 32 random voxels (seed 5).
    Blob volume A (seed 1) at N=96; y = P(volume B, seed 2) + N(0, (0.1 std)^2)
    (seed 3).  The oracle differences use a 32-direction subset of the 2,048
    directions so that they finish in seconds."""
    import torch
    from multigpugridder_amd.autograd import project
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 96, 32
    print(f"finite differences n={n} p={p} voxels=32", flush=True)
    axes_np = random_directions(2048, 104)[:p]
    axes = torch.from_numpy(axes_np).cuda()
    va = blob_volume(n, 1)
    vb = torch.from_numpy(blob_volume(n, 2)).cuda()
    with torch.no_grad():
        yb = project(vb, axes)
    noise = np.random.default_rng(3).standard_normal(tuple(yb.shape)).astype(np.float32)
    y = yb + torch.from_numpy(noise).cuda() * 0.1 * float(yb.std())
    v = torch.from_numpy(va).cuda().requires_grad_(True)
    loss = ((project(v, axes) - y) ** 2).sum()
    loss.backward()
    grad = v.grad.cpu().numpy().astype(np.float64)

    y64 = y.cpu().numpy().astype(np.float64)
    r = oracle.projector(va.astype(np.float64), axes_np, ker32) - y64
    rng = np.random.default_rng(5)
    idx = [tuple(int(c) for c in rng.integers(n // 4, 3 * n // 4, 3)) for _ in range(32)]
    h = 1e-2
    fd = []
    for (z, yy, x) in idx:
        e = np.zeros((n, n, n))
        e[z, yy, x] = 1.0
        pe = oracle.projector(e, axes_np, ker32)
        lp = np.sum((r + h * pe) ** 2)
        lm = np.sum((r - h * pe) ** 2)
        fd.append((lp - lm) / (2 * h))
    fd = np.array(fd)
    g = np.array([grad[i] for i in idx])
    err = rel_l2(g, fd)
    print(f"gradient vs FP64 central differences: rel L2 {err:.3e}", flush=True)
    assert err < 1e-4


def test_descent_on_projection_loss(cuda):
    """Steepest descent with the exact step on ||P v - y||^2 lowers the loss
    every step (autograd gradient, one extra projection per step)."""
    import torch
    from multigpugridder_amd.autograd import project
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 32, 256
    axes = torch.from_numpy(random_directions(p, 9)).cuda()
    with torch.no_grad():
        y = project(torch.from_numpy(blob_volume(n, 2)).cuda(), axes)
    v = torch.zeros((n, n, n), device="cuda", requires_grad=True)
    losses = []
    for _ in range(6):
        loss = ((project(v, axes) - y) ** 2).sum()
        losses.append(float(loss.detach()))
        (g,) = torch.autograd.grad(loss, v)
        with torch.no_grad():
            pg = project(g, axes)
            step = float((g * g).sum()) / (2.0 * float((pg * pg).sum()))
            v -= step * g
    print("losses", [f"{x:.4e}" for x in losses], flush=True)
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.5 * losses[0]
