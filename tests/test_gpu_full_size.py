"""The device-resident step at the plain drop-in configuration's full size
(N=128, interp 2, the 10,000-direction 100 x 100 grid, launches of 2,000 on two
streams, as the bench times it), checked where the oracle cannot run the whole
workload in seconds: a random sample of the 10,000 images against the oracle
(relative L2 and the max-abs guard), the same images bit for bit and the same
volume to FP32 accumulation order under another chunking, and the back
projection additive over a split of the directions (one launch of all 10,000,
two launches of 5,000, five of 2,000 all agree)."""
import numpy as np
import pytest

from conftest import assert_parity

pytestmark = pytest.mark.gpu


def test_c2_full_size_step(cuda, oracle, ker32):
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.hostmath import uniform_axes
    from multigpugridder_amd.pipeline import ProjectionPipeline
    from multigpugridder_amd.synth import blob_volume

    n = 128
    m, v = 2 * n, 2 * n + 6
    print(f"full-size step n={n}, 10,000 grid directions", flush=True)
    cas = oracle.vol_to_cas(blob_volume(n, 1), 2, 3).astype(np.float32)
    axes = np.ascontiguousarray(uniform_axes(100, 100), dtype=np.float32)
    p = axes.shape[0]
    dc, da = torch.from_numpy(cas).to(cuda), torch.from_numpy(axes).to(cuda)
    dk = torch.from_numpy(ker32).to(cuda)

    pipe = ProjectionPipeline(dc, da, n, chunk=2000, num_streams=2, ker=dk)
    pipe.step()
    torch.cuda.synchronize()
    imgs = pipe.imgs.cpu().numpy()
    bp = pipe.bp_vol.cpu().numpy()
    assert np.isfinite(imgs).all() and np.isfinite(bp).all()

    # a random sample of the images against the oracle
    idx = np.sort(np.random.default_rng(7).choice(p, 48, replace=False))
    print("  48 sampled images vs oracle", flush=True)
    ref = oracle.cas_imgs_to_imgs(oracle.forward_project(cas, axes[idx], ker32, m), n)
    assert_parity(imgs[idx], ref)

    # another chunking: three streams, launches of 1,500 and a partial last one
    print("  chunks of 1,500 on three streams", flush=True)
    pipe2 = ProjectionPipeline(dc, da, n, chunk=1500, num_streams=3, ker=dk)
    pipe2.step()
    torch.cuda.synchronize()
    assert np.array_equal(pipe2.imgs.cpu().numpy(), imgs)
    assert_parity(pipe2.bp_vol.cpu().numpy(), bp, tol=1e-6, peak=1e-6)
    del pipe2

    # the back projection over all directions in one launch, and in two halves
    print("  one launch of 10,000 and two of 5,000", flush=True)
    cimg = ops.images_to_cas_images(pipe.imgs, m)
    one = ops.back_project(cimg, da, v, ker=dk)
    h1 = ops.back_project(cimg[: p // 2], da[: p // 2], v, ker=dk)
    h2 = ops.back_project(cimg[p // 2:], da[p // 2:], v, ker=dk)
    torch.cuda.synchronize()
    halves = h1.cpu().numpy().astype(np.float64) + h2.cpu().numpy()
    assert_parity(one.cpu().numpy(), bp, tol=1e-6, peak=1e-6)
    assert_parity(halves, bp, tol=1e-6, peak=1e-6)
