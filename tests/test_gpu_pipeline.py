"""The device-resident projection step (ProjectionPipeline, the bench's step)
against the oracle: forward projection through to images, then back
projection of those images, for one, two and three streams and a direction
count that leaves a partial last chunk; two steps in a row give the same
volume (the step zeroes its accumulator)."""
import numpy as np
import pytest

from conftest import assert_parity, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("streams,chunk", [(1, 64), (2, 64), (3, 96), (2, 300)])
def test_pipeline_step_matches_oracle(cuda, oracle, ker32, streams, chunk):
    import torch

    from multigpugridder_amd.pipeline import ProjectionPipeline
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 32, 300
    m, v = 2 * n, 2 * n + 6
    print(f"pipeline step n={n} p={p} streams={streams} chunk={chunk}", flush=True)
    cas = oracle.vol_to_cas(blob_volume(n, 2), 2, 3).astype(np.float32)
    axes = random_directions(p, 33)
    pipe = ProjectionPipeline(torch.from_numpy(cas).to(cuda), torch.from_numpy(axes).to(cuda), n, chunk=chunk,
                              num_streams=streams, ker=torch.from_numpy(ker32).to(cuda))
    pipe.step()
    torch.cuda.synchronize()
    imgs = pipe.imgs.cpu().numpy()
    ref = oracle.cas_imgs_to_imgs(oracle.forward_project(cas, axes, ker32, m), n)
    assert rel_l2(imgs, ref) < 1e-5
    bp = pipe.bp_vol.cpu().numpy()
    bp_ref = oracle.back_project(oracle.imgs_to_cas_imgs(imgs, m).astype(np.float32), axes, ker32, v)
    assert_parity(bp, bp_ref)
    pipe.step()
    torch.cuda.synchronize()
    assert np.array_equal(pipe.imgs.cpu().numpy(), imgs)
    assert_parity(pipe.bp_vol.cpu().numpy(), bp, tol=1e-6, peak=1e-6)
