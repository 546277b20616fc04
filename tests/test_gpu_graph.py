"""The device-resident projection step captured into a hipGraph: every replay,
with the program's own eager work and synchronisations between replays,
reproduces the eager step exactly, and the eager step matches the oracle."""
import numpy as np
import pytest

from conftest import assert_parity, rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="hipGraph capture of the pipeline is switched off: its one run ended in a "
                                      "host segmentation fault at the end of capture (torch.cuda.graph __exit__), "
                                      "cause not found")]


def test_captured_step_replays_match_eager(cuda, oracle, ker32):
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.pipeline import ProjectionPipeline
    from multigpugridder_amd.synth import blob_volume, random_directions

    n, p = 64, 300
    m, v = 2 * n, 2 * n + 6
    print(f"graph capture n={n} p={p}, chunks of 128 on 2 streams", flush=True)
    vol = blob_volume(n, 1)
    cas = oracle.vol_to_cas(vol, 2, 3).astype(np.float32)
    axes = random_directions(p, 31)
    pipe = ProjectionPipeline(torch.from_numpy(cas).to(cuda), torch.from_numpy(axes).to(cuda), n, chunk=128,
                              num_streams=2, ker=torch.from_numpy(ker32).to(cuda))
    pipe.step()
    torch.cuda.synchronize()
    imgs0, bp0 = pipe.imgs.clone(), pipe.bp_vol.clone()
    ref = oracle.cas_imgs_to_imgs(oracle.forward_project(cas, axes, ker32, m), n)
    assert rel_l2(imgs0.cpu().numpy(), ref) < 1e-5
    bp_ref = oracle.back_project(oracle.imgs_to_cas_imgs(imgs0.cpu().numpy(), m).astype(np.float32), axes, ker32, v)
    assert_parity(bp0.cpu().numpy(), bp_ref)

    pipe.capture()
    for r in range(3):
        # eager work between replays: clobber the outputs, then synchronise
        pipe.imgs.fill_(float("nan"))
        pipe.bp_vol.fill_(123.0)
        ops.forward_project(pipe.cas_vol, pipe.axes[:5], m, ker=pipe.ker)
        torch.cuda.synchronize()
        pipe.replay()
        torch.cuda.synchronize()
        assert torch.equal(pipe.imgs, imgs0), f"replay {r}: images differ"
        assert torch.equal(pipe.bp_vol, bp0), f"replay {r}: back projection differs"
