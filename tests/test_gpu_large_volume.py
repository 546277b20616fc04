"""A CAS volume of more than 4 GB: N=512 at interp 2 gives V = 1030, 1.09e9
FP32 voxels (4.4 GB), so voxel byte offsets pass 2^32 while the reference's
int linear index still fits.  The staged forward gather addresses each box
from its first element; both window widths, and the back projector, must
equal the oracle there."""
import numpy as np
import pytest

from conftest import assert_parity

pytestmark = pytest.mark.gpu


def test_forward_project_volume_over_4gb(cuda, oracle, ker32):
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    n = 512
    m, v = 2 * n, 2 * n + 6
    assert v ** 3 * 4 > 2 ** 32 and v ** 3 < 2 ** 31
    print(f"forward projection, CAS volume {v}^3 FP32 ({v ** 3 * 4 / 2 ** 30:.2f} GiB), 3 directions", flush=True)
    g = torch.Generator(device=cuda).manual_seed(11)
    dc = torch.randn((v, v, v), generator=g, device=cuda)
    axes = random_directions(3, 12)
    da = torch.from_numpy(axes).to(cuda)
    dk = torch.from_numpy(ker32).to(cuda)
    flag = ops.check_volume(dc, dk)
    w4 = ops.forward_project(dc, da, m, ker=dk, check=flag)
    w5 = ops.forward_project(dc, da, m, ker=dk, check=False)
    torch.cuda.synchronize()
    assert int(flag.item()) == 1
    cas = dc.cpu().numpy()
    del dc
    print("  oracle", flush=True)
    ref = oracle.forward_project(cas, axes, ker32, m)
    assert_parity(w4.cpu().numpy(), ref)
    assert_parity(w5.cpu().numpy(), ref)


def test_back_project_volume_over_4gb(cuda, oracle, ker32):
    """The back projector into a 1030^3 FP32 volume (brick indices and voxel
    offsets past 2^32 bytes), against the oracle plane block by plane block."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.synth import random_directions

    n = 512
    m, v = 2 * n, 2 * n + 6
    print(f"back projection into a {v}^3 FP32 volume, 2 directions", flush=True)
    g = torch.Generator(device=cuda).manual_seed(13)
    imgs = torch.randn((2, m, m), generator=g, device=cuda)
    axes = random_directions(2, 14)
    dk = torch.from_numpy(ker32).to(cuda)
    out = ops.back_project(imgs, torch.from_numpy(axes).to(cuda), v, ker=dk)
    torch.cuda.synchronize()
    print("  oracle", flush=True)
    ref = oracle.back_project(imgs.cpu().numpy(), axes, ker32, v)
    err2 = ref2 = worst = peak = 0.0
    for z0 in range(0, v, 64):
        a = out[z0:z0 + 64].cpu().numpy().astype(np.float64)
        b = ref[z0:z0 + 64]
        err2 += float(np.sum((a - b) ** 2))
        ref2 += float(np.sum(b ** 2))
        worst = max(worst, float(np.max(np.abs(a - b))))
        peak = max(peak, float(np.max(np.abs(b))))
    assert ref2 > 0
    assert (err2 / ref2) ** 0.5 < 1e-5
    assert worst <= 2e-6 * peak
