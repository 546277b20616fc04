"""A CAS volume of more than 4 GB: N=512 at interp 2 gives V = 1030, 1.09e9
FP32 voxels (4.4 GB), so voxel byte offsets pass 2^32 while the reference's
int linear index still fits.  The staged forward gather addresses each box
from its first element; both window widths must equal the oracle there."""
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
