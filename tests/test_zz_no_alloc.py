"""The allocation-free image chain (mgg_op_cas_images_to_images_noalloc) against
the allocating entry point it mirrors (mgg_op_cas_images_to_images), through
the C ABI: bit-identical images on two calls in a row, the second smaller, so
the shared fixed scratch left by the first call cannot leak into the second;
the input is not modified; unsupported sizes are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run_pair(cuda, oracle, m, nn, n, seed):
    import torch

    from multigpugridder_amd import _lib

    print(f"allocation-free image chain m={m} nn={nn} n={n}", flush=True)
    c = torch.from_numpy(np.random.default_rng(seed).standard_normal((n, m, m)).astype(np.float32)).to(cuda)
    keep = c.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    got = torch.full((n, nn, nn), np.nan, device=cuda)
    _lib.call("mgg_op_cas_images_to_images_noalloc", ctypes.c_void_p(c.data_ptr()),
              ctypes.c_void_p(got.data_ptr()), n, m, nn, s)
    ref = torch.empty((n, nn, nn), device=cuda)
    work = c.clone()  # the allocating entry point may clobber its input
    _lib.call("mgg_op_cas_images_to_images", ctypes.c_void_p(work.data_ptr()),
              ctypes.c_void_p(ref.data_ptr()), n, m, nn, s)
    torch.cuda.synchronize()
    assert torch.equal(c, keep), "input modified"
    assert torch.equal(got, ref), "differs from mgg_op_cas_images_to_images"
    sub = slice(0, min(n, 40))
    assert rel_l2(got[sub].cpu().numpy(), oracle.cas_imgs_to_imgs(keep[sub].cpu().numpy(), nn)) < 1e-5


def test_noalloc_matches_allocating_chain_twice(cuda, oracle):
    # 700 images at m = 256 span three chunks of the 32 MiB scratch (254 each);
    # the second call is smaller and reuses what the first left behind
    _run_pair(cuda, oracle, 256, 128, 700, 1)
    _run_pair(cuda, oracle, 256, 128, 37, 2)


def test_noalloc_other_sizes(cuda, oracle):
    _run_pair(cuda, oracle, 512, 256, 70, 3)
    _run_pair(cuda, oracle, 128, 64, 9, 4)


def test_noalloc_refuses_unsupported_sizes(cuda):
    import torch

    from multigpugridder_amd import _lib

    c = torch.zeros((2, 200, 200), device=cuda)
    out = torch.zeros((2, 100, 100), device=cuda)
    for m, nn, n in ((200, 100, 2), (256, 257, 2), (256, 128, -1)):
        with pytest.raises(_lib.MggError) as e:
            _lib.call("mgg_op_cas_images_to_images_noalloc", ctypes.c_void_p(c.data_ptr()),
                      ctypes.c_void_p(out.data_ptr()), n, m, nn, None)
        assert e.value.code in (_lib.MGG_ERR_UNSUPPORTED, _lib.MGG_ERR_INVALID_ARGUMENT)
