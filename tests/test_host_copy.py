"""CPU checks of the gridder class's host copy of image stacks (the threaded
path for large C-contiguous float32 stacks, numpy for everything else)."""
import numpy as np
import pytest

from multigpugridder_amd.gridder import _copy_into


@pytest.mark.parametrize("shape", [(3, 8, 8), (40, 128, 128)])
def test_copy_matches_numpy(shape):
    src = np.random.default_rng(0).standard_normal(shape).astype(np.float32)
    dst = np.empty_like(src)
    _copy_into(dst, src)
    assert np.array_equal(dst, src)


def test_copy_of_views_read_only_and_other_dtypes():
    rng = np.random.default_rng(1)
    src = rng.standard_normal((40, 128, 128)).astype(np.float32)
    dst = np.zeros_like(src)
    _copy_into(dst, src[:, ::-1])  # not C-contiguous
    assert np.array_equal(dst, src[:, ::-1])
    ro = src.copy()
    ro.setflags(write=False)
    _copy_into(dst, ro)
    assert np.array_equal(dst, src)
    _copy_into(dst, src.astype(np.float64))  # same-kind cast
    assert np.array_equal(dst, src)
    big = np.zeros((50, 128, 128), np.float32)
    _copy_into(big[:40], src)  # leading slice of a larger stack
    assert np.array_equal(big[:40], src) and not big[40:].any()


def test_copy_refuses_mismatched_shapes():
    with pytest.raises(ValueError):
        _copy_into(np.empty((3, 8, 8), np.float32), np.zeros((4, 8, 8), np.float32))
