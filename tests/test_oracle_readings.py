"""The oracle's arithmetic readings and its mask-radius conversion (CPU).

The reference is built by nvcc with no flags (src/CMakeLists.txt:80-83), and
nvcc's default --fmad=true contracts the multiply-adds that feed rounding
decisions (gpuForwardProject.cu:53-55,67,76,84; gpuBackProject.cu:62-64,
77-94).  The oracle restates several readings of that source
(oracle/gridder_oracle.c) and pins one, "first": contraction without
reassociation, each add of two products fusing its first operand.  These
tests check that the readings are distinct (so a parity test against the
pinned one tells a kernel that follows another reading apart), that they
differ only through rounding decisions, and that they agree wherever the
arithmetic is exact.
"""
import numpy as np
import pytest

from conftest import rel_l2

OTHER = ("source", "second", "reassoc", "index")


def _signed_permutation_frames():
    frames = []
    perms = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (1, 0, 2)]
    signs = [(1, 1, 1), (-1, 1, -1), (1, -1, -1), (-1, -1, 1)]
    for pm in perms:
        for sg in signs:
            r = np.zeros((3, 3), dtype=np.float32)
            for c in range(3):
                r[pm[c], c] = sg[c]
            frames.append(r.T.reshape(9))  # columns u, v, n
    return np.array(frames, dtype=np.float32)


def test_default_reading_is_the_contracted_one(oracle):
    assert oracle.current_reading() == oracle.READING_FIRST
    with oracle.reading("source"):
        assert oracle.current_reading() == oracle.READING_SOURCE
    assert oracle.current_reading() == oracle.READING_FIRST
    with pytest.raises(ValueError):
        with oracle.reading(7):
            pass


@pytest.fixture(scope="module")
def small_case(oracle):
    from multigpugridder_amd.synth import blob_volume, random_directions
    n = 32
    m, v = 2 * n, 2 * n + 6
    cas = oracle.vol_to_cas(blob_volume(n, 1), 2, 3).astype(np.float32)
    axes = random_directions(60, 7)
    noise = np.random.default_rng(3).standard_normal((60, m, m)).astype(np.float32)
    return cas, axes, noise, m, v


def test_readings_are_distinct_and_differ_only_in_decisions(oracle, ker32, small_case):
    cas, axes, noise, m, v = small_case
    out = {}
    for r in ("first",) + OTHER:
        with oracle.reading(r):
            out[r] = (oracle.forward_project(cas, axes, ker32, m), oracle.back_project(noise, axes, ker32, v))
    fp0, bp0 = out["first"]
    for r in OTHER:
        fp, bp = out[r]
        # every reading takes some decision differently ...
        assert np.any(fp != fp0) and np.any(bp != bp0), r
        # ... and nothing else: a flipped table entry or rounded coordinate
        # moves single pixels / voxels, far below a wrong formula's O(1)
        assert rel_l2(fp, fp0) < 1e-4, r
        assert rel_l2(bp, bp0) < 1e-3, r
    # the back projection of noise separates the readings beyond the parity bar
    assert max(rel_l2(out[r][1], bp0) for r in OTHER) > 1e-5


def test_readings_agree_on_exact_geometry(oracle, ker32):
    # Signed-permutation frames: every product is exact, so every reading
    # computes the same coordinates and indices.
    from multigpugridder_amd.synth import blob_volume
    n = 16
    m, v = 2 * n, 2 * n + 6
    cas = oracle.vol_to_cas(blob_volume(n, 2), 2, 3).astype(np.float32)
    axes = _signed_permutation_frames()
    imgs = np.random.default_rng(5).standard_normal((axes.shape[0], m, m)).astype(np.float32)
    fp0 = oracle.forward_project(cas, axes, ker32, m)
    bp0 = oracle.back_project(imgs, axes, ker32, v)
    for r in OTHER:
        with oracle.reading(r):
            np.testing.assert_array_equal(oracle.forward_project(cas, axes, ker32, m), fp0)
            np.testing.assert_array_equal(oracle.back_project(imgs, axes, ker32, v), bp0)


def test_mask_radius_is_truncated_toward_zero(oracle, ker32, small_case):
    # RunKernel takes the radius as an int (gpuForwardProject.h:57,
    # gpuBackProject.h:56), so 19.9 behaves as 19 and differs from 20.
    cas, axes, noise, m, v = small_case
    axes = axes[:8]
    f19 = oracle.forward_project(cas, axes, ker32, m, mask_radius=19.0)
    np.testing.assert_array_equal(oracle.forward_project(cas, axes, ker32, m, mask_radius=19.9), f19)
    assert np.any(oracle.forward_project(cas, axes, ker32, m, mask_radius=20.0) != f19)
    b19 = oracle.back_project(noise[:8], axes, ker32, v, mask_radius=19.0)
    np.testing.assert_array_equal(oracle.back_project(noise[:8], axes, ker32, v, mask_radius=19.5), b19)
    assert np.any(oracle.back_project(noise[:8], axes, ker32, v, mask_radius=20.0) != b19)
    a19 = oracle.back_project_adjoint(noise[:8], axes, ker32, v, mask_radius=19.0)
    np.testing.assert_array_equal(oracle.back_project_adjoint(noise[:8], axes, ker32, v, mask_radius=19.7), a19)
