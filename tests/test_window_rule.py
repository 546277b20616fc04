"""CPU check of the rule behind the 4x4x4 windows of the forward gather and
its transpose (fp_lds_tile_body, adj_lds_tile with W4).

Along one axis the reference gathers the taps X-2..X+2 around X = roundf(x),
each with r = (float)tap - x clamped to [-2, 2] and the table index
roundf(fma(r, 125, 250)) clamped to [0, 500] (gpuForwardProject.cu:63-84, the
contracted reading).  The kernels skip one end tap per axis and gather the four
taps from floor(x) - 1.  This recomputes the reference's indices in exact
arithmetic for random and adversarial positions and checks that the skipped
tap always reads table entry 0 (tap X-2, when x >= X) or 500 (tap X+2, when
x < X), that the kernels' window start X - 2 + (x >= X) is floor(x) - 1, and
that the reference's table has both end entries exactly 0.
"""
import numpy as np


def _roundf(v):
    """C roundf (half away from zero) of float32 values, exactly, as int64."""
    v = v.astype(np.float64)
    return (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)


def _ref_index(tap, x):
    r = (tap.astype(np.float32) - x).astype(np.float32)  # (float)tap - x, one FP32 rounding
    r = np.clip(r, np.float32(-2), np.float32(2))
    s = (r.astype(np.float64) * 125.0 + 250.0).astype(np.float32)  # fma: exact product and sum, one rounding
    return np.clip(_roundf(s), 0, 500)


def _positions():
    rng = np.random.default_rng(11)
    x = [rng.uniform(2.0, 300.0, 2_000_000).astype(np.float32)]
    k = np.arange(2, 300, dtype=np.float32)
    for base in (k, k + np.float32(0.5), k + np.float32(0.004), k - np.float32(0.004),
                 k + np.float32(0.5 - 0.004), k + np.float32(0.5 + 0.004)):
        for n in range(-40, 41):
            v = base
            step = np.inf if n > 0 else -np.inf
            for _ in range(abs(n)):
                v = np.nextafter(v, np.float32(step)).astype(np.float32)
            x.append(v.astype(np.float32))
    return np.concatenate(x)


def test_skipped_tap_reads_a_zero_entry_and_window_starts_at_floor_minus_one():
    x = _positions()
    X = _roundf(x)
    up = x.astype(np.float64) >= X  # x >= X: drop tap X - 2, else drop tap X + 2
    dropped = np.where(up, X - 2, X + 2)
    idx = _ref_index(dropped, x)
    assert np.all(idx[up] == 0)
    assert np.all(idx[~up] == 500)
    start = X - 2 + up.astype(np.int64)
    assert np.array_equal(start, np.floor(x.astype(np.float64)).astype(np.int64) - 1)
    # the four kept taps are exactly the reference's other four
    for t in range(4):
        kept = start + t
        assert np.all((kept >= X - 2) & (kept <= X + 2) & (kept != dropped))


def test_reference_table_ends_are_zero(oracle):
    ker = oracle.kb_table().astype(np.float32)
    assert ker.shape == (501,)
    assert ker[0] == 0.0 and ker[-1] == 0.0
    assert np.all(ker[1:-1] > 0)  # so only the end entries can be skipped
