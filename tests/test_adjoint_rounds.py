"""CPU check of the exact-transpose kernel's collision rule (adj_lds_tile).

The kernel adds each 8x8 tile's taps into an LDS box with plain
read-modify-writes in rounds: a lane waits while a lower pending lane among
its left, up, up-left and up-right tile neighbours has the same stencil
centre (X, Y, Z).  That is only safe if, for an orthonormal frame, two pixels
of a tile with the same rounded centre are always tile neighbours.  This
recomputes the kernel's centres with its arithmetic (the contracted reading of
gpuForwardProject.cu:53-59: X = roundf(fl(fma(u0, fi, fl(v0 * fj)) + c))) for
random rotations and for near-axis and diagonal frames, over every tile of
M = 256 images, and checks the claim, then replays the round rule and checks
that no two lanes with one centre ever run in the same round.
"""
import numpy as np

from multigpugridder_amd.synth import random_directions

M, V = 256, 262


def centres(axes):
    """(P, M, M, 3) int rounded stencil centres, as the kernel computes them."""
    ci, cv = M // 2, np.float32(V // 2)
    f = np.arange(M, dtype=np.float32) - np.float32(ci)
    fi = f[None, None, :]  # pixel column i
    fj = f[None, :, None]  # pixel row j
    out = []
    for k in range(3):
        u = axes[:, k].astype(np.float32)[:, None, None]
        v = axes[:, 3 + k].astype(np.float32)[:, None, None]
        bv = (v * fj).astype(np.float32)                       # fl(v * fj)
        t = (u.astype(np.float64) * fi + bv).astype(np.float32)  # fma: one rounding of an exact sum
        x = (t + cv).astype(np.float32)                        # fl(... + c)
        out.append(np.floor(x.astype(np.float64) + 0.5).astype(np.int64))  # roundf (x > 0 here)
    return np.stack(out, -1)


def frames():
    rnd = random_directions(120, 77)
    special = []
    for a in np.deg2rad([0.0, 0.01, 0.5, 30.0, 45.0, 60.0]):
        for b in np.deg2rad([0.0, 0.02, 35.26, 45.0, 90.0]):
            ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
            rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
            rx = np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])
            r = rz @ rx
            special.append(np.concatenate([r[:, 0], r[:, 1], r[:, 2]]))
    return np.concatenate([rnd, np.array(special, np.float32)])


def test_equal_centres_are_tile_neighbours_and_rounds_separate_them():
    axes = frames()
    worst = 0
    dup_tiles = 0
    for a0 in range(0, axes.shape[0], 50):
        c = centres(axes[a0:a0 + 50])
        p = c.shape[0]
        # tiles: (P, 32, 8, 32, 8, 3) -> (P, 1024 tiles, 64 lanes, 3), lane = i + 8 j
        t = c.reshape(p, M // 8, 8, M // 8, 8, 3).transpose(0, 1, 3, 2, 4, 5).reshape(p, -1, 64, 3)
        key = (t[..., 0] * 1_000_003 + t[..., 1]) * 1_000_003 + t[..., 2]
        ln = np.arange(64)
        ix, jy = ln % 8, ln // 8
        same = key[..., :, None] == key[..., None, :]
        same &= ~np.eye(64, dtype=bool)
        if not same.any():
            continue
        dup_tiles += int(same.any(axis=(-1, -2)).sum())
        di = np.abs(ix[:, None] - ix[None, :])
        dj = np.abs(jy[:, None] - jy[None, :])
        dist = np.maximum(di, dj)
        worst = max(worst, int(dist[np.nonzero(same)[-2:]].max()))
        # replay the rounds on tiles with duplicates: lower pending neighbour with the same key -> wait
        for idx in zip(*np.nonzero(same.any(axis=(-1, -2)))):
            k = key[idx]
            pend = np.ones(64, bool)
            rounds = 0
            while pend.any():
                kk = np.where(pend, k, -1 - ln)
                wait = np.zeros(64, bool)
                for off, ok in ((-1, ix >= 1), (-8, jy >= 1), (-9, (ix >= 1) & (jy >= 1)), (-7, (ix <= 6) & (jy >= 1))):
                    src = np.clip(ln + off, 0, 63)
                    wait |= pend & ok & (kk[src] == kk)
                go = pend & ~wait
                assert go.any()
                gk = k[go]
                assert len(np.unique(gk)) == len(gk), "two lanes with one centre in the same round"
                pend = wait
                rounds += 1
            assert rounds <= 4
    assert dup_tiles > 0  # the rule is exercised: some tiles do have shared centres
    assert worst <= 1, f"equal centres {worst} tiles apart"
