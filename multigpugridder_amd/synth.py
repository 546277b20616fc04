"""Seeded synthetic inputs for tests and benchmarks.

* :func:`blob_volume` - sum of isotropic Gaussian blobs (centres in the inner
  60% of the box, sigma 2-6 voxels, amplitude U(0.5, 1)), float32 (N, N, N)
  indexed [z, y, x].
* :func:`random_directions` - uniformly distributed rotations from unit
  quaternions (Shoemake), returned as (P, 9) float32 rows [u; v; n] (the
  columns of the rotation matrix, the reference's direction format,
  utils/get_axes_from_n_rot.m:36-37).
"""
from __future__ import annotations

import numpy as np


def blob_params(n: int, seed: int, nblobs: int = 20):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.2 * n, 0.8 * n, size=(nblobs, 3))  # (x, y, z)
    sigmas = rng.uniform(2.0, 6.0, size=nblobs)
    amps = rng.uniform(0.5, 1.0, size=nblobs)
    return centres, sigmas, amps


def blob_volume(n: int, seed: int, nblobs: int = 20) -> np.ndarray:
    centres, sigmas, amps = blob_params(n, seed, nblobs)
    x = np.arange(n, dtype=np.float64)
    vol = np.zeros((n, n, n), dtype=np.float64)
    for (cx, cy, cz), s, a in zip(centres, sigmas, amps):
        gx = np.exp(-((x - cx) ** 2) / (2 * s * s))
        gy = np.exp(-((x - cy) ** 2) / (2 * s * s))
        gz = np.exp(-((x - cz) ** 2) / (2 * s * s))
        vol += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return vol.astype(np.float32)


def quaternion_to_columns(q: np.ndarray) -> np.ndarray:
    """(P, 4) unit quaternions (x, y, z, w) -> (P, 9) [R(:,1); R(:,2); R(:,3)]."""
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.empty((q.shape[0], 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z)
    r[:, 0, 1] = 2 * (x * y - z * w)
    r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w)
    r[:, 1, 1] = 1 - 2 * (x * x + z * z)
    r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w)
    r[:, 2, 1] = 2 * (y * z + x * w)
    r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    # columns stacked: element 3*c + r of the 9-vector is R[r, c]
    return np.ascontiguousarray(np.transpose(r, (0, 2, 1)).reshape(-1, 9), dtype=np.float32)


def random_directions(p: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.random((3, p))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    return quaternion_to_columns(q)
