"""Spread of the projectors' results across arithmetic readings of the
reference kernels (CPU, oracle only).

The reference is compiled by nvcc with no flags (src/CMakeLists.txt:80-83), so
nvcc's default --fmad=true may contract the multiply-adds that feed rounding
decisions: the rotated coordinates (gpuForwardProject.cu:53-55,
gpuBackProject.cu:62-64), the tap positions (gpuBackProject.cu:77,84,91) and
the kernel-table index (gpuForwardProject.cu:67,76,84, gpuBackProject.cu:80,
87,94).  This script runs the oracle under each reading it restates on two
workloads and prints, against the pinned reading ("first"), the relative L2
and max-abs / peak differences of the forward-projected CAS images and the
back-projected CAS volume.

Workloads:
  C1: N=64 blob volume (seed 1), interp 2 (M=128, V=134), 500 random
      directions (seed 100); forward project, then back project those images.
  C2 subset: N=128 blob volume (seed 1), interp 2 (M=256, V=262), every 50th
      direction of the 100x100 uniform grid (200 directions).
Back projection runs twice: on the forward-projected images (smooth, the
workloads' own data) and on standard-normal images (seed 3), where the
result carries no smooth signal and decision changes show most.

Usage: python tests/reading_spread.py [c1] [c2]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from multigpugridder_amd.synth import blob_volume, random_directions  # noqa: E402
from oracle import oracle  # noqa: E402


def _diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    rel = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    mx = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    frac = float(np.count_nonzero(a != b) / a.size)
    return rel, mx, frac


def run(name, n, axes):
    ker = oracle.kb_table().astype(np.float32)
    m, v = 2 * n, 2 * n + 6
    cas = oracle.vol_to_cas(blob_volume(n, 1), 2, 3).astype(np.float32)
    res = {}
    t0 = time.time()
    with oracle.reading("first"):
        imgs = oracle.forward_project(cas, axes, ker, m).astype(np.float32)
    noise = np.random.default_rng(3).standard_normal(imgs.shape).astype(np.float32)
    for r in oracle.READINGS:
        with oracle.reading(r):
            fp = oracle.forward_project(cas, axes, ker, m)
            bp = oracle.back_project(imgs, axes, ker, v)
            bn = oracle.back_project(noise, axes, ker, v)
        res[r] = (fp, bp, bn)
    print(f"{name}: N={n} M={m} V={v} P={axes.shape[0]} ({time.time() - t0:.0f} s)")
    ref = res["first"]
    print(f"  {'reading':8s} | {'FP relL2':>9s} {'max/peak':>9s} {'differ%':>8s} | "
          f"{'BP relL2':>9s} {'max/peak':>9s} | {'BPnoise':>9s} {'max/peak':>9s}")
    for r, out in res.items():
        f = _diff(out[0], ref[0])
        b = _diff(out[1], ref[1])
        c = _diff(out[2], ref[2])
        print(f"  {r:8s} | {f[0]:9.2e} {f[1]:9.2e} {100 * f[2]:8.3f} | {b[0]:9.2e} {b[1]:9.2e} | "
              f"{c[0]:9.2e} {c[1]:9.2e}")
    sys.stdout.flush()


def main(argv):
    which = argv[1:] or ["c1", "c2"]
    if "c1" in which:
        run("C1", 64, random_directions(500, 100))
    if "c2" in which:
        run("C2 subset", 128, oracle.uniform_axes(100, 100)[::50].copy())


if __name__ == "__main__":
    main(sys.argv)
