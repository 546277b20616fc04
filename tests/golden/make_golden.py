"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py
The fixtures are small (N=16: M=32, V=38, 8 directions) and pin the oracle
and the GPU path against regressions; the oracle itself is pinned by the
closed-form and identity tests in tests/test_oracle.py.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from multigpugridder_amd.synth import blob_volume, random_directions  # noqa: E402


def main():
    n, m, pad = 16, 32, 3
    v = m + 2 * pad
    ker = oracle.kb_table().astype(np.float32)
    vol = blob_volume(n, 1)
    cas_vol = oracle.vol_to_cas(vol, 2, pad).astype(np.float32)
    axes = random_directions(8, 100)
    fp = oracle.forward_project(cas_vol, axes, ker, m)
    rng = np.random.default_rng(42)
    cas_imgs = rng.standard_normal((8, m, m)).astype(np.float32)
    bp = oracle.back_project(cas_imgs, axes, ker, v)
    adj = oracle.back_project_adjoint(cas_imgs, axes, ker, v)
    imgs = oracle.cas_imgs_to_imgs(fp, n)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "projector_n16.npz")
    np.savez_compressed(out, ker=ker, vol=vol, cas_vol=cas_vol, axes=axes, fp=fp, cas_imgs=cas_imgs, bp=bp,
                        adj=adj, imgs=imgs, n=n, m=m, v=v, pad=pad)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
