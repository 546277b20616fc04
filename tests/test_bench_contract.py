"""CPU checks of the benchmark's definitions: the slab-model byte counts the
roofline divides by (the in-mask CAS pixel counts n_mask(M), computed from
the mask definition, against the table of exact counts for the five image
sizes of the workloads), and the one-GPU refusal of ``--gpus N``."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("m,count", [(128, 12453), (192, 28345), (256, 50617), (384, 114553), (512, 204269)])
def test_mask_pixel_counts(m, count):
    assert bench.n_mask(m) == count


def test_slab_model_bytes_at_n128():
    # 5 slab voxels per in-mask pixel (FP32 read; read-modify-write for the
    # back projection), the 256^2 CAS image and the 9 axis floats
    assert bench.bytes_fp(256) == 4 * 5 * 50617 + 4 * 256 * 256 + 36 == 1274520
    assert bench.bytes_fp(256, sv=2) == 2 * 5 * 50617 + 4 * 256 * 256 + 36
    assert bench.bytes_bp(256) == 8 * 5 * 50617 + 4 * 256 * 256 + 36 == 2286860


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()
