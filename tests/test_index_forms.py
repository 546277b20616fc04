"""The projection kernels replace the reference's table-index arithmetic, as
nvcc contracts it (r = fl(p - v), clamp to [-2, 2], roundf(fma(r, 125, 250));
Projection/gpuBackProject.cu:78-81, gpuForwardProject.cu:65-68), by
floor(s + 0.5) (one v_cvt_rpi_i32_f32) on an unclamped r with a saturated
index (form N), and the back projector by a single FMA of the tap position
itself (form A, s = fma(p, 125, 250 - 125 v), for voxel coordinates v >= 3).
Both must pick the same table entry as the reference.  This runs
`index_forms.c`: form N over every float r in [-4.5, 4.5], form A over every
float tap position p within 4.5 of v for coordinates around its bound and at
large volumes, each with floor(s + 0.5) exact and with s + 0.5 rounded to
FP32.  Form A must also fail at v = 1 (p - v is inexact for p < 1), which
shows the check can fail."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("idx") / "index_forms")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-march=x86-64-v3", "-o", exe,
                    os.path.join(HERE, "index_forms.c"), "-lm"], check=True)
    return exe


def run(checker, *ranges):
    out = subprocess.run([checker, *ranges], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.strip().splitlines():
        f = line.split()
        if f[0] == "r":
            rows["r"] = tuple(int(x) for x in f[1:])
        else:
            lo, hi, n, me, mf = (int(x) for x in f)
            rows[(lo, hi)] = (n, me, mf)
    return rows


def test_form_n_matches_reference_index_for_every_offset(checker):
    n, me, mf = run(checker, "r")["r"]
    assert n > 2_000_000_000
    assert me == 0 and mf == 0


def test_form_a_matches_reference_index(checker):
    rows = run(checker, "4:14", "120:140", "255:262", "515:520", "1027:1030", "2050:2060")
    for (lo, hi), (n, me, mf) in rows.items():
        assert n > 0
        assert me == 0 and mf == 0, f"form A differs for v in [{lo}, {hi}]"


def test_form_a_fails_below_its_bound(checker):
    n, me, mf = run(checker, "1:1")[(1, 1)]
    assert me > 0 and mf > 0
