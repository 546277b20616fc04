"""The back-projection kernel replaces the reference's table-index arithmetic
(r = fl(p - v), clamp, fl(fl(r*125) + 250), roundf; Projection/
gpuBackProject.cu:77-95) by two fused forms that each drop operations from
every tap.  They must pick the same table entry for every float tap position
p within 4.5 of the voxel coordinate v.  This runs `index_forms.c` over every
such float for voxel coordinates near each range boundary and at large
volumes (the kernel's own exhaustive run covered every v up to 2,060).

Form B (v >= 9) must agree everywhere; form A (v >= 130) must agree above
its bound and must NOT agree below it (a double rounding it skips matters
there), which shows the check can fail."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("idx") / "index_forms")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "index_forms.c"), "-lm"],
                   check=True)
    return exe


def run(checker, *ranges):
    out = subprocess.run([checker, *ranges], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.strip().splitlines():
        lo, hi, n, ma, mb = (int(x) for x in line.split())
        rows[(lo, hi)] = (n, ma, mb)
    return rows


def test_form_b_matches_reference_index(checker):
    rows = run(checker, "9:14", "120:140", "255:262", "515:520", "1027:1030")
    for (lo, hi), (n, _, mb) in rows.items():
        assert n > 0
        assert mb == 0, f"form B differs for v in [{lo}, {hi}]"


def test_form_a_matches_above_its_bound(checker):
    rows = run(checker, "130:140", "255:262", "515:520", "1027:1030")
    for (lo, hi), (n, ma, _) in rows.items():
        assert n > 0
        assert ma == 0, f"form A differs for v in [{lo}, {hi}]"


def test_form_a_fails_below_its_bound(checker):
    rows = run(checker, "100:129")
    n, ma, mb = rows[(100, 129)]
    assert ma > 0 and mb == 0
