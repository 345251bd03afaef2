"""The anisotropic-kernel walks divide r by h with one fma correction of r * (1/h) instead of an IEEE division
(div_h, fltk_sph_ice_amd/csrc/aniso.hip).  Their results are bit-exact against the CPU path only if that quotient is
correctly rounded: tests/cpp/div_check.c compares it bit for bit with r / h over every float r in [h/4, 2h] and a
stride through [2^-75, h/4), for radii of the scenes and radii with extreme significands, and shows that the
uncorrected product fails the same check."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RADII = ["0.2", "0.1", "0.04", "0.02", "0.0123", "0.3333333", "0.99999994", "1.99999988", "7.77"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("gcc not found")
    exe = str(tmp_path_factory.mktemp("div") / "div_check")
    r = subprocess.run([cc, "-O2", "-ffp-contract=off", "-o", exe, os.path.join(ROOT, "tests", "cpp", "div_check.c"), "-lm"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, mode, radii):
    r = subprocess.run([exe, str(mode), *radii], capture_output=True, text=True, timeout=300)
    return r.returncode, {float(a): int(b) for a, b in (line.split() for line in r.stdout.split("\n") if line)}


def test_one_correction_quotient_is_correctly_rounded(checker):
    rng = np.random.default_rng(4)
    radii = RADII + [repr(float(np.float32(v))) for v in 10.0 ** rng.uniform(-3, 1, 6)]
    rc, bad = _run(checker, 1, radii)
    assert len(bad) == len(radii)
    assert rc == 0 and not any(bad.values()), bad


def test_uncorrected_product_fails_the_check(checker):
    rc, bad = _run(checker, 0, ["0.2", "0.0123"])
    assert rc == 1 and all(v > 0 for v in bad.values()), bad
