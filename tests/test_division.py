"""The exact-division forms of the step kernels are the IEEE quotient bit for
bit on the range where the kernels use them, and are not outside it, which is
why the kernels fall back to the IEEE division there:

* reciprocal and correction (divisors rho, J): divisor in [2^-20, 2^20],
  numerator zero or 2^-100 <= |x| <= 2^100; misses for subnormal quotients;
* fma(x, zh, x * zl) (divisors 8 dx, 8 dz): divisors admitted by the host's
  check of one binade, numerator zero or T <= |x| <= 2^100 with the host's
  threshold T, signed zeros included; misses below T.

Checked on the CPU with C's fmaf, which rounds like the GPU's v_fma_f32."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("div") / "division_check")
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-fno-fast-math", os.path.join(HERE, "division_check.c"),
                    "-o", exe, "-lm"], check=True)
    return exe


def test_guarded_form_is_ieee_division_on_its_range(checker):
    r = subprocess.run([checker, "guarded", "7"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout


def test_unguarded_form_misses_subnormal_quotients(checker):
    r = subprocess.run([checker, "unguarded"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode > 0, "expected one-unit misses for subnormal quotients"


def test_constant_divisor_form_is_ieee_division_on_its_range(checker):
    r = subprocess.run([checker, "bmr", "7"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout


def test_constant_divisor_form_misses_below_its_threshold(checker):
    r = subprocess.run([checker, "bmr_below"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode > 0, "expected misses for numerators below the threshold"
