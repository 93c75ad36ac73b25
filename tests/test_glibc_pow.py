"""The device pow used by the characteristic pressure boundary state
(cffc_amd/csrc/glibc_pow.h) against this host's glibc pow(), bit for bit.

The reference evaluates rho_b = rho_i * pow(p_b/p_i, 1/gamma)
(src_2D/Euler2D/Euler2DState.cc:1432) with the C library's pow; the oracle
does too.  glibc's pow is not correctly rounded (about 0.52 ulp), so only the
same method with the same roundings gives the same bits.  A replica whose
final s + s*tmp is left unfused (one rounding moved) differs from libm on
about 7 in 10,000 of these inputs, so the samples below would see a wrong
step.  The GPU test (test_gpu_parity.py) runs the same header on the device.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "glibc_pow_check.c")


def _has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
    except OSError:
        return False
    return " fma " in flags and " avx2 " in flags


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not _has_fma():
        pytest.skip("glibc selects its FMA pow only on CPUs with FMA and AVX2")
    exe = str(tmp_path_factory.mktemp("pow") / "glibc_pow_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-mfma", "-o", exe, SRC, "-lm"], check=True)
    return exe


def _run(exe, *args):
    out = subprocess.run([exe] + [str(a) for a in args], check=True, capture_output=True, text=True).stdout
    n, bad = (int(v) for v in out.strip().splitlines()[-1].split())
    return n, bad, out


def test_boundary_exponent_matches_libm(checker):
    # pressure ratios of the characteristic outflow, exponent 1/gamma = 1/1.4
    n, bad, out = _run(checker, 4_000_000, 0.3, 3.0, repr(1.0 / 1.4))
    assert n == 4_000_000 and bad == 0, out


def test_random_exponents_match_libm(checker):
    n, bad, out = _run(checker, 4_000_000, 0.01, 100.0)
    assert bad == 0, out


def test_tables_follow_their_construction_rule():
    """The committed tables are what the generator derives from the rule."""
    pytest.importorskip("mpmath")
    sys.path.insert(0, os.path.join(ROOT, "cffc_amd", "csrc"))
    try:
        import gen_glibc_pow_tables as g
    finally:
        sys.path.pop(0)
    with open(os.path.join(ROOT, "cffc_amd", "csrc", "glibc_pow_tables.h")) as f:
        assert f.read() == g.render()
    # spot values any reader can check: 2^(1/128) split as H(1+T), and the
    # first sub-interval's 1/c = 181/128
    ex = g.exp_table()
    assert ex[0] == (0, 0x3FF0000000000000)
    assert ex[1] == (0x3C9B3B4F1A88BF6E, 0x3FEFF63DA9FB3335)
    lg = g.log_table()
    assert lg[0][0] == 181 / 128 and lg[74][:2] == (1.0, 0.0) and lg[75][:2] == (1.0, 0.0)
