"""The FP64 division, square root, minimum and maximum of the stage kernels,
checked on the device against the compiler's own operations and the host's
IEEE arithmetic.

The stage kernels divide with gfx950's correctly rounded sequence written
out (v_rcp_f64, two Newton steps, the quotient, its FMA remainder and
correction, v_div_fixup_f64) without the operand scaling of
v_div_scale_f64, so that quotients over one divisor can share the Newton
steps; the square root is the compiler's sequence without its input
scaling; the limiter's extrema are single v_min_f64/v_max_f64.  The scaling
acts only when a/b or 1/b is subnormal, |a/b| >= 2^767 or 0 < |a| < 2^-969
(x < 2^-767 for the square root): inside that domain the results must be
the compiler's bit for bit, and the compiler's are IEEE round-to-nearest,
which numpy reproduces on the host.  Zeros of either sign, infinities and
NaNs go through v_div_fixup_f64 as in the compiler's sequence.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def euler2d():
    from cffc_amd import euler2d as E
    E.load()
    return E


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def in_domain_pairs(rng, n):
    """Operand pairs inside the exactness domain: random signs, mantissas and
    exponents (numerator 2^-960 .. 2^700, divisor 2^-700 .. 2^700, quotient
    2^-1000 .. 2^760), plus the mantissa edge cases 1 and 2 - ulp."""
    ea = rng.integers(-960, 700, n)
    eb = rng.integers(-700, 700, n)
    d = ea - eb
    keep = (d > -1000) & (d < 760)
    ea, eb = ea[keep], eb[keep]
    m = ea.size
    ma = rng.uniform(1.0, 2.0, m)
    mb = rng.uniform(1.0, 2.0, m)
    edge = rng.integers(0, 8, m)
    ma = np.where(edge == 0, 1.0, np.where(edge == 1, np.nextafter(2.0, 0.0), ma))
    mb = np.where(edge == 2, 1.0, np.where(edge == 3, np.nextafter(2.0, 0.0), mb))
    sa = np.where(rng.random(m) < 0.5, -1.0, 1.0)
    sb = np.where(rng.random(m) < 0.5, -1.0, 1.0)
    a = sa * np.ldexp(ma, ea)
    b = sb * np.ldexp(mb, eb)
    return a, b


def flow_pairs(rng, n):
    """Magnitudes of the quantities the kernels divide: densities, pressures,
    sound speeds, areas, face-length sums and differences of nearby values."""
    a = rng.lognormal(0.0, 3.0, n) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    b = rng.lognormal(0.0, 3.0, n) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    close = rng.random(n) < 0.25
    b = np.where(close, a * (1.0 + rng.normal(0.0, 1e-12, n)), b)
    a = np.where(rng.random(n) < 0.1, np.nextafter(b, np.inf) - b, a)
    return a, b


def test_division_and_sqrt_equal_the_compilers_in_their_domain(euler2d):
    rng = np.random.default_rng(20261203)
    a1, b1 = in_domain_pairs(rng, 400000)
    a2, b2 = flow_pairs(rng, 200000)
    a = np.concatenate([a1, a2])
    b = np.concatenate([b1, b2])
    print("in-domain pairs:", a.size, flush=True)
    out = euler2d.check_arith(a, b)
    with np.errstate(all="ignore"):
        host = a / b
    assert np.all(np.isfinite(host))
    assert np.array_equal(bits(out[:, 1]), bits(host)), "compiler's a / b is not IEEE round-to-nearest"
    nq = int(np.sum(bits(out[:, 0]) != bits(out[:, 1])))
    assert nq == 0, "%d quotients differ from a / b" % nq
    # square root of |a| where |a| >= 2^-767
    x = np.abs(a)
    ok = x >= 2.0 ** -767
    assert np.array_equal(bits(out[ok, 3]), bits(np.sqrt(x[ok])))
    ns = int(np.sum(bits(out[ok, 2]) != bits(out[ok, 3])))
    assert ns == 0, "%d square roots differ from sqrt" % ns
    # exact squares and the mantissa edges
    sq = np.concatenate([np.arange(1.0, 2001.0) ** 2, np.ldexp(1.0, np.arange(-766, 1000, 2)),
                         np.nextafter(np.ldexp(1.0, np.arange(-700, 700, 7)), 0.0)])
    o2 = euler2d.check_arith(sq, np.ones_like(sq))
    assert np.array_equal(bits(o2[:, 2]), bits(np.sqrt(sq)))
    assert np.array_equal(bits(o2[:, 2]), bits(o2[:, 3]))


def test_division_special_operands_follow_ieee(euler2d):
    """Zeros of either sign, infinities and NaN go through v_div_fixup_f64
    as in the compiler's sequence: the results equal the compiler's and
    IEEE's (NaN compared as NaN)."""
    vals = np.array([0.0, -0.0, 1.0, -1.0, 3.5, -2.25e-7, 7.0e10, np.inf, -np.inf, np.nan])
    a = np.repeat(vals, vals.size)
    b = np.tile(vals, vals.size)
    out = euler2d.check_arith(a, b)
    with np.errstate(all="ignore"):
        host = a / b
    nan = np.isnan(host)
    assert np.array_equal(np.isnan(out[:, 0]), nan) and np.array_equal(np.isnan(out[:, 1]), nan)
    assert np.array_equal(bits(out[~nan, 0]), bits(host[~nan]))
    assert np.array_equal(bits(out[~nan, 1]), bits(host[~nan]))
    # square root of |a|: zeros keep their value, +inf stays +inf
    x = np.abs(a)
    sq = np.sqrt(x)
    nn = ~np.isnan(sq)
    assert np.array_equal(bits(out[nn, 2]), bits(sq[nn]))
    assert np.all(np.isnan(out[~nn, 2]))


def test_single_instruction_extrema_equal_fmin_fmax(euler2d):
    rng = np.random.default_rng(7)
    a = rng.normal(0.0, 1.0, 100000)
    b = np.where(rng.random(a.size) < 0.2, a, rng.normal(0.0, 1.0, a.size))
    z = np.array([0.0, -0.0, 0.0, -0.0])
    a = np.concatenate([a, z])
    b = np.concatenate([b, z[[0, 0, 1, 1]]])
    out = euler2d.check_arith(a, b)
    assert np.array_equal(bits(out[:, 4]), bits(out[:, 5]))
    assert np.array_equal(bits(out[:, 6]), bits(out[:, 7]))
    # values (zero signs aside) are the ordinary minimum and maximum
    assert np.array_equal(out[:, 4], np.minimum(a, b))
    assert np.array_equal(out[:, 6], np.maximum(a, b))


def test_outside_the_domain_only_the_scaled_cases_may_differ(euler2d):
    """Documented limit, not a requirement of the solver: tiny numerators
    and subnormal quotients are where the unscaled sequence may differ from
    a / b.  Every differing pair must lie there."""
    rng = np.random.default_rng(11)
    n = 100000
    ea = rng.integers(-1070, -900, n)
    eb = rng.integers(-60, 60, n)
    a = np.ldexp(rng.uniform(1.0, 2.0, n), ea) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    b = np.ldexp(rng.uniform(1.0, 2.0, n), eb)
    out = euler2d.check_arith(a, b)
    diff = bits(out[:, 0]) != bits(out[:, 1])
    print("pairs differing below the domain: %d of %d" % (int(diff.sum()), n))
    with np.errstate(all="ignore"):
        q = np.abs(a / b)
    outside = (np.abs(a) < 2.0 ** -969) | (q < 2.0 ** -1022)
    assert not np.any(diff & ~outside)
