"""The oracle's state/CFD functions against the reference's own outputs.

tests/golden/state_kat.txt holds inputs and outputs of the reference's
Euler2DState.cc / CFD.cc compiled from source (tests/golden/make_state_kat.py).
Every value must match bit for bit.
"""
import os

import numpy as np
import pytest

import oracle

KAT = os.path.join(os.path.dirname(__file__), "golden", "state_kat.txt")


def _load():
    rows = {}
    with open(KAT) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            parts = line.split()
            rows.setdefault(parts[0], []).append([float.fromhex(x) for x in parts[1:]])
    return {k: np.array(v) for k, v in rows.items()}


ROWS = _load()


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _check(tag, nin, fn):
    rows = ROWS[tag]
    assert len(rows) > 0
    bad = 0
    for r in rows:
        got = np.atleast_1d(fn(r[:nin]))
        exp = r[nin:]
        if not np.array_equal(_bits(got), _bits(exp)):
            bad += 1
    assert bad == 0, "%s: %d of %d vectors differ from the reference" % (tag, bad, len(rows))


def test_kat_flux_roe_n(oracle_lib):
    _check("ROE_N", 10, lambda v: oracle.flux_roe_n(v[0:4], v[4:8], v[8], v[9]))


def test_kat_roe_average(oracle_lib):
    _check("ROE_AVG", 8, lambda v: oracle.roe_average(v[0:4], v[4:8]))


def test_kat_venkatakrishnan(oracle_lib):
    _check("VENKAT", 7, lambda v: oracle.limiter_venkatakrishnan(v[0:4], v[4], v[5], v[6]))


def test_kat_harten(oracle_lib):
    _check("HARTEN", 3, lambda v: [oracle.harten_fix_pos(*v), oracle.harten_fix_neg(*v)])


def test_kat_reflect(oracle_lib):
    _check("REFLECT", 6, lambda v: oracle.reflect(v[0:4], v[4], v[5]))


def test_kat_bc_characteristic_pressure(oracle_lib):
    _check("BCCHARP", 10, lambda v: oracle.bc_characteristic_pressure(v[0:4], v[4:8], v[8], v[9]))


def test_kat_conversions(oracle_lib):
    _check("U_OF_W", 4, oracle.U_of_W)
    _check("W_OF_U", 4, oracle.W_of_U)
    _check("FX", 4, oracle.Fx)


def test_kat_stage_coefficients(oracle_lib):
    _check("RK", 2, lambda v: oracle.runge_kutta(int(v[0]), int(v[1])))
    _check("OPTSM", 3, lambda v: oracle.multistage_optimal_smoothing(int(v[0]), int(v[1]), int(v[2])))


def test_kat_covers_branches():
    """The vectors exercise the branches a plausible mistake would hide in."""
    v = ROWS["VENKAT"]
    d = v[:, 0:4] - v[:, 4:5]
    assert (d > 1e-7).any() and (d < -1e-7).any() and (np.abs(d) <= 1e-7).any()
    assert (v[:, 7] < 1).sum() > 50 and (v[:, 7] == 1).sum() > 10
    h = ROWS["HARTEN"]
    # entropy fix active (|la| <= delta and delta >= TOLER) in some rows
    delta = np.maximum(0, 2 * (h[:, 2] - h[:, 1]))
    assert ((np.abs(h[:, 0]) <= delta) & (delta >= 1e-7)).sum() > 20
    b = ROWS["BCCHARP"]
    assert b.shape[0] >= 400


# --- closed forms and invariants the mathematics fixes ------------------------

def test_roe_consistency_equals_physical_flux(oracle_lib):
    """F_Roe(W, W, n) = Fx rotated (test_Euler2DState.cc:80 analogue)."""
    rng = np.random.default_rng(1)
    for _ in range(100):
        W = np.array([rng.uniform(0.1, 2), rng.normal(), rng.normal(), rng.uniform(0.1, 2)])
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        F = oracle.flux_roe_n(W, W, nx, ny)
        un = W[1] * nx + W[2] * ny
        E = W[3] / 0.4 + 0.5 * W[0] * (W[1] ** 2 + W[2] ** 2)
        Fe = np.array([W[0] * un, W[0] * W[1] * un + W[3] * nx, W[0] * W[2] * un + W[3] * ny, un * (E + W[3])])
        np.testing.assert_allclose(F, Fe, rtol=1e-12, atol=1e-12)


def test_roe_supersonic_upwind(oracle_lib):
    """All eigenvalues positive (un > a on both sides): F = F(W_L)."""
    Wl = np.array([1.0, 3.0, 0.2, 1.0])
    Wr = np.array([0.8, 2.9, -0.1, 0.7])
    F = oracle.flux_roe_n(Wl, Wr, 1.0, 0.0)
    np.testing.assert_array_equal(F, oracle.Fx(Wl))


def test_roe_rotational_invariance(oracle_lib):
    rng = np.random.default_rng(2)
    for _ in range(50):
        Wl = np.array([rng.uniform(0.2, 2), rng.normal(), rng.normal(), rng.uniform(0.2, 2)])
        Wr = np.array([rng.uniform(0.2, 2), rng.normal(), rng.normal(), rng.uniform(0.2, 2)])
        th = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(th), np.sin(th)
        rot = lambda W: np.array([W[0], W[1] * c + W[2] * s, -W[1] * s + W[2] * c, W[3]])
        Fr = oracle.flux_roe_x(rot(Wl), rot(Wr))
        F = oracle.flux_roe_n(Wl, Wr, c, s)
        back = np.array([Fr[0], Fr[1] * c - Fr[2] * s, Fr[1] * s + Fr[2] * c, Fr[3]])
        np.testing.assert_array_equal(F, back)


def test_venkatakrishnan_closed_forms(oracle_lib):
    f = lambda y: (y * y + 2 * y) / (y * y + y + 2)
    # y = 1 -> 3/4, y = 2 -> 1, y >= 2 -> capped at 1, dead band -> 1
    assert oracle.limiter_venkatakrishnan([1.0, 0, 0, 0], 0.0, -1.0, 1.0) == 0.75
    assert oracle.limiter_venkatakrishnan([0.5, 0, 0, 0], 0.0, -1.0, 1.0) == 1.0
    assert oracle.limiter_venkatakrishnan([0.25, 0, 0, 0], 0.0, -1.0, 1.0) == 1.0
    assert oracle.limiter_venkatakrishnan([5e-8, -5e-8, 0, 0], 0.0, -1.0, 1.0) == 1.0
    assert oracle.limiter_venkatakrishnan([0.4, -0.8, 0, 0], 0.0, -0.2, 1.0) == f(0.25)


def test_harten_reduces_to_split_eigenvalues(oracle_lib):
    """delta below TOLER: the fix returns max(la,0) / min(la,0) exactly."""
    for la in (-2.0, -0.3, 0.0, 0.7):
        assert oracle.harten_fix_pos(la, 1.0, 1.0) == max(la, 0.0)
        assert oracle.harten_fix_neg(la, 1.0, 1.0) == min(la, 0.0)


def test_state_round_trip(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(100):
        W = np.array([rng.uniform(0.1, 2), rng.normal(), rng.normal(), rng.uniform(0.1, 2)])
        np.testing.assert_allclose(oracle.W_of_U(oracle.U_of_W(W)), W, rtol=1e-14, atol=1e-14)


def test_reflect_involution(oracle_lib):
    W = np.array([1.2, 0.3, -0.7, 0.9])
    th = 0.37
    R = oracle.reflect(W, np.cos(th), np.sin(th))
    np.testing.assert_allclose(oracle.reflect(R, np.cos(th), np.sin(th)), W, rtol=1e-15, atol=1e-15)
    un = W[1] * np.cos(th) + W[2] * np.sin(th)
    unr = R[1] * np.cos(th) + R[2] * np.sin(th)
    assert abs(un + unr) < 1e-15


def test_rk4_coefficients(oracle_lib):
    assert [oracle.runge_kutta(i, 4) for i in (1, 2, 3, 4)] == [0.5, 0.5, 1.0, 1.0 / 6.0]


def test_quad_centroid_closed_form(oracle_lib):
    # unit square shifted: centroid at its centre, area 1; a trapezoid
    np.testing.assert_allclose(oracle.quad_centroid([2, 3, 3, 2], [5, 5, 6, 6]), [2.5, 5.5, 1.0])
    # trapezoid with parallel sides 2 (bottom) and 1 (top), height 1
    x, y, A = oracle.quad_centroid([0, 2, 1.5, 0.5], [0, 0, 1, 1])
    assert abs(A - 1.5) < 1e-15 and abs(x - 1.0) < 1e-15 and abs(y - 4.0 / 9.0) < 1e-15


def test_kat_flux_hlle_n(oracle_lib):
    _check("HLLE_N", 10, lambda v: oracle.flux_hlle_n(v[0:4], v[4:8], v[8], v[9]))


def test_kat_axisymmetric_source(oracle_lib):
    _check("AXISRC", 6, lambda v: oracle.axisym_source(v[0:4], v[4], v[5]))


def _bilin_scalar(v):
    U = np.zeros((4, 4))
    U[:, 0] = v[0:4]
    e, out = oracle.bilinear(U, np.asarray(v[4:12]).reshape(4, 2), v[12], v[13])
    return [float(e), 0.0 if e else out[0]]


def test_kat_bilinear_interpolation(oracle_lib):
    _check("BILIN", 14, _bilin_scalar)


def test_hlle_consistency_and_upwinding(oracle_lib):
    """F_HLLE(W, W) = Fn(W); supersonic to the right (left) gives the left
    (right) state's normal flux exactly."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        W = np.array([rng.uniform(0.2, 2), rng.normal(), rng.normal(), rng.uniform(0.2, 2)])
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        F = oracle.flux_hlle_n(W, W, nx, ny)
        un = W[1] * nx + W[2] * ny
        H = 1.4 / 0.4 * W[3] + 0.5 * W[0] * (W[1] ** 2 + W[2] ** 2)
        ref = [W[0] * un, W[0] * un * W[1] + W[3] * nx, W[0] * un * W[2] + W[3] * ny, un * H]
        np.testing.assert_allclose(F, ref, rtol=1e-13, atol=1e-13)
    Wl = np.array([1.0, 3.0, 0.2, 1.0])
    Wr = np.array([0.5, 2.5, -0.1, 0.4])
    np.testing.assert_array_equal(oracle.flux_hlle_n(Wl, Wr, 1.0, 0.0), oracle.Fx(Wl))
