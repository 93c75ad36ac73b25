"""The oracle's single-level block functions, pinned by what the mathematics
fixes (CPU).

The block-level files of the reference (Euler2DQuad.cc,
Euler2DQuadSingleBlock.cc, the message passing) cannot be built here, so no
reference binary checks these functions.  Each test below states a
property that holds whatever the implementation, chosen so that a plausible
mistake (a wrong index or sign, a transposed operand, a dropped face term)
breaks it:
  * a5 least squares: every interior cell's gradient is the least-squares
    fit over its 3 x 3 window of cells, ghosts included (numpy's lstsq on
    the same centroids and states), and is exact for linear fields;
  * a6 on a uniform mesh a linear field is not limited (phi = 1);
  * a4 CFL: on a uniform mesh dt = min(dx/(a+|u|), dy/(a+|v|)) per cell;
  * a11 messages: every ghost cell next to a neighbour block (faces and
    corners) holds that neighbour's interior cell;
  * a10 BCs: reflecting, fixed and extrapolating ghosts by their definition
    on a straight wall;
  * a14 norms: L1, L2 and max of the chosen variable of dUdt;
  * a7 / a13 / a15: mass and energy are conserved in a closed box with
    reflecting walls on a disturbed mesh, free stream is preserved with
    fixed, extrapolating and characteristic boundaries, and the 4-quadrant
    Riemann problem keeps its diagonal mirror symmetry (x, y, u, v) ->
    (y, x, v, u).
"""
import numpy as np
import pytest

import oracle
from cffc_amd import abi
from cffc_amd import mesh as M

NG = 2


def _smooth_W(c, seed=0):
    rng = np.random.default_rng(seed)
    k = rng.uniform(1.0, 3.0, size=(4, 2))
    x, y = c[:, 0], c[:, 1]
    return np.stack([1.0 + 0.2 * np.sin(k[0, 0] * x + k[0, 1] * y),
                     0.3 + 0.2 * np.cos(k[1, 0] * x - k[1, 1] * y),
                     -0.1 + 0.2 * np.sin(k[2, 0] * x * y + k[2, 1]),
                     1.0 + 0.3 * np.cos(k[3, 0] * x + k[3, 1] * y * y)], axis=1)


GRAD = np.array([[0.3, -0.2], [-0.1, 0.25], [0.15, 0.05], [0.2, 0.1]])   # d/dx, d/dy of (rho, u, v, p)


def _linear_W(c):
    return np.array([1.0, 0.4, -0.2, 1.0]) + c[:, 0:1] * GRAD[:, 0] + c[:, 1:2] * GRAD[:, 1]


def _interior(a, m, b):
    ni, nj = int(m.ni[b]), int(m.nj[b])
    return a[NG:nj + NG, NG:ni + NG]


def _block_slices(m):
    off = np.concatenate([[0], np.cumsum(m.ni * m.nj)])
    return [slice(off[b], off[b + 1]) for b in range(m.nblk)]


def _centroids(o, m):
    """Interior cell centroids in the oracle's cell order."""
    g = [o.block_geometry(b) for b in range(m.nblk)]
    return np.concatenate([np.stack([_interior(x, m, b).ravel(), _interior(y, m, b).ravel()], axis=1)
                           for b, (x, y, _) in enumerate(g)])


def _stage1(m, W, scheme=None):
    o = oracle.Oracle(m, scheme or abi.Scheme())
    o.set_W(W)
    o.exchange_and_bcs()
    o.cfl()
    assert o.dudt_stage(1) == 0
    return o


def test_lsq_gradient_is_the_least_squares_fit_of_the_window():
    bcs = dict(W=abi.BC_REFLECTION, E=abi.BC_CONSTANT_EXTRAPOLATION, S=abi.BC_FIXED, N=abi.BC_REFLECTION)
    m = M.cartesian(3, 2, 6, 6, x1=1.5, bcs=bcs, jitter=0.15, seed=5)
    m.wo[:] = M.freestream(0.5)
    c = m.cell_centres()
    o = _stage1(m, _smooth_W(c))
    gx, gy = o.get_recon(0), o.get_recon(1)
    worst = 0.0
    for b, sl in enumerate(_block_slices(m)):
        Wp = o.block_W_padded(b)
        xc, yc, _ = o.block_geometry(b)
        ni, nj = int(m.ni[b]), int(m.nj[b])
        k = 0
        for j in range(NG, nj + NG):
            for i in range(NG, ni + NG):
                nb = [(j + dj, i + di) for dj in (-1, 0, 1) for di in (-1, 0, 1) if (di, dj) != (0, 0)]
                A = np.array([[xc[q] - xc[j, i], yc[q] - yc[j, i]] for q in nb])
                rhs = np.array([Wp[q] - Wp[j, i] for q in nb])
                g, *_ = np.linalg.lstsq(A, rhs, rcond=None)
                got = np.stack([gx[sl][k], gy[sl][k]])
                worst = max(worst, float(np.max(np.abs(got - g) / (1.0 + np.abs(g)))))
                k += 1
    assert worst < 1e-10, worst


def test_lsq_exact_and_unlimited_for_linear_fields_on_a_uniform_mesh():
    """Away from the physical boundary (block interfaces included) a linear
    field's gradient is exact; on a uniform mesh the face-midpoint values
    lie halfway to the neighbours, so Venkatakrishnan's y >= 2 and phi = 1."""
    for jitter in (0.15, 0.0):
        m = M.cartesian(3, 3, 6, 6, jitter=jitter, seed=7)
        o = oracle.Oracle(m, abi.Scheme())
        c = _centroids(o, m)                 # area centroids, not node averages
        o.set_W(_linear_W(c))
        o.exchange_and_bcs()
        o.cfl()
        assert o.dudt_stage(1) == 0
        h = 1.0 / 18
        away = (c[:, 0] > h) & (c[:, 0] < 1 - h) & (c[:, 1] > h) & (c[:, 1] < 1 - h)
        np.testing.assert_allclose(o.get_recon(0)[away], np.broadcast_to(GRAD[:, 0], (away.sum(), 4)), atol=1e-11)
        np.testing.assert_allclose(o.get_recon(1)[away], np.broadcast_to(GRAD[:, 1], (away.sum(), 4)), atol=1e-11)
        if jitter == 0.0:
            np.testing.assert_allclose(o.get_recon(2)[away], 1.0, rtol=0, atol=1e-12)


def test_cfl_closed_form_on_a_uniform_mesh():
    m = M.cartesian(2, 2, 8, 4, x1=2.0)           # dx = 1/8, dy = 1/8
    W = _smooth_W(m.cell_centres(), seed=3)
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(W)
    dtmin = o.cfl(set_global=False)
    a = np.sqrt(1.4 * W[:, 3] / W[:, 0])
    ref = np.minimum(0.125 / (a + np.abs(W[:, 1])), 0.125 / (a + np.abs(W[:, 2])))
    np.testing.assert_allclose(o.get_dt(), ref, rtol=1e-14)
    assert dtmin == pytest.approx(ref.min(), rel=1e-14)


def test_messages_fill_ghosts_with_the_neighbours_interior():
    n = 4
    m = M.cartesian(3, 3, n, n, jitter=0.1, seed=2)
    W = np.random.default_rng(1).uniform(0.5, 1.5, size=(m.ncell, 4))
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(W)
    o.exchange_and_bcs()
    P = [o.block_W_padded(b) for b in range(m.nblk)]
    e = n + NG
    # (ghost rows, ghost cols) of the block <- (rows, cols) of the neighbour
    regions = {abi.NBR_E: ((slice(2, e), slice(e, e + 2)), (slice(2, e), slice(2, 4))),
               abi.NBR_W: ((slice(2, e), slice(0, 2)), (slice(2, e), slice(n, e))),
               abi.NBR_N: ((slice(e, e + 2), slice(2, e)), (slice(2, 4), slice(2, e))),
               abi.NBR_S: ((slice(0, 2), slice(2, e)), (slice(n, e), slice(2, e))),
               abi.NBR_NE: ((slice(e, e + 2), slice(e, e + 2)), (slice(2, 4), slice(2, 4))),
               abi.NBR_NW: ((slice(e, e + 2), slice(0, 2)), (slice(2, 4), slice(n, e))),
               abi.NBR_SE: ((slice(0, 2), slice(e, e + 2)), (slice(n, e), slice(2, 4))),
               abi.NBR_SW: ((slice(0, 2), slice(0, 2)), (slice(n, e), slice(n, e)))}
    checked = 0
    for b in range(m.nblk):
        for d, (mine, theirs) in regions.items():
            nb = int(m.nbr[b, d])
            if nb < 0:
                continue
            np.testing.assert_allclose(P[b][mine], P[nb][theirs], rtol=1e-14, atol=0, err_msg="block %d dir %d" % (b, d))
            checked += 1
    assert checked == 24 + 16   # directed face and corner neighbours of a 3 x 3 layout


def test_boundary_ghosts_by_definition_on_straight_walls():
    bcs = dict(W=abi.BC_REFLECTION, E=abi.BC_FIXED, S=abi.BC_CONSTANT_EXTRAPOLATION, N=abi.BC_REFLECTION)
    m = M.cartesian(1, 1, 5, 4, bcs=bcs)
    wo = np.array([1.2, 0.3, -0.1, 0.9])
    m.wo[:] = wo
    W = np.random.default_rng(4).uniform(0.5, 1.5, size=(m.ncell, 4))
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(W)
    o.exchange_and_bcs()
    P = o.block_W_padded(0)
    ni, nj = 5, 4
    rows, cols = slice(2, nj + 2), slice(2, ni + 2)
    flip_u, flip_v = np.array([1, -1, 1, 1]), np.array([1, 1, -1, 1])
    # west wall (normal -x): ghost k mirrors interior k, u reversed
    np.testing.assert_allclose(P[rows, 1], P[rows, 2] * flip_u, rtol=1e-15)
    np.testing.assert_allclose(P[rows, 0], P[rows, 3] * flip_u, rtol=1e-15)
    # east: the fixed reference state in both layers
    np.testing.assert_allclose(P[rows, ni + 2], np.broadcast_to(wo, (nj, 4)), rtol=1e-15)
    np.testing.assert_allclose(P[rows, ni + 3], np.broadcast_to(wo, (nj, 4)), rtol=1e-15)
    # south: both layers copy the boundary cell
    np.testing.assert_allclose(P[0, cols], P[2, cols], rtol=1e-15)
    np.testing.assert_allclose(P[1, cols], P[2, cols], rtol=1e-15)
    # north wall (normal +y): v reversed
    np.testing.assert_allclose(P[nj + 2, cols], P[nj + 1, cols] * flip_v, rtol=1e-15)
    np.testing.assert_allclose(P[nj + 3, cols], P[nj, cols] * flip_v, rtol=1e-15)


def test_periodic_ghosts_by_definition():
    """BC_PERIODIC wraps the block onto itself (SingleBlock.cc:2421-2425,
    2579-2583, 2734-2738, 2891-2895): W ghost g takes cell ICu-g+1, E ghost g
    cell ICl+g-1, and the S/N loop, which runs over whole rows after the W/E
    one (:2362, :2678), wraps rows, so the corners come out doubly periodic."""
    P_ = abi.BC_PERIODIC
    m = M.cartesian(1, 1, 5, 4, bcs=dict(W=P_, E=P_, S=P_, N=P_))
    W = np.random.default_rng(12).uniform(0.5, 1.5, size=(m.ncell, 4))
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(W)
    o.exchange_and_bcs()
    P = o.block_W_padded(0)
    ni, nj = 5, 4
    inner = P[2:nj + 2, 2:ni + 2]
    # the padded block equals the interior tiled periodically
    wrap = np.take(np.take(inner, (np.arange(ni + 4) - 2) % ni, axis=1), (np.arange(nj + 4) - 2) % nj, axis=0)
    np.testing.assert_array_equal(P, wrap)


def test_residual_norms_closed_form():
    m = M.cartesian(2, 2, 6, 6, jitter=0.1, seed=8)
    o = _stage1(m, _smooth_W(m.cell_centres(), seed=6))
    d = o.get_dUdt(0)
    for var in (1, 2, 3, 4):
        v = d[:, var - 1]
        np.testing.assert_allclose(o.residual_norms(var), [np.abs(v).sum(), np.sqrt((v * v).sum()), np.abs(v).max()],
                                   rtol=1e-13)
    assert np.abs(d).max() > 1e-3          # a non-trivial residual


def _totals(o, m):
    U = o.get_U()
    A = np.concatenate([_interior(o.block_geometry(b)[2], m, b).ravel() for b in range(m.nblk)])
    return (U * A[:, None]).sum(axis=0)


def test_closed_box_conserves_mass_and_energy():
    bcs = dict(W=abi.BC_REFLECTION, E=abi.BC_REFLECTION, S=abi.BC_REFLECTION, N=abi.BC_REFLECTION)
    m = M.cartesian(3, 2, 8, 8, x1=1.5, bcs=bcs, jitter=0.1, seed=9)
    W = _smooth_W(m.cell_centres(), seed=2)
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(W)
    t0 = _totals(o, m)
    rc, n, norms, _ = o.run(30)
    assert rc == 0 and n == 30 and norms[-1, 1] > 1e-6      # the flow moves
    t1 = _totals(o, m)
    assert abs(t1[0] - t0[0]) < 1e-13 * abs(t0[0])
    assert abs(t1[3] - t0[3]) < 1e-13 * abs(t0[3])
    assert abs(t1[1] - t0[1]) > 1e-6                          # walls exert pressure forces


def test_free_stream_preserved_on_a_disturbed_mesh():
    bcs = dict(W=abi.BC_FIXED, E=abi.BC_CHARACTERISTIC, S=abi.BC_CONSTANT_EXTRAPOLATION, N=abi.BC_FIXED)
    m = M.cartesian(3, 2, 8, 8, x1=1.5, bcs=bcs, jitter=0.2, seed=10)
    ws = np.array([1.4, 0.45, 0.15, 1.0])
    m.wo[:] = ws
    W = np.broadcast_to(ws, (m.ncell, 4)).copy()
    for s in (abi.Scheme(), abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING)):
        o = oracle.Oracle(m, s)
        o.set_W(W)
        rc, n, norms, _ = o.run(20)
        assert rc == 0 and n == 20
        assert np.max(np.abs(o.get_W() - W)) < 1e-13


def test_riemann_4quadrant_diagonal_mirror_symmetry():
    m = M.cartesian(4, 4, 8, 8)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(W0)
    rc, n, _, _ = o.run(25)
    assert rc == 0 and n == 25
    W = o.get_W()
    c = m.cell_centres()
    key = {(round(x * 64), round(y * 64)): k for k, (x, y) in enumerate(c)}
    mirror = np.array([key[(round(y * 64), round(x * 64))] for x, y in c])
    Wm = W[mirror][:, [0, 2, 1, 3]]
    assert np.max(np.abs(W - W0)) > 0.1                       # the waves have moved
    np.testing.assert_allclose(W, Wm, rtol=0, atol=1e-12)


def test_axisymmetric_source_is_the_only_residual_of_a_uniform_flow():
    """Axisymmetric (y the radius): a uniform state has zero net face flux,
    so the first stage's residual is (CFL*dt)*S(W, Xc) in every cell, with
    S the reference's source (Euler2DState.h:3179, bitwise against it in
    test_oracle_kat.py) at the cell's area centroid; planar, it is zero."""
    bcs = dict(W=abi.BC_FIXED, E=abi.BC_FIXED, S=abi.BC_FIXED, N=abi.BC_FIXED)
    m = M.cartesian(3, 2, 6, 6, x1=1.5, y0=0.3, y1=1.3, bcs=bcs, jitter=0.15, seed=12)
    ws = np.array([1.2, 0.35, -0.25, 0.9])
    m.wo[:] = ws
    W = np.broadcast_to(ws, (m.ncell, 4)).copy()
    for axi in (0, 1):
        s = abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING, axisymmetric=axi)
        o = oracle.Oracle(m, s)
        o.set_W(W)
        o.exchange_and_bcs()
        o.cfl(set_global=False)
        assert o.dudt_stage(1) == 0
        d = o.get_dUdt(0)
        if axi:
            yc = _centroids(o, m)[:, 1]
            ref = np.stack([s.cfl_number * dt * oracle.axisym_source(ws, 0.0, y) for dt, y in zip(o.get_dt(), yc)])
            np.testing.assert_allclose(d, ref, rtol=1e-12, atol=1e-15)
            assert np.abs(ref).min() > 1e-4
        else:
            assert np.abs(d).max() < 1e-14


def test_set_global_timestep_sets_every_cell():
    m = M.cartesian(2, 2, 4, 4, jitter=0.1, seed=3)
    o = oracle.Oracle(m, abi.Scheme())
    o.set_W(_smooth_W(m.cell_centres()))
    dtmin = o.cfl(set_global=False)
    assert np.ptp(o.get_dt()) > 0
    o.set_global_dt(0.5 * dtmin)
    np.testing.assert_array_equal(o.get_dt(), 0.5 * dtmin)


# Reduced least-squares stencils of the first ghost ring at physical
# boundaries, as the reference lists them (Euler2DQuadSingleBlock.cc:3692-3906,
# the same table as the Green-Gauss one at :3195-3410): offsets (di, dj) of
# the points used, per side, per BC class and per position along the side.
# "extrap" = PERIODIC, CONSTANT_EXTRAPOLATION, LINEAR_EXTRAPOLATION,
# CHARACTERISTIC; "other" = any other physical BC.
_ALL8 = {(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)}
_REF_STENCILS = {
    ("W", "extrap"): ({(-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)}, {(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0)}, _ALL8),
    ("W", "other"): ({(1, 0), (0, 1), (1, 1)}, {(0, -1), (1, -1), (1, 0)}, {(0, -1), (1, -1), (1, 0), (0, 1), (1, 1)}),
    ("E", "extrap"): ({(-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)}, {(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0)}, _ALL8),
    ("E", "other"): ({(-1, 0), (-1, 1), (0, 1)}, {(-1, -1), (0, -1), (-1, 0)},
                     {(-1, -1), (0, -1), (-1, 0), (-1, 1), (0, 1)}),
    ("S", "extrap"): ({(0, -1), (1, -1), (1, 0), (0, 1), (1, 1)}, {(-1, -1), (0, -1), (-1, 0), (-1, 1), (0, 1)}, _ALL8),
    ("S", "other"): ({(1, 0), (0, 1), (1, 1)}, {(-1, 0), (-1, 1), (0, 1)}, {(-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)}),
    ("N", "extrap"): ({(0, -1), (1, -1), (1, 0), (0, 1), (1, 1)}, {(-1, -1), (0, -1), (-1, 0), (-1, 1), (0, 1)}, _ALL8),
    ("N", "other"): ({(0, -1), (1, -1), (1, 0)}, {(-1, -1), (0, -1), (-1, 0)},
                     {(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0)}),
}


@pytest.mark.parametrize("bc,cls", [(abi.BC_CONSTANT_EXTRAPOLATION, "extrap"), (abi.BC_CHARACTERISTIC, "extrap"),
                                    (abi.BC_PERIODIC, "extrap"), (abi.BC_REFLECTION, "other"),
                                    (abi.BC_FIXED, "other")])
def test_boundary_ring_reduced_stencils(bc, cls):
    """Which cells the reconstruction of each first-ring ghost cell reads:
    on a linear field every stencil gives the exact gradient; perturbing one
    neighbour changes the gradient exactly when that neighbour is in the
    stencil.  The sets found must be the reference's table above; the
    outermost ring and the first ring's corners get no reconstruction."""
    n = 6
    m = M.cartesian(1, 1, n, n, bcs=dict(W=bc, E=bc, S=bc, N=bc), jitter=0.1, seed=9)
    W0 = M.uniform_state(m, M.freestream(0.4))
    m.set_wo_from_initial(W0)
    o = oracle.Oracle(m, abi.Scheme())
    xc, yc, _ = o.block_geometry(0)
    grad = np.array([[0.3, -0.2], [0.1, 0.05], [-0.07, 0.2], [0.4, 0.1]])
    base = np.stack([1.0 + grad[v, 0] * xc + grad[v, 1] * yc for v in range(4)], axis=-1)
    o.set_block_W_padded(0, base)
    o.reconstruct_block(0)
    gx0, gy0 = o.block_recon_padded(0, 0), o.block_recon_padded(0, 1)
    lo, hi = 2, n + 1                      # ICl, ICu (= JCl, JCu)
    ring = [(i, j) for j in range(1, n + 3) for i in range(1, n + 3) if i in (1, n + 2) or j in (1, n + 2)]
    for i, j in ring:
        corner = i in (1, n + 2) and j in (1, n + 2)
        if corner:
            assert not gx0[j, i].any() and not gy0[j, i].any(), (i, j)
            continue
        np.testing.assert_allclose(gx0[j, i], grad[:, 0], rtol=0, atol=1e-11)
        np.testing.assert_allclose(gy0[j, i], grad[:, 1], rtol=0, atol=1e-11)
    for j in range(n + 4):                 # outermost ring: n_pts = 0
        for i in range(n + 4):
            if i in (0, n + 3) or j in (0, n + 3):
                assert not gx0[j, i].any() and not gy0[j, i].any()
    for i, j in ring:
        if i in (1, n + 2) and j in (1, n + 2):
            continue
        side = "W" if i == 1 else "E" if i == n + 2 else "S" if j == 1 else "N"
        along = j if side in "WE" else i
        pos = 0 if along == lo else 1 if along == hi else 2
        used = set()
        for di, dj in _ALL8:
            W = base.copy()
            W[j + dj, i + di, 0] += 0.37
            o.set_block_W_padded(0, W)
            o.reconstruct_block(0)
            if o.block_recon_padded(0, 0)[j, i, 0] != gx0[j, i, 0] or \
                    o.block_recon_padded(0, 1)[j, i, 0] != gy0[j, i, 0]:
                used.add((di, dj))
        assert used == _REF_STENCILS[(side, cls)][pos], (side, cls, i, j, sorted(used))


def _lts_halving_case():
    """A strong symmetric expansion (u = -/+3, p = 0.1) under scalar local
    time stepping at CFL 0.9: some RK4 stage updates leave rho, E or e <= 0."""
    m = M.cartesian(4, 1, 8, 8, x1=1.0, y1=0.25)
    c = m.cell_centres()
    W0 = np.zeros((m.ncell, 4))
    W0[:, 0], W0[:, 3] = 1.0, 0.1
    W0[:, 1] = np.where(c[:, 0] < 0.5, -3.0, 3.0)
    m.set_wo_from_initial(W0)
    s = abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING, cfl_number=0.9)
    return m, W0, s


def _unphysical(U):
    e = U[:, 3] / U[:, 0] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / U[:, 0] ** 2
    return (U[:, 0] <= 0) | (U[:, 3] <= 0) | (e <= 0)


def test_local_time_stepping_halving_rule():
    """Update_Solution_Multistage_Explicit's scalar-local-time-stepping
    branch (SingleBlock.cc:5387-5425): a cell whose update is unphysical has
    its dt and dUdt[k] scaled by a factor that is halved before each of up to
    10 tries (so by 1/2, 1/8, 1/64, ... in total) until U = Uo + omega*dUdt is
    physical.  Checked cell by cell, stage by stage, from the values the
    oracle holds before and after each update."""
    m, W0, s = _lts_halving_case()
    o = oracle.Oracle(m, s)
    o.set_W(W0)
    omega = {1: 0.5, 2: 0.5, 3: 1.0, 4: 1.0 / 6.0}
    totals = [2.0 ** -(t * (t + 1) // 2) for t in range(1, 11)]
    halved = 0
    for step in range(12):
        o.cfl(set_global=False)
        Uo = o.get_U()
        for st in range(1, 5):
            k = 0 if st == 4 else st - 1
            o.exchange_and_bcs()
            o.dudt_stage(st)
            dt_b, d_b = o.get_dt(), o.get_dUdt(k)
            assert o.update_stage(st)[0] == 0
            dt_a, d_a, U_a = o.get_dt(), o.get_dUdt(k), o.get_U()
            r = dt_a / dt_b
            same = r == 1.0
            np.testing.assert_array_equal(d_a[same], d_b[same])
            for c in np.nonzero(~same)[0]:
                t = totals.index(r[c]) + 1            # the factor is one of the compounded halvings
                np.testing.assert_array_equal(d_a[c], d_b[c] * r[c])
                prev = 1.0 if t == 1 else totals[t - 2]
                assert _unphysical((Uo[c] + omega[st] * (d_b[c] * prev))[None])[0]
                halved += 1
            np.testing.assert_array_equal(U_a, Uo + omega[st] * d_a)
            assert not _unphysical(U_a).any()
    assert halved > 10
