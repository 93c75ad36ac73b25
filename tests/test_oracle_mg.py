"""The oracle's FAS multigrid (FASMultigrid2D/FASMultigrid2D.h) and the host
coarsening of meshes (CPU).

The reference publishes no multigrid trajectories and its full solver cannot
be built here, so the restatement is pinned by:
  * FluxHLLE_n and Bilinear_Interpolation bitwise against the reference's own
    compiled Euler2DState.cc / CFD.h (test_oracle_kat.py);
  * coarsening: on a uniform lattice the coarse level, ghost nodes included,
    is the half-resolution lattice exactly (Update_Exterior_Nodes' point
    reflection and wall mirror both continue a uniform lattice), and areas
    are conserved on jittered meshes;
  * Restrict_Solution_Blocks reproduces a linear field at the coarse
    centroids of a uniform mesh;
  * free-stream preservation through V and W cycles on a jittered mesh;
  * the FAS fixed point: on a bump channel the cycles converge to a state
    whose single-grid residual is round-off (the fine-grid steady state),
    and further cycles leave it unchanged while the coarse forcing terms
    stay O(1) -- a wrong sign or scaling of P, of the forcing or of the
    restriction breaks one or the other;
  * convergence: three levels reduce the residual 1000-fold in a quarter of
    the steps the single-grid smoother needs.
"""
import numpy as np
import pytest

import oracle
from cffc_amd import abi
from cffc_amd import mesh as M
from cffc_amd import multigrid as MG
from cffc_amd import steady


def _box(n=16, nbi=4, nbj=2, jitter=0.0, seed=3, bc=abi.BC_FIXED, blob=0.2):
    m = M.cartesian(nbi, nbj, n, n, x1=2.0, jitter=jitter, seed=seed)
    W0s = M.freestream(0.5)
    m.bc[:] = np.where(m.bc != abi.BC_NONE, bc, abi.BC_NONE)
    m.wo[:] = W0s
    W0 = M.uniform_state(m, W0s)
    c = m.cell_centres()
    W0[:, 0] *= 1.0 + blob * np.exp(-((c[:, 0] - 0.6) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.02)
    return m, W0


def _scheme(cfl=1.0, limiter=abi.LIMITER_VENKATAKRISHNAN):
    s = steady.bump_scheme(cfl=cfl)
    s.i_limiter = limiter
    return s


@pytest.mark.parametrize("bcs", [None, dict(W=abi.BC_FIXED, E=abi.BC_CHARACTERISTIC, S=abi.BC_REFLECTION,
                                            N=abi.BC_REFLECTION)])
def test_coarsen_uniform_lattice_exact(bcs):
    m = M.cartesian(3, 2, 16, 8, x1=1.5, bcs=bcs)
    c = MG.coarsen(m)
    ref = M.cartesian(3, 2, 8, 4, x1=1.5, bcs=bcs)
    np.testing.assert_array_equal(c.xnode, ref.xnode)
    np.testing.assert_array_equal(c.ynode, ref.ynode)
    np.testing.assert_array_equal(c.bc, ref.bc)
    np.testing.assert_array_equal(c.nbr, ref.nbr)


def test_coarsen_areas_and_interface_ghosts():
    m = M.cartesian(3, 3, 16, 16, jitter=0.1, seed=4)
    levels = MG.hierarchy(m, 3)
    tot = []
    for q in levels:
        o = oracle.Oracle(q, abi.Scheme())
        tot.append(sum(o.block_geometry(b)[2][2:int(q.nj[b]) + 2, 2:int(q.ni[b]) + 2].sum() for b in range(q.nblk)))
    np.testing.assert_allclose(tot, 1.0, rtol=1e-13)
    # every interface ghost node is the neighbour's node (the centre block)
    c = levels[1]
    X, Y = c.block_nodes(4)
    Xe, Ye = c.block_nodes(5)            # east neighbour
    n = int(c.ni[4])
    np.testing.assert_array_equal(X[2:n + 3, n + 3:n + 5], Xe[2:n + 3, 3:5])
    np.testing.assert_array_equal(Y[2:n + 3, n + 3:n + 5], Ye[2:n + 3, 3:5])


def test_coarsen_rejects_odd_and_small_blocks():
    with pytest.raises(ValueError):
        MG.coarsen(M.cartesian(2, 2, 6, 5))
    with pytest.raises(ValueError):
        MG.coarsen(M.cartesian(2, 2, 2, 2))   # 1-cell coarse blocks
    assert MG.coarsen(M.cartesian(2, 2, 4, 4)).ni[0] == 2


def test_restriction_exact_for_linear_fields():
    """A linear conserved field restricted to the coarse level equals the
    field at the coarse centroids (uniform mesh: the area-weighted mean of
    the four fine centroids is the coarse centroid)."""
    m, _ = _box(n=8)
    levels = MG.hierarchy(m, 2)
    mg = oracle.OracleMG(levels, _scheme(), MG.MGParams(levels=2))
    c = m.cell_centres()
    W = np.stack([1.0 + 0.2 * c[:, 0] - 0.1 * c[:, 1], 0 * c[:, 0], 0 * c[:, 0], 2.0 + 0.3 * c[:, 1]], axis=1)
    mg.set_W(W)
    mg.initialize()
    cc = levels[1].cell_centres()
    Uc = mg.levels[1].get_U()
    np.testing.assert_allclose(Uc[:, 0], 1.0 + 0.2 * cc[:, 0] - 0.1 * cc[:, 1], rtol=0, atol=1e-14)
    np.testing.assert_allclose(Uc[:, 3], (2.0 + 0.3 * cc[:, 1]) / 0.4, rtol=0, atol=1e-13)


@pytest.mark.parametrize("cycle", [abi.MULTIGRID_V_CYCLE, abi.MULTIGRID_W_CYCLE])
def test_free_stream_preserved(cycle):
    m, W0 = _box(n=8, jitter=0.1, blob=0.0)
    mg = oracle.OracleMG(MG.hierarchy(m, 3), _scheme(), MG.MGParams(levels=3, i_cycle=cycle))
    mg.set_W(W0)
    mg.initialize()
    for _ in range(3):
        assert mg.cycle() == 0
    assert np.max(np.abs(mg.get_W() - W0)) < 1e-13


def test_converges_to_the_fine_steady_state_and_stays():
    """On a bump channel (a non-trivial steady flow, Roe, unlimited): the
    cycles drive the fine residual to round-off, the single-grid residual of
    the result is round-off too (the FAS solution is the fine-grid steady
    state, not a coarse-grid one), and further cycles leave it unchanged
    while the coarse forcing terms stay O(1)."""
    m, W0 = steady.bump_case(0.08, 0.5, nbi=3, nbj=1, ni=8, nj=8)
    s = _scheme(limiter=abi.LIMITER_ONE)
    mg = oracle.OracleMG(MG.hierarchy(m, 3), s, MG.MGParams(levels=3))
    mg.set_W(W0)
    mg.initialize()
    rc, cycles, norms = mg.execute(320)
    assert rc == 0 and norms[-1, 1] < 1e-11 * norms[1, 1]
    Wc = mg.get_W()
    o = oracle.Oracle(m, s)
    o.set_W(Wc)
    _, _, sg, _ = o.run(1)
    assert sg[0, 1] < 1e-12
    for _ in range(2):
        assert mg.cycle() == 0
    assert np.max(np.abs(mg.get_W() - Wc)) < 1e-13
    assert np.max(np.abs(mg.get_P(1))) > 0.1 and np.max(np.abs(mg.get_P(2))) > 0.1


def test_multigrid_converges_faster_than_single_grid():
    """Bump channel, Venkatakrishnan limiter: three levels reduce the
    residual 1000-fold in far fewer cycles than the single-grid smoother
    needs steps (each cycle costs about two fine smoothings)."""
    m, W0 = steady.bump_case(0.08, 0.5, nbi=3, nbj=1, ni=8, nj=8)
    s = _scheme()
    mg = oracle.OracleMG(MG.hierarchy(m, 3), s, MG.MGParams(levels=3))
    mg.set_W(W0)
    mg.initialize()
    rc, cycles, norms = mg.execute(100)
    assert rc == 0 and cycles == 100
    l2 = norms[1:, 1]
    assert np.any(l2 < 1e-3 * l2[0]), l2[::10]
    hit = int(np.argmax(l2 < 1e-3 * l2[0])) + 1
    o = oracle.Oracle(m, s)
    o.set_W(W0)
    _, n, sg, _ = o.run(1500)
    first = sg[0, 1]
    sg_hit = int(np.argmax(sg[:, 1] < 1e-3 * first)) + 1 if np.any(sg[:, 1] < 1e-3 * first) else len(sg) + 1
    print("multigrid: 1e-3 drop after %d cycles; single grid after %d steps" % (hit, sg_hit))
    assert 4 * hit < sg_hit


def test_full_multigrid_prolongs_linear_fields_exactly():
    """Execute with fmg_cycles > 0 and no cycle budget: level 1 (restricted
    from level 0 at initialisation) is prolonged onto level 0 and nothing
    else runs.  On a uniform mesh a linear conserved field comes back exactly
    in every fine cell whose interpolation quadrilateral is interior (all but
    the outermost fine layer along the physical boundary): a fine cell
    written at the wrong index, an added rather than assigned solution or a
    coarse centroid in place of the fine one breaks it."""
    m, _ = _box(n=8)
    mg = oracle.OracleMG(MG.hierarchy(m, 3), _scheme(), MG.MGParams(levels=3, fmg_cycles=4))
    c = m.cell_centres()
    W = np.stack([1.0 + 0.2 * c[:, 0] - 0.1 * c[:, 1], 0 * c[:, 0], 0 * c[:, 0], 2.0 + 0.3 * c[:, 1]], axis=1)
    mg.set_W(W)
    mg.initialize()
    U0 = mg.levels[0].get_U()
    U0[:] = 7.0                          # the prolongation must overwrite it
    mg.levels[0].set_U(U0)
    rc, n, norms = mg.execute(0)
    assert rc == 0 and n == 0 and len(norms) == 2
    U = mg.levels[0].get_U()
    h = 1.0 / 16                         # 32 x 16 cells on [0, 2] x [0, 1]
    inner = (c[:, 0] > h) & (c[:, 0] < 2.0 - h) & (c[:, 1] > h) & (c[:, 1] < 1.0 - h)
    assert inner.sum() == 30 * 14
    np.testing.assert_allclose(U[inner, 0], W[inner, 0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(U[inner, 3], W[inner, 3] / 0.4, rtol=0, atol=1e-13)
    assert np.all(U[~inner] != 7.0)


def test_full_multigrid_start_reaches_the_regular_fixed_point():
    """Bump channel, unlimited Roe: the full-multigrid start (five cycles
    with level 1 on top, checked on level 1, then the solution prolonged)
    converges to the state the regular cycles reach, within round-off; its
    first checks are level-1 norms, and free stream stays free stream
    through the start-up on a uniform mesh.  (Not on a jittered one:
    Restrict_Solution_Blocks divides the area-weighted sum by the coarse
    cell's area, which there differs from the four fine areas, so the
    coarse start-up level begins from a non-uniform state.)"""
    m, W0 = steady.bump_case(0.08, 0.5, nbi=3, nbj=1, ni=8, nj=8)
    s = _scheme(limiter=abi.LIMITER_ONE)
    out = {}
    for fmg in (0, 5):
        mg = oracle.OracleMG(MG.hierarchy(m, 3), s, MG.MGParams(levels=3, fmg_cycles=fmg))
        mg.set_W(W0)
        mg.initialize()
        rc, n, norms = mg.execute(320)
        assert rc == 0 and n == 320 and len(norms) == 321
        assert norms[-1, 1] < 1e-11 * norms[1, 1]
        out[fmg] = (mg.get_W(), norms)
    np.testing.assert_allclose(out[5][0], out[0][0], rtol=0, atol=1e-12)
    assert not np.allclose(out[5][1][:5], out[0][1][:5])
    mb, Wb = _box(n=8, jitter=0.0, blob=0.0)
    mg = oracle.OracleMG(MG.hierarchy(mb, 3), _scheme(), MG.MGParams(levels=3, fmg_cycles=3))
    mg.set_W(Wb)
    mg.initialize()
    rc, n, _ = mg.execute(6)
    assert rc == 0 and n == 6
    assert np.max(np.abs(mg.get_W() - Wb)) < 1e-13
