"""Blocks of uneven size on the GPU (lattice meshes, one launch per stage for
all block sizes).

A merged quadtree (merge_quadtree) holds the same cells and nodes as the
equal-block quadtree it comes from, with blocks of n^2, (2n)^2 and (4n)^2
cells at every level.  Its discretisation is the quadtree's: the same
ghost values (copies, restrictions, prolongations), the same face fluxes and
the same flux corrections.  So the merged run must reproduce the oracle's
run of the quadtree (the reference's message rules) bit for bit, cell for
cell through the merge's cell map.  Also: the lattice path on the equal
quadtree equals the neighbour-table path; mass is conserved; the mixed
path is held to a bound against FP64.
"""
import numpy as np
import pytest

import oracle
from cffc_amd import abi
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from cffc_amd import euler2d as E
    E.load()
    return E


def nonbitwise(a, b):
    return int(np.sum(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64)))


def case(nroot=8, n=8, seed=5, radius=0.3):
    m = M.quadtree_box(nroot=nroot, n=n, max_level=2, refine=M.circle_refinement(radius, band=0.02), jitter=0.1,
                       seed=seed)
    W0 = M.density_perturbation(m, M.cylindrical_explosion(m, radius=radius), seed=seed)
    mm, cmap = M.merge_quadtree(m)
    return m, W0, mm, cmap


def scheme(**kw):
    s = abi.Scheme(cfl_number=0.4)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_lattice_path_equals_neighbour_tables_on_equal_blocks(E):
    m, W0, _, _ = case()
    lat = M.Mesh(**{k: getattr(m, k) for k in ("nblk", "ni", "nj", "xnode", "ynode", "nbr", "bc", "wo", "level",
                                                 "nbr2", "lattice")}, info=dict(m.info, use_lattice=True))
    out = []
    for mesh in (m, lat):
        g = E.Euler2DQuadBlocks(mesh, scheme())
        g.set_W(W0)
        rc, n, norms, _ = g.run_steps(20)
        assert rc == 0 and n == 20
        out.append((g.get_W(), norms))
        g.close()
    nb = nonbitwise(out[0][0], out[1][0])
    print("equal blocks, lattice vs neighbour tables: %d non-bitwise values" % nb, flush=True)
    assert nb == 0
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("graph", [False, True])
def test_equal_blocks_with_walls_next_to_level_changes_match_the_oracle(E, graph):
    """Refinement that reaches the blocks along the walls: the fine blocks'
    ghosts are prolonged from coarse cells whose stencils hold wall ghosts,
    which the coarse-to-fine messages read as the previous BCs call left
    them (the fused path keeps them in both W buffers)."""
    m, W0, _, _ = case()
    assert any(m.level[b] > 0 and m.level[o] == 0 and (m.bc[m.bc_offsets()[o]:m.bc_offsets()[o + 1]] != 0).any()
               for b in range(m.nblk) for o in m.nbr[b] if o >= 0)
    s = scheme()
    orc = oracle.Oracle(m, s)
    orc.set_W(W0)
    assert orc.run(30)[0] == 0
    g = E.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    rc, n, _, _ = g.run_steps(30, graph=graph)
    assert rc == 0 and n == 30
    nb = nonbitwise(g.get_W(), orc.get_W())
    g.close()
    print("equal blocks, walls next to level changes, 30 steps: %d non-bitwise" % nb, flush=True)
    assert nb == 0


@pytest.mark.parametrize("graph", [False, True])
def test_merged_blocks_match_the_oracle_on_the_quadtree(E, graph):
    m, W0, mm, cmap = case()
    s = scheme()
    orc = oracle.Oracle(m, s)
    orc.set_W(W0)
    rc, n, norms_o, dt_o = orc.run(30)
    assert rc == 0 and n == 30
    Wo = orc.get_W()
    g = E.Euler2DQuadBlocks(mm, s)
    g.set_W(W0[cmap])
    rc, n, norms, dts = g.run_steps(30, graph=graph)
    assert rc == 0 and n == 30
    Wg = g.get_W()
    g.close()
    nb = nonbitwise(Wg, Wo[cmap])
    print("merged mesh (%d blocks of %s cells) vs the oracle on the quadtree (%d blocks), 30 steps: %d non-bitwise"
          % (mm.nblk, sorted(set(mm.ni.tolist())), m.nblk, nb), flush=True)
    assert nb == 0
    assert np.array_equal(dts, dt_o)
    assert np.allclose(norms, norms_o, rtol=1e-12, atol=0.0)


def test_merged_granular_calls_match_the_oracle(E):
    """The reference-named calls (CFL, exchange + BCs, dUdt, flux
    corrections, update) in the reference's loop order on the merged mesh."""
    m, W0, mm, cmap = case(nroot=4, n=8)
    s = scheme()
    orc = oracle.Oracle(m, s)
    orc.set_W(W0)
    g = E.Euler2DQuadBlocks(mm, s)
    g.set_W(W0[cmap])
    rc, _ = g.step_reference_order(5)
    assert rc == 0
    assert orc.run(5)[0] == 0
    nb = nonbitwise(g.get_W(), orc.get_W()[cmap])
    print("merged mesh, granular calls, 5 steps: %d non-bitwise" % nb, flush=True)
    assert nb == 0
    g.close()


def test_merged_mass_conserved_and_mixed_bounded(E):
    m, W0, mm, cmap = case(nroot=8, n=8, seed=7)
    Wi = W0[cmap]
    out = {}
    for prec in (abi.FLUX_FP64, abi.FLUX_FP32):
        g = E.Euler2DQuadBlocks(mm, scheme(flux_precision=prec))
        g.set_W(Wi)
        U0 = g.get_U()
        rc, n, _, _ = g.run_steps(40)
        assert rc == 0 and n == 40
        out[prec] = (g.get_W(), g.get_U(), U0)
        g.close()
    import bench
    A = bench.cell_areas(mm)
    W64, U64, U0 = out[abi.FLUX_FP64]
    drift = abs(np.sum(A * U64[:, 0]) - np.sum(A * U0[:, 0])) / np.sum(A * U0[:, 0])
    d = np.abs(out[abi.FLUX_FP32][0] - W64)
    print("merged mesh: mass drift %.3e, mixed vs FP64 L1 %s Linf %s" % (drift, d.mean(axis=0), d.max(axis=0)),
          flush=True)
    assert drift < 1e-14
    assert np.all(d.mean(axis=0) < 1e-5) and np.all(d.max(axis=0) < 1e-2)


def test_c3_full_size_uneven_blocks_match_the_oracle(E):
    """The bench's C3 workload at full size: 2,956 blocks of 16^2, 32^2 and
    64^2 cells from three levels (5,131,264 cells), one stage launch for all
    of them, 3 graph-replayed FP64 steps, against the oracle's run of the
    20,044-block equal quadtree it was merged from: bit for bit through the
    cell map, and the same time steps."""
    import os
    import bench
    mm, W0m, q, cmap = bench.amr_case(with_quadtree=True)
    print("C3 mesh: %d blocks (%s), %d cells; quadtree %d blocks" % (
        mm.nblk, dict(zip(*[a.tolist() for a in np.unique(mm.ni, return_counts=True)])), mm.ncell, q.nblk), flush=True)
    assert mm.nblk == 2956 and mm.ncell == 5131264 and sorted(set(mm.ni.tolist())) == [16, 32, 64]
    W0 = np.empty_like(W0m)
    W0[cmap] = W0m
    s = abi.Scheme()
    orc = oracle.Oracle(q, s)
    orc.set_W(W0)
    rc, n, norms_o, dt_o = orc.run(3, nthreads=min(16, os.cpu_count() or 1))
    assert rc == 0 and n == 3
    Wo = orc.get_W()
    g = E.Euler2DQuadBlocks(mm, s)
    g.set_W(W0m)
    rc, n, norms, dts = g.run_steps(3, graph=True)
    assert rc == 0 and n == 3
    Wg = g.get_W()
    g.close()
    nb = nonbitwise(Wg, Wo[cmap])
    print("C3 full size, 3 FP64 steps: %d of %d values not bit-identical" % (nb, Wg.size), flush=True)
    assert nb == 0
    assert np.array_equal(dts, dt_o)
    assert np.allclose(norms, norms_o, rtol=1e-12, atol=0.0)
