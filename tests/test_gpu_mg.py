"""FAS multigrid on the GPU against the oracle (FASMultigrid2D.h restated).

The device cycle (cffc_mg_* through the C ABI: granular FP64 stage kernels
with HLLE / the zero limiter on coarse levels, the forcing term, the level
transfers of euler2d_mg.hip) must reproduce the oracle bit for bit on every
level: the solution, the forcing terms and the residual norms (norms to
1e-12 relative: their reduction order differs).

Bump channels are compared with both outflows the reference's bump grids
name: constant extrapolation and BC_Characteristic_Pressure.  The latter
calls pow(); the device evaluates it by glibc's own method
(cffc_amd/csrc/glibc_pow.h), so it too is compared bit for bit.
"""
import numpy as np
import pytest

import oracle
from cffc_amd import abi
from cffc_amd import mesh as M
from cffc_amd import multigrid as MG
from cffc_amd import steady

pytestmark = pytest.mark.gpu


def _box(n=8, nbi=3, nbj=2, jitter=0.1, seed=3, blob=0.2):
    m = M.cartesian(nbi, nbj, n, n, x1=1.5, jitter=jitter, seed=seed)
    W0s = M.freestream(0.5)
    m.bc[:] = np.where(m.bc != abi.BC_NONE, abi.BC_FIXED, abi.BC_NONE)
    m.wo[:] = W0s
    W0 = M.uniform_state(m, W0s)
    c = m.cell_centres()
    W0[:, 0] *= 1.0 + blob * np.exp(-((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.02)
    return m, W0


def _scheme(limiter=abi.LIMITER_VENKATAKRISHNAN):
    s = steady.bump_scheme(cfl=1.0)
    s.i_limiter = limiter
    return s


def _bump(ni=8, outflow=abi.BC_CONSTANT_EXTRAPOLATION):
    """Bump channel (3 x 1 blocks of ni^2 cells): reflecting walls, fixed
    inflow, and the given outflow."""
    m, W0 = steady.bump_case(0.08, 0.5, nbi=3, nbj=1, ni=ni, nj=ni)
    m.bc[m.bc == abi.BC_CHARACTERISTIC] = outflow
    return m, W0


def _nonbitwise(a, b):
    return int(np.sum(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64)))


def _pair(m, W0, s, p, rs=None):
    meshes = MG.hierarchy(m, p.levels)
    g = MG.FASMultigrid(m, s, p, meshes=meshes)
    o = oracle.OracleMG(meshes, s, p)
    if rs:
        g.set_residual_smoothing(*rs)
        o.set_residual_smoothing(*rs)
    g.set_W(W0)
    o.set_W(W0)
    g.initialize()
    o.initialize()
    return g, o


def _compare_levels(g, o, tag):
    for l in range(len(g.levels)):
        Wg, Wo = g.levels[l].get_W(), o.levels[l].get_W()
        Ug, Uo = g.levels[l].get_U(), o.levels[l].get_U()
        print("%s level %d: W %d non-bitwise, U %d non-bitwise of %d" %
              (tag, l, _nonbitwise(Wg, Wo), _nonbitwise(Ug, Uo), Wg.size), flush=True)
        np.testing.assert_array_equal(Ug, Uo)
        np.testing.assert_array_equal(Wg, Wo)
        if l > 0:
            np.testing.assert_array_equal(g.get_P(l), o.get_P(l))


@pytest.mark.parametrize("cycle", [abi.MULTIGRID_V_CYCLE, abi.MULTIGRID_W_CYCLE])
def test_cycles_match_oracle(cycle):
    m, W0 = _box()
    p = MG.MGParams(levels=3, i_cycle=cycle, post_smooths=1 if cycle == abi.MULTIGRID_W_CYCLE else 0)
    g, o = _pair(m, W0, _scheme(), p)
    print("cycle type %d, %d cells on 3 levels" % (cycle, m.ncell), flush=True)
    _compare_levels(g, o, "initialized")
    for n in range(4):
        assert g.cycle() == o.cycle() == 0
    _compare_levels(g, o, "after 4 cycles")
    g.close()


def test_axisymmetric_cycles_match_oracle():
    """Axisymmetric levels (the source in every smoothing stage and in the
    residual evaluation, dUdt_Residual_Evaluation :1016-1020)."""
    m, W0 = _box(seed=6)
    s = _scheme()
    s.axisymmetric = 1
    g, o = _pair(m, W0, s, MG.MGParams(levels=3))
    for n in range(3):
        assert g.cycle() == o.cycle() == 0
    _compare_levels(g, o, "axisymmetric")
    g.close()


def test_variants_match_oracle():
    """Second-order coarse levels (Roe + limiter), injection prolongation,
    no defect correction, residual smoothing."""
    m, W0 = _box(seed=5)
    for kw, rs in ((dict(first_order_coarse=0), None), (dict(prolong_using_injection=1), None),
                   (dict(defect_correction=0, smooths_on_coarsest_level=2), None), ({}, (0.4, 2))):
        p = MG.MGParams(levels=3, **kw)
        g, o = _pair(m, W0, _scheme(), p, rs)
        print("variant", kw, rs, flush=True)
        for n in range(3):
            assert g.cycle() == o.cycle() == 0
        _compare_levels(g, o, str(kw))
        g.close()


@pytest.mark.parametrize("outflow", [abi.BC_CONSTANT_EXTRAPOLATION, abi.BC_CHARACTERISTIC])
def test_execute_bump_matches_oracle(outflow):
    """Bump channel: 150 cycles of Execute (the unfrozen limiter stalls the
    residual near 1e-5 here), same norms and state as the oracle, bit for bit
    on every level, with either outflow."""
    m, W0 = _bump(outflow=outflow)
    p = MG.MGParams(levels=3)
    g, o = _pair(m, W0, _scheme(), p)
    rc, n, norms = g.execute(150)
    rc2, n2, norms2 = o.execute(150)
    print("bump, outflow %d: %d cycles on the GPU, %d in the oracle; final L2 %.3e" % (outflow, n, n2, norms[-1, 1]),
          flush=True)
    assert rc == rc2 == 0 and n == n2 == 150
    np.testing.assert_allclose(norms, norms2, rtol=1e-12, atol=0)
    _compare_levels(g, o, "bump, outflow %d" % outflow)
    g.close()


@pytest.mark.parametrize("outflow", [abi.BC_CONSTANT_EXTRAPOLATION, abi.BC_CHARACTERISTIC])
def test_execute_bump_with_freezing_matches_oracle(outflow):
    """The same with the limiters frozen at 1e-4 (graph recaptured at the
    freeze), bit for bit with either outflow."""
    m, W0 = _bump(outflow=outflow)
    p = MG.MGParams(levels=3, absolute_tolerance=1e-9)
    g, o = _pair(m, W0, _scheme(), p)
    rc, n, norms = g.execute(300, freeze_level=1e-4)
    rc2, n2, norms2 = o.execute(300, freeze_level=1e-4)
    print("bump, outflow %d, frozen at 1e-4: %d / %d cycles" % (outflow, n, n2), flush=True)
    assert rc == rc2 == 0 and n == n2
    if outflow == abi.BC_CHARACTERISTIC:
        assert n < 300
    np.testing.assert_allclose(norms, norms2, rtol=1e-12, atol=0)
    _compare_levels(g, o, "frozen, outflow %d" % outflow)
    g.close()


def test_variants_on_the_bump_match_oracle():
    """Injection prolongation (with the Neumann copies of the extrapolated
    outflow skipped) and second-order coarse levels on the bump channel,
    30 cycles each, every level bit for bit."""
    m, W0 = _bump()
    for kw in (dict(prolong_using_injection=1), dict(first_order_coarse=0)):
        g, o = _pair(m, W0, _scheme(), MG.MGParams(levels=3, **kw))
        rc, n, _ = g.execute(30)
        rc2, n2, _ = o.execute(30)
        print("bump variant", kw, flush=True)
        assert rc == rc2 == 0 and n == n2 == 30
        _compare_levels(g, o, str(kw))
        g.close()


def test_multigrid_solution_is_the_fine_steady_state():
    """The converged multigrid state has a round-off single-grid residual on
    the GPU's own fused solver (the fine-grid steady state)."""
    from cffc_amd import euler2d
    m, W0 = steady.bump_case(0.08, 0.5, nbi=3, nbj=1, ni=8, nj=8)
    s = _scheme(limiter=abi.LIMITER_ONE)
    p = MG.MGParams(levels=3)
    g = MG.FASMultigrid(m, s, p)
    g.set_W(W0)
    g.initialize()
    rc, n, norms = g.execute(320)
    assert rc == 0 and norms[-1, 1] < 1e-11 * norms[1, 1]
    W = g.get_W()
    g.close()
    f = euler2d.Euler2DQuadBlocks(m, s)
    f.set_W(W)
    rc, steps, sg, _ = f.run_steps(1)
    print("single-grid residual of the multigrid state: %.3e" % sg[0, 1], flush=True)
    assert rc == 0 and sg[0, 1] < 1e-12
    f.close()


def test_graph_replays_equal_eager_launches():
    """Execute replays one captured cycle (recaptured when the limiters
    freeze), with the norm kernels and host reads between the replays; the
    same runs issued launch by launch give the same bits."""
    m, W0 = steady.bump_case(0.08, 0.5, nbi=3, nbj=1, ni=8, nj=8)
    for freeze in (0.0, 1e-4):
        out = []
        for graph in (True, False):
            g = MG.FASMultigrid(m, _scheme(), MG.MGParams(levels=3))
            g.set_graph(graph)
            g.set_W(W0)
            g.initialize()
            rc, n, norms = g.execute(120, freeze_level=freeze)
            out.append((rc, n, norms, g.get_W(), [g.levels[l].get_U() for l in range(3)]))
            g.close()
        (rc1, n1, nr1, W1, U1), (rc2, n2, nr2, W2, U2) = out
        print("freeze %g: graph %d cycles, eager %d cycles; W %d non-bitwise" % (freeze, n1, n2, _nonbitwise(W1, W2)),
              flush=True)
        assert rc1 == rc2 == 0 and n1 == n2 == 120
        np.testing.assert_array_equal(nr1, nr2)
        np.testing.assert_array_equal(W1, W2)
        for a, b in zip(U1, U2):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("graph", [True, False])
def test_full_multigrid_execute_matches_oracle(graph):
    """Execute with the full-multigrid start-up (Number_of_Cycles_per_Stage_
    for_Full_Multigrid): six cycles with level 1 on top (the smoother's
    scheme there, HLLE and the zero limiter below), the solution prolonged
    to level 0, then the regular cycles -- replayed from captured graphs
    (recaptured at the new top level) or launched one by one; norms, state
    and every level bit for bit as in the oracle."""
    m, W0 = _bump()
    p = MG.MGParams(levels=3, fmg_cycles=6)
    g, o = _pair(m, W0, _scheme(), p)
    g.set_graph(graph)
    rc, n, norms = g.execute(80)
    rc2, n2, norms2 = o.execute(80)
    print("bump, full multigrid (graph %s): %d / %d cycles, %d checks, final L2 %.3e" %
          (graph, n, n2, len(norms), norms[-1, 1]), flush=True)
    assert rc == rc2 == 0 and n == n2 == 80 and len(norms) == len(norms2) == 81
    np.testing.assert_allclose(norms, norms2, rtol=1e-12, atol=0)
    _compare_levels(g, o, "full multigrid")
    g.close()


@pytest.mark.parametrize("budget", [3, 12])
def test_full_multigrid_four_levels_w_cycle_matches_oracle(budget):
    """Four levels, W cycles with a post-smoothing: start-up on levels 2 and
    1 (four cycles each), then level 0; with a budget of 3 the cycles run
    out on level 2 and the solution is still prolonged through levels 1 and
    0, each stopping at its first check."""
    m, W0 = _bump(ni=16)
    p = MG.MGParams(levels=4, fmg_cycles=4, i_cycle=abi.MULTIGRID_W_CYCLE, post_smooths=1)
    g, o = _pair(m, W0, _scheme(), p)
    rc, n, norms = g.execute(budget)
    rc2, n2, norms2 = o.execute(budget)
    print("4 levels, W, budget %d: %d / %d cycles, %d / %d checks" % (budget, n, n2, len(norms), len(norms2)),
          flush=True)
    assert rc == rc2 == 0 and n == n2 == budget and len(norms) == len(norms2)
    np.testing.assert_allclose(norms, norms2, rtol=1e-12, atol=0)
    _compare_levels(g, o, "4 levels, budget %d" % budget)
    g.close()


def test_serving_batch_with_multigrid_converges():
    """serving.serve_multigrid on 4 cases of two mesh sizes (96 x 32 and
    192 x 64 cells in blocks of 32^2): one multigrid batch, run in segments,
    reaches every case's single-grid-equivalent tolerance, per case and with
    the batch norm against the strictest case."""
    from cffc_amd import serving
    cases = [(0.05, 0.4, 1), (0.08, 0.55, 2), (0.06, 0.35, 1), (0.09, 0.5, 2)]
    r = serving.serve_multigrid(cases, levels=3, segment=200, log=lambda s: print(s, flush=True))
    print("per case: %d cycles, %.2f s, cycles to convergence %s" % (r["cycles"], r["wall_seconds"],
                                                                      r["case_cycles"]), flush=True)
    assert r["rc"] == 0 and r["converged"] == 4 and r["cycles"] == max(r["case_cycles"])
    assert np.all(np.array(r["case_final_R"]) < np.array(r["tol_R"]))
    b = serving.serve_multigrid(cases, levels=3, segment=200, per_case=False)
    print("batch norm: %d cycles, %.2f s, R %.3e (tol %.3e)" % (b["cycles"], b["wall_seconds"], b["final_R"],
                                                                 b["tol_R"]), flush=True)
    assert b["rc"] == 0 and b["converged"] == 4 and b["final_R"] < b["tol_R"]


@pytest.mark.parametrize("limiter", [abi.LIMITER_ZERO, abi.LIMITER_VENKATAKRISHNAN])
def test_per_case_stopping_equals_each_case_alone(limiter):
    """cffc_mg_set_parts: a batch of three bump cases of two mesh sizes,
    each with its own tolerance (and, with the Venkatakrishnan limiter, its
    own limiter freeze level), reaches each case's tolerance after exactly
    the cycles the case takes alone, and stops with the last of them.  A
    converged case is left alone by the rest of the batch's cycles, so every
    case ends in its state alone, bit for bit, with its norm alone."""
    from cffc_amd import serving
    cases = [(0.05, 0.4, 1), (0.08, 0.55, 2), (0.06, 0.35, 1)]
    zero = limiter == abi.LIMITER_ZERO
    tols = np.array([3e-8, 1e-8, 2e-8]) if zero else np.array([3e-9, 1e-9, 2e-9])
    freeze = None if zero else 1e5 * tols

    def run(parts):
        meshes = [serving.case_mesh(*cases[q]) for q in parts]
        m, starts = M.concat([mm[0] for mm in meshes])
        g = MG.FASMultigrid(m, _scheme(limiter), MG.MGParams(levels=3, absolute_tolerance=1e-30))
        g.set_W(np.concatenate([mm[1] for mm in meshes]))
        g.initialize()
        g.set_parts(list(starts) + [m.nblk], tols[parts], None if zero else freeze[parts])
        rc, n, norms = g.execute(6000, freeze_level=0.0 if zero else 1.0)
        l2, done = g.part_status()
        W = g.get_W()
        g.close()
        assert rc == 0
        offs = np.cumsum([0] + [mm[0].ncell for mm in meshes])
        return n, l2, done, [W[offs[k]:offs[k + 1]] for k in range(len(parts))]

    alone = [run([q]) for q in range(3)]
    n, l2, done, Ws = run([0, 1, 2])
    print("alone:", [(a[0], a[2][0]) for a in alone], "batch:", n, done.tolist(), flush=True)
    assert [a[2][0] for a in alone] == done.tolist()
    assert n == max(a[0] for a in alone) == done.max()
    for q in range(3):
        assert l2[q] == alone[q][1][0]
        assert np.array_equal(Ws[q].view(np.uint64), alone[q][3][0].view(np.uint64)), "case %d" % q


def test_initialize_restarts_the_cases():
    """cffc_mg_initialize after a batch run with parts starts every case
    afresh (none converged or skipped): the same state run again takes the
    same cycles per case and ends in the same states."""
    from cffc_amd import serving
    cases = [(0.05, 0.4, 1), (0.08, 0.55, 2)]
    meshes = [serving.case_mesh(*c) for c in cases]
    m, starts = M.concat([mm[0] for mm in meshes])
    W0 = np.concatenate([mm[1] for mm in meshes])
    g = MG.FASMultigrid(m, _scheme(abi.LIMITER_ZERO), MG.MGParams(levels=3, absolute_tolerance=1e-30))
    out = []
    for rep in range(2):
        g.set_W(W0)
        g.initialize()
        if rep == 0:
            g.set_parts(list(starts) + [m.nblk], [3e-8, 1e-8])
        rc, n, _ = g.execute(4000)
        l2, done = g.part_status()
        assert rc == 0 and (done >= 0).all()
        out.append((n, done.tolist(), g.get_W()))
    g.close()
    print("runs:", [(o[0], o[1]) for o in out], flush=True)
    assert out[0][:2] == out[1][:2]
    assert np.array_equal(out[0][2].view(np.uint64), out[1][2].view(np.uint64))
