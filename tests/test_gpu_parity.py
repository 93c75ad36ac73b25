"""Parity of the HIP path (through the C ABI) with the CPU oracle.

FP64 runs must agree with the oracle within 1e-12 relative per cell (the
kernels keep the reference's operation order, so bit identity is expected and
reported); the residual norms are reduced in a different order and are
compared at 1e-12 relative.  Mixed FP32-flux runs report their L1/Linf error
against the FP64 oracle and are held to a loose bound.
"""
import numpy as np
import pytest

import oracle
from cffc_amd import abi
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.fixture(scope="module")
def euler2d():
    from cffc_amd import euler2d as E
    E.load()
    return E


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def nonbitwise(a, b):
    return int(np.sum(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64)))


def run_both(euler2d, mesh, W0, scheme, nsteps, nthreads=8, tile=None):
    gpu = euler2d.Euler2DQuadBlocks(mesh, scheme)
    if tile:
        gpu.set_tile(*tile)
    gpu.set_W(W0)
    rc, n, norms, dts = gpu.run_steps(nsteps)
    orc = oracle.Oracle(mesh, scheme)
    orc.set_W(W0)
    rc2, n2, norms2, dts2 = orc.run(nsteps, nthreads=nthreads)
    return gpu, orc, (rc, n, norms, dts), (rc2, n2, norms2, dts2)


def test_riemann_4quadrant_small_fp64(euler2d):
    m = M.cartesian(4, 4, 8, 8)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    gpu, orc, g, o = run_both(euler2d, m, W0, abi.Scheme(), 20)
    assert g[0] == 0 and o[0] == 0 and g[1] == o[1] == 20
    np.testing.assert_array_equal(g[3], o[3])          # global dt: min is order independent
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("W rel %.3e, %d non-bitwise" % (relerr(Wg, Wo), nonbitwise(Wg, Wo)))
    assert relerr(Wg, Wo) <= REL
    assert relerr(gpu.get_U(), orc.get_U()) <= REL
    assert relerr(g[2], o[2]) <= REL


def test_granular_calls_match_oracle_stage_by_stage(euler2d):
    """The reference-named entry points, one call at a time."""
    m = M.cartesian(3, 2, 8, 8, jitter=0.1, seed=3)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    s = abi.Scheme()
    gpu = euler2d.Euler2DQuadBlocks(m, s)
    orc = oracle.Oracle(m, s)
    gpu.set_W(W0)
    orc.set_W(W0)
    for n in range(3):
        dg, do = gpu.CFL(), orc.cfl()
        assert dg == do
        for st in range(1, 5):
            gpu.Send_All_Messages_and_BCs()
            orc.exchange_and_bcs()
            gpu.dUdt_Multistage_Explicit(st)
            orc.dudt_stage(st)
            k = 0 if st == 4 else st - 1
            assert relerr(gpu.get_dUdt(k), orc.get_dUdt(k)) <= REL, (n, st)
            assert gpu.Update_Solution_Multistage_Explicit(st) == 0
            assert orc.update_stage(st)[0] == 0
            assert relerr(gpu.get_W(), orc.get_W()) <= REL, (n, st)
        ng, no = gpu.Residual_Norms(1), orc.residual_norms(1)
        assert relerr(ng, no) <= REL


def test_riemann_4quadrant_C1_full_fp64(euler2d):
    """8 x 8 blocks of 32^2 cells, 200 RK4 steps, CFL 0.5 (the C1 workload)."""
    m = M.cartesian(8, 8, 32, 32)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    gpu, orc, g, o = run_both(euler2d, m, W0, abi.Scheme(), 200)
    assert g[0] == 0 and o[0] == 0
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("C1 200 steps: W rel %.3e, %d non-bitwise of %d" % (relerr(Wg, Wo), nonbitwise(Wg, Wo), Wg.size))
    assert relerr(Wg, Wo) <= REL
    assert relerr(g[2], o[2]) <= REL


@pytest.mark.parametrize("bcs", [
    dict(W=abi.BC_REFLECTION, E=abi.BC_REFLECTION, S=abi.BC_REFLECTION, N=abi.BC_REFLECTION),
    dict(W=abi.BC_FIXED, E=abi.BC_CONSTANT_EXTRAPOLATION, S=abi.BC_REFLECTION, N=abi.BC_CONSTANT_EXTRAPOLATION),
])
def test_boundary_conditions_jittered_mesh(euler2d, bcs):
    m = M.cartesian(3, 3, 8, 8, x0=0, x1=1.5, y0=0, y1=1.0, bcs=bcs, jitter=0.1, seed=11)
    W0 = M.riemann_4quadrant(m, split=(0.75, 0.5))
    m.set_wo_from_initial(W0)
    gpu, orc, g, o = run_both(euler2d, m, W0, abi.Scheme(), 30)
    assert g[0] == 0 and o[0] == 0
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("bcs %s: rel %.3e, %d non-bitwise" % (bcs, relerr(Wg, Wo), nonbitwise(Wg, Wo)))
    assert relerr(Wg, Wo) <= REL


def test_characteristic_bc_bitwise(euler2d):
    """BC_Characteristic_Pressure (Euler2DState.cc:1383-1465) at a subsonic
    outflow calls pow(p_b/p_i, 1/gamma).  The device takes glibc's own pow
    method (glibc_pow.h), so the state stays bit for bit equal to the
    oracle, which calls the host's pow."""
    bcs = dict(W=abi.BC_FIXED, E=abi.BC_CHARACTERISTIC, S=abi.BC_REFLECTION, N=abi.BC_CONSTANT_EXTRAPOLATION)
    m = M.cartesian(3, 3, 8, 8, x0=0, x1=1.5, y0=0, y1=1.0, bcs=bcs, jitter=0.1, seed=11)
    W0 = M.riemann_4quadrant(m, split=(0.75, 0.5))
    m.set_wo_from_initial(W0)
    gpu, orc, g, o = run_both(euler2d, m, W0, abi.Scheme(), 30)
    assert g[0] == 0 and o[0] == 0
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("characteristic BC, 30 steps: rel %.3e, %d non-bitwise" % (relerr(Wg, Wo), nonbitwise(Wg, Wo)))
    assert nonbitwise(Wg, Wo) == 0
    assert relerr(g[2], o[2]) <= REL


def test_steady_bump_characteristic_outflow_bitwise(euler2d):
    """The steady bump channel of the warm-start cases (fixed inflow,
    reflecting walls, characteristic outflow) with their scheme: 5-stage
    optimal smoothing, local time stepping, CFL 1.5.  400 steps, bit for bit."""
    from cffc_amd import steady
    m, W0 = steady.bump_case(0.08, 0.45, nbi=6, nbj=2, ni=8, nj=8)
    s = steady.bump_scheme()
    gpu, orc, g, o = run_both(euler2d, m, W0, s, 400)
    assert g[0] == 0 and o[0] == 0 and g[1] == o[1] == 400
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("bump, characteristic outflow, 400 steps: L2 %.3e, %d non-bitwise" % (g[2][-1, 1], nonbitwise(Wg, Wo)))
    assert nonbitwise(Wg, Wo) == 0
    assert relerr(g[2], o[2]) <= REL


def test_local_time_stepping_and_other_schemes(euler2d):
    m = M.cartesian(2, 2, 12, 12, jitter=0.05, seed=5)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    for s in (abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING),
              abi.Scheme(i_time_integration=abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING, n_stage=5, cfl_number=0.8),
              abi.Scheme(n_stage=2, cfl_number=0.4),
              abi.Scheme(i_time_integration=abi.TIME_STEPPING_EXPLICIT_EULER, n_stage=1, cfl_number=0.3)):
        gpu, orc, g, o = run_both(euler2d, m, W0, s, 15)
        assert g[0] == o[0] == 0
        assert relerr(gpu.get_W(), orc.get_W()) <= REL, s
        assert relerr(g[2], o[2]) <= REL, s


def test_graph_replay_equals_stream_loop(euler2d):
    m = M.cartesian(4, 4, 16, 16)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    a = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
    b = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
    a.set_W(W0)
    b.set_W(W0)
    ra = a.run_steps(9)
    rb = b.run_steps(9, graph=True)
    np.testing.assert_array_equal(a.get_W(), b.get_W())
    np.testing.assert_array_equal(ra[2], rb[2])
    np.testing.assert_array_equal(ra[3], rb[3])


def test_tile_shapes_agree(euler2d):
    m = M.cartesian(2, 3, 32, 16, jitter=0.05, seed=9)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    ref = None
    for ti, tj in euler2d.TILE_SHAPES:
        g = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
        g.set_tile(ti, tj)
        g.set_W(W0)
        g.run_steps(6)
        W = g.get_W()
        if ref is None:
            ref = W
        np.testing.assert_array_equal(W, ref)


def test_autotune_keeps_state_and_results(euler2d):
    """The autotuner times every tile shape on the live state and restores
    it: a run after autotuning equals a run without it, bit for bit."""
    m = M.cartesian(3, 2, 32, 32, jitter=0.05, seed=4)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    g = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
    g.set_W(W0)
    g.run_steps(3)
    (ti, tj), ms = g.autotune(nsteps=3)
    print("autotune picked %dx%d; ms/step %s" % (ti, tj, ms))
    assert (ti, tj) in euler2d.TILE_SHAPES and g.get_tile() == (ti, tj)
    assert len(ms) == len(euler2d.TILE_SHAPES) and all(v > 0 for v in ms.values())
    assert min(ms, key=ms.get) == (ti, tj)
    g.run_steps(5)
    W = g.get_W()
    ref = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
    ref.set_W(W0)
    ref.run_steps(8)
    np.testing.assert_array_equal(W, ref.get_W())


def test_mixed_precision_error_stated(euler2d):
    m = M.cartesian(4, 4, 16, 16)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    g = euler2d.Euler2DQuadBlocks(m, abi.Scheme(flux_precision=abi.FLUX_FP32))
    g.set_W(W0)
    rc, n, norms, dts = g.run_steps(50)
    orc = oracle.Oracle(m, abi.Scheme())
    orc.set_W(W0)
    orc.run(50, nthreads=4)
    Wg, Wo = g.get_W(), orc.get_W()
    d = np.abs(Wg - Wo)
    l1 = d.mean(axis=0)
    linf = d.max(axis=0)
    print("mixed vs FP64 oracle after 50 steps: L1 %s Linf %s" % (l1, linf))
    assert rc == 0
    assert np.all(linf < 5e-3) and np.all(l1 < 1e-4)


def test_negative_density_reports_reference_error(euler2d):
    """A strong symmetric expansion (u = -/+5, p = 1e-3) drives the density
    negative in the first stage; both report the reference's code, the
    i-index of the first failing cell (Euler2DQuadSingleBlock.cc:5375-5385)."""
    m = M.cartesian(2, 2, 8, 8)
    c = m.cell_centres()
    W0 = M.uniform_state(m, (1.0, 0.0, 0.0, 1e-3))
    W0[c[:, 0] < 0.5, 1] = -5.0
    W0[c[:, 0] > 0.5, 1] = 5.0
    m.set_wo_from_initial(W0)
    g = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
    g.set_W(W0)
    rc, n, _, _ = g.run_steps(5)
    orc = oracle.Oracle(m, abi.Scheme())
    orc.set_W(W0)
    rc2, n2, _, _ = orc.run(5)
    assert rc2 > 0, "the oracle should fail on this input"
    assert rc == rc2
    assert n == n2 == 0


def test_two_level_step_reduction_large_mesh(euler2d):
    """16,384 tiles or more: the step's norms are reduced in 256 chunks
    first; norms and dt still match the oracle and the state stays bitwise."""
    m = M.cartesian(64, 32, 32, 32)
    W0 = M.riemann_4quadrant(m)
    W0 = M.density_perturbation(m, W0, amplitude=0.01, seed=3)
    m.set_wo_from_initial(W0)
    gpu, orc, g, o = run_both(euler2d, m, W0, abi.Scheme(), 2, tile=(16, 8))
    assert gpu.get_tile() == (16, 8) and m.ncell // 128 >= 16384
    assert g[0] == o[0] == 0
    np.testing.assert_array_equal(g[3], o[3])
    assert relerr(g[2], o[2]) <= REL
    assert nonbitwise(gpu.get_W(), orc.get_W()) == 0


def _axisymmetric_case():
    bcs = dict(W=abi.BC_FIXED, E=abi.BC_CONSTANT_EXTRAPOLATION, S=abi.BC_REFLECTION, N=abi.BC_REFLECTION)
    m = M.cartesian(3, 2, 8, 8, x0=0, x1=1.5, y0=0.2, y1=1.2, bcs=bcs, jitter=0.1, seed=13)
    W0 = M.riemann_4quadrant(m, split=(0.75, 0.7))
    m.set_wo_from_initial(W0)
    return m, W0


def test_axisymmetric_fp64_matches_oracle(euler2d):
    """Axisymmetric (y the radius, source (CFL*dt)*S(W, Xc) after each
    cell's east face): the fused loop and the granular reference-named
    calls agree with the oracle bit for bit; the mixed path stays close."""
    m, W0 = _axisymmetric_case()
    s = abi.Scheme(axisymmetric=1)
    gpu, orc, g, o = run_both(euler2d, m, W0, s, 30)
    assert g[0] == 0 and o[0] == 0 and g[1] == o[1] == 30
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("axisymmetric, 30 steps: rel %.3e, %d non-bitwise" % (relerr(Wg, Wo), nonbitwise(Wg, Wo)), flush=True)
    assert nonbitwise(Wg, Wo) == 0
    assert relerr(g[2], o[2]) <= REL
    gr = euler2d.Euler2DQuadBlocks(m, s)
    gr.set_W(W0)
    rc, _ = gr.step_reference_order(5)
    o5 = oracle.Oracle(m, s)
    o5.set_W(W0)
    o5.run(5)
    assert rc == 0 and nonbitwise(gr.get_W(), o5.get_W()) == 0
    mx = euler2d.Euler2DQuadBlocks(m, abi.Scheme(axisymmetric=1, flux_precision=abi.FLUX_FP32))
    mx.set_W(W0)
    rc, n, _, _ = mx.run_steps(30)
    d = np.abs(mx.get_W() - Wo)
    print("axisymmetric mixed vs FP64 oracle: L1 %s Linf %s" % (d.mean(axis=0), d.max(axis=0)), flush=True)
    assert rc == 0 and np.all(d.mean(axis=0) < 1e-4) and np.all(d.max(axis=0) < 5e-3)
    # the planar run differs: the source is really applied
    pl = oracle.Oracle(m, abi.Scheme())
    pl.set_W(W0)
    pl.run(30)
    assert np.max(np.abs(pl.get_W() - Wo)) > 1e-3


def test_axisymmetric_mesh_on_the_axis_is_refused(euler2d):
    m = M.cartesian(2, 2, 4, 4, y0=-0.5, y1=0.5)
    with pytest.raises(euler2d.CffcError, match="centroid y <= 0"):
        euler2d.Euler2DQuadBlocks(m, abi.Scheme(axisymmetric=1))


def _oracle_time_accurate(m, s, W0, t_max, time=0.0, rs=None, freeze_level=0.0):
    """The reference's time-accurate loop (Euler2DQuadSolvers.cc:526-898)
    driven through the oracle's calls, its branches as written: the clip's
    else-if uses MultiStage_Optimally_Smoothing(N, N, limiter) for every
    scheme (:662-680), the time update only for optimal smoothing
    (:884-896); Freeze_Limiters after the first step (:796-803); residual
    smoothing between the flux corrections and the update (:861-867)."""
    o = oracle.Oracle(m, s)
    o.set_W(W0)
    if rs:
        o.set_residual_smoothing(*rs)
    optsm = s.i_time_integration == abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING
    beta = oracle.multistage_optimal_smoothing(s.n_stage, s.n_stage, s.i_limiter)
    n, first, frozen = 0, True, False
    o.frozen_at = None
    while time < t_max:
        dt = o.cfl(set_global=False)
        if not optsm and time + s.cfl_number * dt > t_max:
            dt = (t_max - time) / s.cfl_number
        elif time + s.cfl_number * dt * beta > t_max:
            dt = (t_max - time) / (s.cfl_number * beta)
        o.set_global_dt(dt)
        if freeze_level > 0 and not first and not frozen and o.residual_norms(1)[1] <= freeze_level:
            o.freeze_limiters(True)
            frozen = True
            o.frozen_at = n
        for st in range(1, s.n_stage + 1):
            o.exchange_and_bcs()
            o.dudt_stage(st)
            o.apply_flux_corrections(st)
            if rs:
                o.residual_smoothing(st)
            assert o.update_stage(st)[0] == 0
        first = False
        n += 1
        time = time + s.cfl_number * dt * beta if optsm else time + s.cfl_number * dt
    return o, time, n


def _ta_case():
    m = M.cartesian(3, 3, 8, 8, jitter=0.1, seed=21)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    return m, W0


def test_time_accurate_run_reaches_t_max_like_the_oracle(euler2d):
    """Time_Accurate: the loop stops at exactly the oracle's step count and
    final time, the last step clipped to reach t_max; the state is bit for
    bit the oracle's."""
    m, W0 = _ta_case()
    t_max = 0.05
    s = abi.Scheme()
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    rc, t, n = g.run_time_accurate(t_max)
    o, t2, n2 = _oracle_time_accurate(m, s, W0, t_max)
    print("time-accurate to %g: %d / %d steps, time %.17g / %.17g, %d non-bitwise" %
          (t_max, n, n2, t, t2, nonbitwise(g.get_W(), o.get_W())), flush=True)
    assert rc == 0 and n == n2 > 3 and t == t2 and abs(t - t_max) <= 1e-15
    assert nonbitwise(g.get_W(), o.get_W()) == 0


@pytest.mark.parametrize("scheme", [
    # RK4 with the zero limiter: MultiStage_Optimally_Smoothing(4,4,0) = 2.0,
    # so the else-if shortens the steps before t_max
    dict(i_limiter=abi.LIMITER_ZERO, cfl_number=0.5),
    # 3-stage optimal smoothing with the zero limiter (factor 1.5)
    dict(i_time_integration=abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING, n_stage=3,
         i_limiter=abi.LIMITER_ZERO, cfl_number=0.5),
])
def test_time_accurate_clip_branches_like_the_reference(euler2d, scheme):
    m, W0 = _ta_case()
    s = abi.Scheme(**scheme)
    t_max = 0.04
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    rc, t, n = g.run_time_accurate(t_max)
    o, t2, n2 = _oracle_time_accurate(m, s, W0, t_max)
    print("%s: %d / %d steps, time %.17g / %.17g" % (scheme, n, n2, t, t2), flush=True)
    assert rc == 0 and n == n2 and t == t2
    assert nonbitwise(g.get_W(), o.get_W()) == 0


def test_time_accurate_with_residual_smoothing_and_freezing(euler2d):
    """Residual smoothing (step 6 of the stage loop) and Freeze_Limiters in
    the time-accurate loop, 5-stage optimal smoothing, bit for bit."""
    m, W0 = _ta_case()
    s = abi.Scheme(i_time_integration=abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING, n_stage=5, cfl_number=0.8)
    rs = (0.125, 2)
    t_max, freeze = 0.1, 0.95
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    g.set_residual_smoothing(*rs)
    rc, t, n = g.run_time_accurate(t_max, freeze_level=freeze)
    o, t2, n2 = _oracle_time_accurate(m, s, W0, t_max, rs=rs, freeze_level=freeze)
    print("smoothing + freeze: %d / %d steps, frozen at step %s, %d non-bitwise" %
          (n, n2, o.frozen_at, nonbitwise(g.get_W(), o.get_W())), flush=True)
    assert rc == 0 and n == n2 and t == t2
    assert o.frozen_at is not None and 1 < o.frozen_at < n - 1
    assert nonbitwise(g.get_W(), o.get_W()) == 0


def test_time_accurate_resumes_from_a_start_time_and_caps_steps(euler2d):
    """A run resumed at Time = t0 (a restart's time) continues to t_max like
    the reference; a step cap that ends the run first is reported."""
    m, W0 = _ta_case()
    s = abi.Scheme()
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    rc, t, n = g.run_time_accurate(0.05, time=0.03)
    o, t2, n2 = _oracle_time_accurate(m, s, W0, 0.05, time=0.03)
    assert rc == 0 and n == n2 and t == t2 and abs(t - 0.05) <= 1e-15
    assert nonbitwise(g.get_W(), o.get_W()) == 0
    g.set_W(W0)
    rc, t, n = g.run_time_accurate(0.05, max_steps=2)
    assert rc == abi.E_STEP_LIMIT and n == 2 and 0.0 < t < 0.05


def _lts_halving_case():
    """Strong symmetric expansion under scalar local time stepping at CFL 0.9:
    RK4 stage updates go unphysical in some cells and take the reference's
    dt / dUdt halving branch (SingleBlock.cc:5387-5425)."""
    m = M.cartesian(4, 1, 8, 8, x1=1.0, y1=0.25)
    c = m.cell_centres()
    W0 = np.zeros((m.ncell, 4))
    W0[:, 0], W0[:, 3] = 1.0, 0.1
    W0[:, 1] = np.where(c[:, 0] < 0.5, -3.0, 3.0)
    m.set_wo_from_initial(W0)
    return m, W0, abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING, cfl_number=0.9)


def test_local_time_stepping_halving_matches_oracle(euler2d):
    """The halving branch on the GPU: granular calls stage by stage (dt,
    dUdt[k] and W after every update, the cells halved counted from dt),
    then the fused loop, bit for bit with the oracle."""
    m, W0, s = _lts_halving_case()
    g = euler2d.Euler2DQuadBlocks(m, s)
    o = oracle.Oracle(m, s)
    g.set_W(W0)
    o.set_W(W0)
    halved = 0
    for n in range(12):
        assert g.CFL(set_global=False) == o.cfl(set_global=False)
        for st in range(1, 5):
            k = 0 if st == 4 else st - 1
            g.Send_All_Messages_and_BCs()
            o.exchange_and_bcs()
            g.dUdt_Multistage_Explicit(st)
            o.dudt_stage(st)
            dt_b = o.get_dt()
            assert g.Update_Solution_Multistage_Explicit(st) == 0
            assert o.update_stage(st)[0] == 0
            halved += int(np.sum(o.get_dt() != dt_b))
            assert nonbitwise(g.get_dt(), o.get_dt()) == 0, (n, st)
            assert nonbitwise(g.get_dUdt(k), o.get_dUdt(k)) == 0, (n, st)
            assert nonbitwise(g.get_W(), o.get_W()) == 0, (n, st)
    print("LTS halving: %d halved cell updates in 12 steps" % halved, flush=True)
    assert halved > 10
    gpu, orc, gr, orr = run_both(euler2d, m, W0, s, 12)
    assert gr[0] == 0 and orr[0] == 0
    assert nonbitwise(gpu.get_W(), orc.get_W()) == 0


def test_periodic_bc_matches_oracle(euler2d):
    """BC_PERIODIC (the in-block wrap of SingleBlock.cc:2421-2895 and the
    reduced stencils of its first ghost ring, :3207): a density wave carried
    across the periodic W/E boundaries of one block column, reflecting S/N
    walls; FP64 bit for bit with the oracle, mixed close to it."""
    P_, R_ = abi.BC_PERIODIC, abi.BC_REFLECTION
    m = M.cartesian(1, 2, 16, 8, bcs=dict(W=P_, E=P_, S=R_, N=R_), jitter=0.1, seed=13)
    c = m.cell_centres()
    W0 = M.uniform_state(m, M.freestream(0.5))
    W0[:, 0] *= 1.0 + 0.2 * np.sin(2 * np.pi * c[:, 0]) * np.cos(np.pi * c[:, 1])
    m.set_wo_from_initial(W0)
    s = abi.Scheme()
    gpu, orc, g, o = run_both(euler2d, m, W0, s, 40)
    assert g[0] == 0 and o[0] == 0
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("periodic, 40 steps: %d non-bitwise" % nonbitwise(Wg, Wo), flush=True)
    assert nonbitwise(Wg, Wo) == 0
    assert relerr(g[2], o[2]) <= REL
    gr = euler2d.Euler2DQuadBlocks(m, s)
    gr.set_W(W0)
    rc, _ = gr.step_reference_order(5)
    o5 = oracle.Oracle(m, s)
    o5.set_W(W0)
    o5.run(5)
    assert rc == 0 and nonbitwise(gr.get_W(), o5.get_W()) == 0
    mx = euler2d.Euler2DQuadBlocks(m, abi.Scheme(flux_precision=abi.FLUX_FP32))
    mx.set_W(W0)
    assert mx.run_steps(40)[0] == 0
    d = np.abs(mx.get_W() - Wo)
    assert np.all(d.max(axis=0) < 1e-4)
