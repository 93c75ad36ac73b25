"""Implicit residual smoothing on the GPU (Residual_Smoothing,
Euler2DQuadSingleBlock.cc:4521-4564) against the oracle, bit for bit: the
granular call stage by stage, the fused time loop (residual, smoothing and
update as three launches per stage) and batched steady runs."""
import numpy as np
import pytest

import oracle
from cffc_amd import abi, steady
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu

OPT5 = abi.Scheme(i_time_integration=abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING, n_stage=5, cfl_number=1.0)


@pytest.fixture(scope="module")
def euler2d():
    from cffc_amd import euler2d as E
    E.load()
    return E


def _case(nbi=3, nbj=2, n=12, jitter=0.08, seed=6):
    m = M.cartesian(nbi, nbj, n, n, jitter=jitter, seed=seed)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    return m, W0


def test_granular_smoothing_stage_by_stage(euler2d):
    m, W0 = _case()
    gpu = euler2d.Euler2DQuadBlocks(m, OPT5)
    orc = oracle.Oracle(m, OPT5)
    gpu.set_residual_smoothing(0.125, 2)
    orc.set_residual_smoothing(0.125, 2)
    gpu.set_W(W0)
    orc.set_W(W0)
    for n in range(2):
        assert gpu.CFL() == orc.cfl()
        for st in range(1, 6):
            gpu.Send_All_Messages_and_BCs()
            orc.exchange_and_bcs()
            gpu.dUdt_Multistage_Explicit(st)
            orc.dudt_stage(st)
            gpu.Residual_Smoothing(st)
            orc.residual_smoothing(st)
            np.testing.assert_array_equal(gpu.get_dUdt(0), orc.get_dUdt(0))
            np.testing.assert_array_equal(gpu.get_dUdt(1), orc.get_dUdt(1))
            assert gpu.Update_Solution_Multistage_Explicit(st) == 0
            assert orc.update_stage(st)[0] == 0
            np.testing.assert_array_equal(gpu.get_W(), orc.get_W())


@pytest.mark.parametrize("lts", [abi.GLOBAL_TIME_STEPPING, abi.SCALAR_LOCAL_TIME_STEPPING])
def test_fused_loop_with_smoothing(euler2d, lts):
    m, W0 = _case(4, 3, 16)
    s = abi.Scheme(i_time_integration=OPT5.i_time_integration, n_stage=5, cfl_number=1.2, local_time_stepping=lts)
    gpu = euler2d.Euler2DQuadBlocks(m, s)
    gpu.set_residual_smoothing(0.25, 2)
    gpu.set_W(W0)
    rc, n, norms, dts = gpu.run_steps(15)
    orc = oracle.Oracle(m, s)
    orc.set_residual_smoothing(0.25, 2)
    orc.set_W(W0)
    rc2, n2, norms2, dts2 = orc.run(15, nthreads=4)
    assert rc == rc2 == 0 and n == n2 == 15
    Wg, Wo = gpu.get_W(), orc.get_W()
    print("smoothing fused loop: %d non-bitwise" % int(np.sum(Wg.view(np.uint64) != Wo.view(np.uint64))))
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_allclose(norms, norms2, rtol=1e-12)
    # the smoothing changes the solution (the test is not vacuous)
    ref = euler2d.Euler2DQuadBlocks(m, s)
    ref.set_W(W0)
    ref.run_steps(15)
    assert not np.array_equal(ref.get_W(), Wg)


def test_rk4_refused(euler2d):
    m, W0 = _case(2, 1, 8)
    g = euler2d.Euler2DQuadBlocks(m, abi.Scheme())
    with pytest.raises(euler2d.CffcError):
        g.set_residual_smoothing(0.125, 2)


def test_batched_steady_with_smoothing_equals_single(euler2d):
    sch = steady.bump_scheme()
    cases = [steady.bump_case(0.06, 0.4, nbi=6, nbj=2, ni=8, nj=8), steady.bump_case(0.09, 0.55, nbi=6, nbj=2, ni=8, nj=8)]
    singles = []
    for mm, W0 in cases:
        g = euler2d.Euler2DQuadBlocks(mm, sch)
        g.set_residual_smoothing(0.125, 2)
        g.set_W(W0)
        r = steady.solve_steady(g, tol=1e-6, max_steps=20000, freeze_level=1e-3, history=False)
        singles.append((r.steps, g.get_W()))
        g.close()
    mc, starts = M.concat([mm for mm, _ in cases])
    g = euler2d.Euler2DQuadBlocks(mc, sch)
    g.set_residual_smoothing(0.125, 2)
    g.set_parts(starts + [mc.nblk])
    g.set_W(np.concatenate([W for _, W in cases]))
    rc, n, steps, l2, _ = g.run_steady(20000, 1e-6, 1e-3)
    W = g.get_W()
    print("single", [x[0] for x in singles], "batch", steps.tolist())
    assert rc == 0 and steps.tolist() == [x[0] for x in singles]
    assert np.array_equal(W[:cases[0][0].ncell], singles[0][1])
    assert np.array_equal(W[cases[0][0].ncell:], singles[1][1])
