"""Batched steady runs (cffc_set_parts / cffc_run_steady) on the GPU.

The device-driven steady loop must take exactly the steps of the reference's
host-driven order of calls (run to Freeze_Limiter_Residual_Level,
Freeze_Limiters, run on; Euler2DQuadSolvers.cc:796-803), and a case must
converge in the same steps to the same state whether it runs alone or as
one part of a batch.
"""
import numpy as np
import pytest

from cffc_amd import abi, euler2d, steady
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu

SMALL = dict(nbi=6, nbj=2, ni=8, nj=8)   # 48 x 16 cells


def _case(h, mach, **kw):
    return steady.bump_case(h, mach, **SMALL, **kw)


def test_device_loop_equals_host_loop():
    m, W0 = _case(0.08, 0.5)
    s = steady.bump_scheme()
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    ref = steady.solve_steady_host_loop(g, tol=1e-7, max_steps=20000, freeze_level=1e-3)
    W_ref = g.get_W()
    g.close()
    assert ref.rc == 0 and ref.l2 <= 1e-7 and ref.frozen_at > 0
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    r = steady.solve_steady(g, tol=1e-7, max_steps=20000, freeze_level=1e-3)
    W = g.get_W()
    g.close()
    print("host loop %d steps (frozen at %d), device loop %d steps" % (ref.steps, ref.frozen_at, r.steps))
    assert r.rc == 0
    assert r.steps == ref.steps
    assert r.frozen_at == ref.frozen_at
    np.testing.assert_array_equal(r.norms[:, 0], ref.norms[:, 1])
    assert np.array_equal(W, W_ref)


@pytest.mark.parametrize("n_stage,ti", [(4, abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING),
                                        (5, abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING)])
def test_batch_equals_single(n_stage, ti):
    """Odd stage counts make converged parts copy their state forward."""
    s = steady.bump_scheme(time_integration=ti, n_stage=n_stage, cfl=1.0)
    cases = [(0.05, 0.35), (0.09, 0.55), (0.07, 0.45)]
    singles = []
    for h, mach in cases:
        m, W0 = _case(h, mach)
        g = euler2d.Euler2DQuadBlocks(m, s)
        g.set_W(W0)
        r = steady.solve_steady(g, tol=1e-6, max_steps=20000, freeze_level=1e-3, history=False)
        singles.append((r.steps, g.get_W()))
        g.close()
        assert r.rc == 0 and r.steps > 0
    b = steady.solve_batch([_case(h, mach) for h, mach in cases], scheme=s, tol=1e-6, max_steps=20000,
                           freeze_level=1e-3)
    print("single steps", [x[0] for x in singles], "batch", b.part_steps.tolist())
    assert b.rc == 0
    assert b.part_steps.tolist() == [x[0] for x in singles]
    assert b.steps_done == max(x[0] for x in singles)
    for (n, W), Wb in zip(singles, b.states):
        assert np.array_equal(W, Wb)


def test_parts_must_be_disjoint():
    m, W0 = _case(0.06, 0.4)
    g = euler2d.Euler2DQuadBlocks(m, steady.bump_scheme())
    with pytest.raises(euler2d.CffcError):
        g.set_parts([0, 3, m.nblk])   # cuts through a block row: neighbours in two parts
    g.close()


def test_unconverged_parts_report_minus_one():
    b = steady.solve_batch([_case(0.08, 0.5), _case(0.05, 0.3)], tol=1e-12, max_steps=50, freeze_level=0.0,
                           history=True)
    assert b.rc == 0 and b.steps_done == 50
    assert b.part_steps.tolist() == [-1, -1]
    assert b.history.shape == (50, 2)
    np.testing.assert_array_equal(b.history[-1], b.part_l2)


def test_segments_equal_one_run():
    """A steady run driven in resumed segments takes the same steps to the
    same state as one call."""
    cases = [_case(0.06, 0.4), _case(0.09, 0.58)]
    one = steady.solve_batch(cases, tol=1e-6, max_steps=20000, freeze_level=1e-3)
    seg = steady.solve_batch(cases, tol=1e-6, max_steps=20000, freeze_level=1e-3, segment=700)
    print("one call", one.part_steps.tolist(), "segments", seg.part_steps.tolist())
    assert one.rc == 0 and seg.rc == 0
    assert seg.part_steps.tolist() == one.part_steps.tolist()
    assert seg.steps_done == one.steps_done
    for a, b in zip(one.states, seg.states):
        assert np.array_equal(a, b)


def test_fp32_flux_then_fp64_finish():
    """FP32 reconstruction/flux down to 1e-5 (above its round-off floor),
    then the FP64 path to the tolerance: every part converges, and the steady
    states agree with an FP64-only solve to the level the frozen limiters
    allow."""
    cases = [_case(0.05, 0.35), _case(0.09, 0.55)]
    first, second = steady.solve_batch_fp32_then_fp64(cases, tol=1e-8, switch_level=1e-5, max_steps=40000)
    assert first.rc == 0 and second is not None and second.rc == 0
    assert np.all(first.part_steps > 0) and np.all(second.part_steps > 0)
    ref = steady.solve_batch(cases, tol=1e-8, max_steps=40000)
    print("fp32 steps", first.part_steps.tolist(), "fp64 finish", second.part_steps.tolist(),
          "fp64 only", ref.part_steps.tolist())
    for a, b in zip(second.states, ref.states):
        assert np.max(np.abs(a - b)) < 1e-3


def _stall_step(hist, window, freeze_level):
    """The step at which the stall rule fires on a part's L2 history (the
    finish kernel's rule restated): the first step, while the part runs,
    that comes window steps after its last new minimum (1% below the best
    before it); None if never."""
    best, best_step = np.inf, 0
    for s, l2 in enumerate(hist):
        if freeze_level > 0.0 and l2 <= freeze_level:
            return None
        if l2 < 0.99 * best:
            best, best_step = l2, s
        elif s - best_step >= window:
            return s
    return None


def test_stall_freeze_fires_where_the_rule_says():
    """cffc_set_stall_freeze: up to and including a part's stall step the
    run is the reference's bit for bit (equal L2 histories); the next step
    loads the frozen limiters, so the part's history leaves the
    reference's; parts that never stall are untouched."""
    cases = [_case(0.08, 0.5), _case(0.05, 0.3), _case(0.09, 0.58)]
    window, n = 40, 1500
    a = steady.solve_batch(cases, tol=1e-14, max_steps=n, freeze_level=0.0, history=True)
    b = steady.solve_batch(cases, tol=1e-14, max_steps=n, freeze_level=0.0, history=True, stall_window=window)
    assert a.rc == 0 and b.rc == 0 and a.history.shape == b.history.shape == (n, 3)
    fired = 0
    for q in range(3):
        f = _stall_step(a.history[:, q], window, 0.0)
        print("part %d: stall step %s" % (q, f), flush=True)
        if f is None:
            np.testing.assert_array_equal(b.history[:, q], a.history[:, q])
            continue
        fired += 1
        np.testing.assert_array_equal(b.history[:f + 1, q], a.history[:f + 1, q])
        assert np.any(b.history[f + 1:, q] != a.history[f + 1:, q])
    assert fired >= 1
