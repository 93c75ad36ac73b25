"""Steady-state solves of the bump channel on the GPU.

The steady loop is the reference's explicit loop (Euler2DQuadSolvers.cc:
549-898) with local (scalar) time stepping and limiter freezing: once the
density-residual L2 norm falls to Freeze_Limiter_Residual_Level the limiter
values of the last evaluated stage are frozen (Freeze_Limiters,
Euler2DQuadSolvers.cc:796-803) and the solve continues to the target
residual.  The residual is the reference's: the L2 norm over all interior
cells of dUdt[..][0][rho], which carries the CFL*dt scaling.
"""
import time
from dataclasses import dataclass

import numpy as np

from . import abi
from . import mesh as M


# Time marching of the steady bump runs: the reference's bump examples march
# with 5-stage multistage optimal smoothing and local time stepping
# (examples/Euler2D/bump_flow/subsonic/explicit/bump.in).  Without residual
# smoothing, 5 stages at CFL 1.6 reached 1e-10 in 51k-101k steps on eight
# sampled cases (4 stages at 1.5: 74k to over 150k); 2.0 (5 stages) and 1.8
# (4 stages) diverge.  CFL 1.5 keeps a margin below the measured limit.
STEADY_TIME_INTEGRATION = abi.TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING
STEADY_N_STAGE = 5
STEADY_CFL = 1.5
STEADY_FREEZE_LEVEL = 1e-5


def bump_scheme(flux_precision=abi.FLUX_FP64, cfl=STEADY_CFL, time_integration=STEADY_TIME_INTEGRATION,
                n_stage=STEADY_N_STAGE):
    """Least squares, Venkatakrishnan, Roe, scalar local time stepping."""
    return abi.Scheme(i_time_integration=time_integration, n_stage=n_stage,
                      local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING, cfl_number=cfl,
                      flux_precision=flux_precision)


@dataclass
class SteadyResult:
    rc: int
    steps: int
    seconds: float
    l2: float
    norms: np.ndarray
    frozen_at: int


def solve_steady_host_loop(blocks, tol=1e-10, max_steps=100000, freeze_level=STEADY_FREEZE_LEVEL):
    """The reference's order of calls: run until the L2 density residual
    reaches freeze_level, Freeze_Limiters, run on to tol (host-driven)."""
    t0 = time.perf_counter()
    blocks.Freeze_Limiters(False)
    blocks.store_limiters(True)
    rc, n1, norms1, _ = blocks.run_steps(max_steps, tol=freeze_level)
    frozen_at = -1
    norms = norms1
    n = n1
    if rc == 0 and n1 < max_steps and freeze_level > tol:
        blocks.Freeze_Limiters(True)
        frozen_at = n1
        rc, n2, norms2, _ = blocks.run_steps(max_steps - n1, tol=tol)
        norms = np.concatenate([norms1, norms2])
        n = n1 + n2
    blocks.synchronize()
    dt = time.perf_counter() - t0
    blocks.store_limiters(False)
    blocks.Freeze_Limiters(False)
    l2 = float(norms[-1, 1]) if len(norms) else float("nan")
    return SteadyResult(rc=rc, steps=n, seconds=dt, l2=l2, norms=norms, frozen_at=frozen_at)


def solve_steady(blocks, tol=1e-10, max_steps=100000, freeze_level=STEADY_FREEZE_LEVEL, history=True):
    """Run blocks (an Euler2DQuadBlocks holding the initial state) to an L2
    density residual <= tol, freezing the limiters at freeze_level, all on
    the device (cffc_run_steady, one part).  steps counts all time steps
    taken (the reference's number_of_time_steps); -1 if not converged."""
    t0 = time.perf_counter()
    rc, n, steps, l2, hist = blocks.run_steady(max_steps, tol, freeze_level, history=history)
    dt = time.perf_counter() - t0
    norms = hist if hist is not None else np.zeros((0, 1))
    frozen = -1
    if hist is not None and freeze_level > 0:
        below = np.nonzero(hist[:, 0] <= freeze_level)[0]
        frozen = int(below[0]) + 1 if len(below) else -1
    return SteadyResult(rc=rc, steps=int(steps[0]), seconds=dt, l2=float(l2[0]), norms=norms, frozen_at=frozen)


@dataclass
class BatchResult:
    rc: int
    steps_done: int
    part_steps: np.ndarray
    part_l2: np.ndarray
    seconds: float
    states: list
    history: np.ndarray
    part_seconds: np.ndarray   # device time from the start to each part's convergence (nan: not converged)


def solve_batch(cases, scheme=None, tol=1e-10, max_steps=200000, freeze_level=STEADY_FREEZE_LEVEL, history=False,
                segment=0, log=None, stall_window=0):
    """Solve the cases [(mesh, W0), ...] together: one mesh of disjoint
    parts, one launch per stage for all of them, each part converging (and
    freezing its limiters) on its own.  segment > 0 drives the run in
    resumed segments of that many steps (same result), logging progress
    after each.  stall_window > 0 also freezes a part's limiters after that
    many steps without a new L2 minimum 1% below the best (a serving policy; the reference
    freezes at freeze_level only).  Returns a BatchResult with the steps of
    every case and the final state arrays."""
    from . import euler2d
    meshes = [m for m, _ in cases]
    mc, starts = M.concat(meshes)
    W0 = np.concatenate([W for _, W in cases])
    g = euler2d.Euler2DQuadBlocks(mc, scheme or bump_scheme())
    try:
        g.set_parts(starts + [mc.nblk])
        if stall_window:
            g.set_stall_freeze(stall_window)
        g.set_W(W0)
        g.synchronize()
        t0 = time.perf_counter()
        if segment <= 0 or history:
            rc, n, steps, l2, hist = g.run_steady(max_steps, tol, freeze_level, history=history)
            clock = g.step_times(n) if n > 0 else np.zeros(0)
            part_s = np.array([clock[k - 1] if k > 0 else np.nan for k in steps])
        else:
            n, hist, part_s, rc = 0, None, None, 0
            while n < max_steps:
                seg = min(segment, max_steps - n)
                rc, nd, steps, l2, _ = g.run_steady(seg, tol, freeze_level, resume=n > 0)
                clock = g.step_times(nd) if nd > 0 else np.zeros(0)
                base = time.perf_counter() - t0 - (clock[-1] if nd else 0.0)
                if part_s is None:
                    part_s = np.full(len(steps), np.nan)
                for q, k in enumerate(steps):
                    if k > n and k <= n + nd:
                        part_s[q] = base + clock[k - n - 1]
                n += nd
                if log:
                    log("  %d steps: %d of %d parts converged, max l2 %.3e, %.1f s" %
                        (n, int(np.sum(steps > 0)), len(steps), float(np.max(l2)), time.perf_counter() - t0))
                if rc != 0 or np.all(steps > 0) or nd < seg:
                    break
        dt = time.perf_counter() - t0
        W = g.get_W()
    finally:
        g.close()
    offs = np.cumsum([0] + [m.ncell for m in meshes])
    states = [W[offs[q]:offs[q + 1]].copy() for q in range(len(meshes))]
    return BatchResult(rc=rc, steps_done=n, part_steps=steps, part_l2=l2, seconds=dt, states=states, history=hist,
                       part_seconds=part_s)


def solve_batch_fp32_then_fp64(cases, tol=1e-10, switch_level=1e-5, max_steps=200000,
                               freeze_level=STEADY_FREEZE_LEVEL, segment=0, log=None, stall_window=0):
    """Mixed-precision steady batch: every part first marches with FP32
    reconstruction/limiter/Roe flux (FP64 state and update) down to
    switch_level, above the FP32 round-off floor of the residual, then the
    FP64 path takes the conserved state and finishes to tol (limiters frozen
    at freeze_level as usual).  Steps and latencies of a part add over the
    two phases; the second phase starts when the first has ended for all
    parts.  stall_window applies to the FP64 phase (solve_batch)."""
    first = solve_batch(cases, scheme=bump_scheme(flux_precision=abi.FLUX_FP32), tol=switch_level,
                        max_steps=max_steps, freeze_level=0.0, segment=segment, log=log)
    if first.rc != 0:
        return first, None
    # the FP64 phase restarts from the FP64 state the first phase ended in
    meshes = [m for m, _ in cases]
    second = solve_batch([(m, W) for m, W in zip(meshes, first.states)], tol=tol,
                         max_steps=max_steps, freeze_level=freeze_level, segment=segment, log=log,
                         stall_window=stall_window)
    return first, second


def bump_case(h, mach, outflow=abi.BC_CHARACTERISTIC, **kw):
    """Mesh, freestream initial state and boundary states of one bump case.
    The outflow is the reference's characteristic (pressure) condition: with
    the fixed inflow and a constant-extrapolation outflow the subsonic
    channel's pressure level is left free and its residual stalls on a
    uniform drift instead of converging to 1e-10."""
    m = M.bump_channel(h, outflow=outflow, **kw)
    W0 = M.uniform_state(m, M.freestream(mach))
    m.set_wo_from_initial(W0)
    return m, W0


def sample_cases(n, seed):
    """(h, M) pairs, h ~ U[0.04, 0.10], M ~ U[0.3, 0.6], seeded."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.04, 0.10, n), rng.uniform(0.3, 0.6, n)], axis=1)
