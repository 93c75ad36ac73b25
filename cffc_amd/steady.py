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


def bump_scheme(flux_precision=abi.FLUX_FP64, cfl=0.8):
    """RK4, Roe, least squares, Venkatakrishnan, scalar local time stepping."""
    return abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING, cfl_number=cfl,
                      flux_precision=flux_precision)


@dataclass
class SteadyResult:
    rc: int
    steps: int
    seconds: float
    l2: float
    norms: np.ndarray
    frozen_at: int


def solve_steady(blocks, tol=1e-10, max_steps=100000, freeze_level=1e-5):
    """Run blocks (an Euler2DQuadBlocks holding the initial state) to an L2
    density residual <= tol.  Returns a SteadyResult; steps counts all time
    steps taken (the reference's number_of_time_steps)."""
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
    l2 = float(norms[-1, 1]) if len(norms) else float("nan")
    return SteadyResult(rc=rc, steps=n, seconds=dt, l2=l2, norms=norms, frozen_at=frozen_at)


def bump_case(h, mach, **kw):
    """Mesh, freestream initial state and boundary states of one bump case."""
    m = M.bump_channel(h, **kw)
    W0 = M.uniform_state(m, M.freestream(mach))
    m.set_wo_from_initial(W0)
    return m, W0


def sample_cases(n, seed):
    """(h, M) pairs, h ~ U[0.04, 0.10], M ~ U[0.3, 0.6], seeded."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.04, 0.10, n), rng.uniform(0.3, 0.6, n)], axis=1)
