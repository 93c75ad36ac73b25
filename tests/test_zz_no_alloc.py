"""cffc_run_steps_noalloc (the fused loop with its per-step records in the
library's fixed device array) against cffc_run_steps, through the C ABI:
two calls in a row on two contexts, the second on a smaller mesh with fewer
steps, each compared bit for bit (state, per-step norms and time steps)
with cffc_run_steps from the same initial state.  Also the limits: more
than 4096 steps and a run with kernel timing on are refused."""
import numpy as np
import pytest

from cffc_amd import abi
from cffc_amd import euler2d as E
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu


def _case(nb, nc, seed):
    m = M.cartesian(nb, nb, nc, nc, jitter=0.05, seed=seed)
    W0 = M.riemann_4quadrant(m)
    W0 = M.density_perturbation(m, W0, amplitude=0.01, seed=seed)
    m.set_wo_from_initial(W0)
    return m, W0


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _compare(m, W0, nsteps, scheme):
    ref = E.Euler2DQuadBlocks(m, scheme)
    ref.set_W(W0)
    rc_r, n_r, norms_r, dt_r = ref.run_steps(nsteps)
    W_r = ref.get_W()
    ref.close()
    g = E.Euler2DQuadBlocks(m, scheme)
    g.set_W(W0)
    rc, n, norms, dts = g.run_steps_noalloc(nsteps)
    W = g.get_W()
    g.close()
    print("noalloc: %d cells, %d steps: rc %d/%d, %d non-bitwise W values"
          % (m.ncell, nsteps, rc, rc_r, int(np.sum(_bits(W) != _bits(W_r)))), flush=True)
    assert rc == rc_r == 0 and n == n_r == nsteps
    assert np.array_equal(_bits(W), _bits(W_r))
    assert np.array_equal(_bits(norms), _bits(norms_r))
    assert np.array_equal(_bits(dts), _bits(dt_r))


@pytest.mark.parametrize("precision", [abi.FLUX_FP64, abi.FLUX_FP32])
def test_noalloc_run_equals_run_steps_twice(precision):
    s = abi.Scheme(flux_precision=precision)
    m1, W1 = _case(4, 16, 11)
    _compare(m1, W1, 40, s)
    # second call, smaller input: the shared array keeps the first call's
    # records, which must not leak into this one
    m2, W2 = _case(2, 8, 12)
    _compare(m2, W2, 12, s)


def test_noalloc_limits():
    m, W0 = _case(2, 8, 13)
    g = E.Euler2DQuadBlocks(m, abi.Scheme())
    g.set_W(W0)
    with pytest.raises(E.CffcError, match="4096"):
        g.run_steps_noalloc(4097)
    g.kernel_timing(True)
    with pytest.raises(E.CffcError, match="timing"):
        g.run_steps_noalloc(2)
    g.kernel_timing(False)
    assert g.run_steps_noalloc(2)[0] == 0
    g.close()
