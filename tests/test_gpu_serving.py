"""The C5 serving batch at full size on the GPU: 256 bump cases of 96 x 32
to 384 x 128 cells (blocks of 32^2, 6,008,832 cells), solved as one mesh of
disjoint parts by the FAS multigrid solver (4 levels, V cycle, CFL 1.0,
FP64), as the bench's serving leg and cffc_amd.serving do.

Batching must not couple the cases: multigrid works block by block with
local time steps, so after the same number of cycles every case of the
batch holds exactly the state it reaches when solved alone (checked bit
for bit on three cases of different sizes).  The norms of the batch only
decide when to stop, and here the cycle count is fixed."""
import numpy as np
import pytest

from cffc_amd import multigrid, serving, steady
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu


def nonbitwise(a, b):
    return int(np.sum(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64)))


def test_c5_256_case_multigrid_batch_equals_cases_alone():
    cases = serving.serving_cases(256, 5)
    batch = [serving.case_mesh(h, m, k) for h, m, k in cases]
    mesh, _ = M.concat([b[0] for b in batch])
    W0 = np.concatenate([b[1] for b in batch])
    assert mesh.ncell == 6008832
    params = multigrid.MGParams(levels=4)   # no tolerance: every cycle runs
    scheme = steady.bump_scheme(cfl=1.0)
    ncyc = 40
    g = multigrid.FASMultigrid(mesh, scheme, params)
    g.set_W(W0)
    g.initialize()
    rc, n, norms = g.execute(ncyc)
    W = g.get_W()
    g.close()
    # the first check precedes any residual evaluation (R = 0); the next
    # ones follow cycles 1, 2, ...
    print("256 cases, %d cells: %d cycles, batch R %.3e after cycle 1, %.3e after cycle %d"
          % (mesh.ncell, n, norms[1, 1], norms[-1, 1], n), flush=True)
    assert rc == 0 and n == ncyc
    assert np.all(np.isfinite(norms)) and np.all(np.isfinite(W))
    offs = np.cumsum([0] + [b[0].ncell for b in batch])
    picks = [int(np.argmin([k for _, _, k in cases])), 97, int(np.argmax([k for _, _, k in cases]))]
    for q in picks:
        m1, W1 = batch[q]
        s = multigrid.FASMultigrid(m1, scheme, params)
        s.set_W(W1)
        s.initialize()
        rc1, n1, _ = s.execute(ncyc)
        Wq = s.get_W()
        s.close()
        assert rc1 == 0 and n1 == ncyc
        nb = nonbitwise(W[offs[q]:offs[q + 1]], Wq)
        print("case %d (k = %d, %d cells): %d values of the batch differ from the case alone"
              % (q, cases[q][2], m1.ncell, nb), flush=True)
        assert nb == 0
