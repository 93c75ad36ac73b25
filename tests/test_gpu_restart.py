"""A state read from the reference's own restart file, run on the GPU.

The 40 x 40 Ringleb-flow block of the reference's unit tests
(tests/golden/restart, a curvilinear mesh and a converged transonic state
written by CFFC) is loaded through cffc_amd.restart, its two inner ghost
layers kept, and advanced 20 RK4 steps by the HIP path and by the oracle.
The file's Ringleb-flow boundary condition (11001) is not on this path:
constant extrapolation stands in, on both sides.  FP64 must agree to
1e-12 relative (bit identity reported).
"""
import os

import numpy as np
import pytest

import oracle
from cffc_amd import abi, restart

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "restart")


def test_ringleb_restart_state_matches_oracle():
    from cffc_amd import euler2d as E
    B = restart.read_soln(os.path.join(G, "Current_Ringleb_Flow_40x40_blk000000.soln.gz"))
    m, U = restart.to_mesh([B])
    m.bc[:] = np.where(m.bc != abi.BC_NONE, abi.BC_CONSTANT_EXTRAPOLATION, abi.BC_NONE)
    s = abi.Scheme()
    gpu = E.Euler2DQuadBlocks(m, s)
    gpu.set_U(U)
    rc, n, norms, dts = gpu.run_steps(20)
    orc = oracle.Oracle(m, s)
    orc.set_U(U)
    rc2, n2, norms2, dts2 = orc.run(20, nthreads=4)
    assert rc == rc2 == 0 and n == n2 == 20
    np.testing.assert_array_equal(dts, dts2)
    Wg, Wo = gpu.get_W(), orc.get_W()
    rel = float(np.max(np.abs(Wg - Wo) / np.abs(Wo)))
    nb = int(np.sum(Wg.view(np.uint64) != Wo.view(np.uint64)))
    print("Ringleb restart: W rel %.3e, %d non-bitwise of %d; L2 %.3e -> %.3e" % (rel, nb, Wg.size, norms[0, 1],
                                                                                    norms[-1, 1]))
    assert rel <= 1e-12
    gpu.close()
