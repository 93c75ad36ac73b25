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


def test_axisymmetric_checkpoint_resume_continues_the_run(tmp_path):
    """Checkpoint an axisymmetric run after 6 steps, resume it with the
    scheme taken from the file (its Axisymmetric flag) and run 6 more: the
    state follows the uninterrupted 12-step run (the restart text keeps 15
    significant digits, hence a tolerance, not bit identity), while the same
    file resumed as a planar run departs from it."""
    from cffc_amd import euler2d as E
    from cffc_amd import mesh as M
    m = M.cartesian(2, 2, 8, 8, y0=0.2, y1=1.2, jitter=0.05, seed=4)
    W0 = M.uniform_state(m, M.freestream(0.3))
    c = m.cell_centres()
    bump = np.exp(-((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.7) ** 2) / 0.02)
    # a pressure bump drives a radial velocity, so the source terms act
    W0[:, 0] *= 1.0 + 0.1 * bump
    W0[:, 3] *= 1.0 + 0.2 * bump
    m.set_wo_from_initial(W0)
    s = abi.Scheme(axisymmetric=1)
    ref = E.Euler2DQuadBlocks(m, s)
    ref.set_W(W0)
    assert ref.run_steps(12)[0] == 0
    a = E.Euler2DQuadBlocks(m, s)
    a.set_W(W0)
    assert a.run_steps(6)[0] == 0
    prefix = str(tmp_path / "axi")
    restart.checkpoint(prefix, a, nsteps=6)
    a.close()
    g, n, _ = restart.resume(prefix)
    assert g.scheme.axisymmetric == 1 and n == 6
    assert g.run_steps(6)[0] == 0
    Ur, Ug = ref.get_U(), g.get_U()
    # per-variable scale: the transverse momentum is exactly zero in places
    scale = np.maximum(np.abs(Ur), 1e-3 * np.abs(Ur).max(axis=0))
    rel = float(np.max(np.abs(Ug - Ur) / scale))
    # the file's state and mesh advanced as a planar flow (what resume() did
    # before it read the flag): clearly not the same run
    _, _, m2, U2 = restart.read_restart(prefix)
    planar = E.Euler2DQuadBlocks(m2, abi.Scheme())
    planar.set_U(U2)
    assert planar.run_steps(6)[0] == 0
    rel_planar = float(np.max(np.abs(planar.get_U() - Ur) / scale))
    print("axisymmetric resume: rel %.3e (resumed as planar: %.3e)" % (rel, rel_planar))
    assert rel <= 1e-10 and rel_planar > 1e-6
    for x in (ref, g, planar):
        x.close()
