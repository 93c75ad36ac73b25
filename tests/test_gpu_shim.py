"""The binding with the reference's signatures, compiled and run on the GPU.

tests/native/euler2d_hip_shim.cpp is the C++ a CFFC maintainer adds next to
Euler2DQuadMultiBlock.cc: CFL, Set_Global_TimeStep, L2_Norm_Residual,
Send_All_Messages, BCs, dUdt_Multistage_Explicit,
Apply_Boundary_Flux_Corrections_Multistage_Explicit and
Update_Solution_Multistage_Explicit with CFFC's argument lists, over CFFC's
block types.  tests/native/shim_c1 builds the blocks and their adaptive
block list from a case file and runs the reference's explicit loop through
those calls.  The 4-quadrant Riemann problem on 8 x 8 blocks of 32^2 cells,
200 RK4 steps at CFL 0.5, must equal the oracle's run bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from cffc_amd import abi
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "native", "shim_c1")


def write_case(path, m, W0, n_stage, cfl):
    level = m.level if m.level is not None else np.zeros(m.nblk, np.int32)
    with open(path, "wb") as f:
        np.array([m.nblk, n_stage], np.int32).tofile(f)
        np.array([cfl], np.float64).tofile(f)
        for a in (m.ni, m.nj, m.nbr.reshape(-1), level, m.bc):
            np.ascontiguousarray(a, np.int32).tofile(f)
        for a in (m.xnode, m.ynode, m.wo.reshape(-1), np.asarray(W0).reshape(-1)):
            np.ascontiguousarray(a, np.float64).tofile(f)


def test_reference_signatures_run_c1_like_the_oracle(tmp_path):
    import oracle
    if not os.path.exists(SHIM):
        pytest.fail("tests/native/shim_c1 is not built (make -C tests/native)")
    m = M.cartesian(8, 8, 32, 32)
    W0 = M.riemann_4quadrant(m)
    m.set_wo_from_initial(W0)
    case, out = str(tmp_path / "c1.bin"), str(tmp_path / "w.bin")
    write_case(case, m, W0, 4, 0.5)
    nsteps = 200
    r = subprocess.run([SHIM, case, str(nsteps), out], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = np.fromfile(out, np.float64)
    Wg = res[:4 * m.ncell].reshape(m.ncell, 4)
    l2g = res[4 * m.ncell:]
    o = oracle.Oracle(m, abi.Scheme(cfl_number=0.5))
    o.set_W(W0)
    rc, n, norms, _ = o.run(nsteps, nthreads=min(16, os.cpu_count() or 1))
    assert rc == 0 and n == nsteps
    Wo = o.get_W()
    nb = int(np.sum(Wg.view(np.uint64) != Wo.view(np.uint64)))
    print("shim, C1, %d steps: %d of %d values not bit-identical" % (nsteps, nb, Wg.size), flush=True)
    assert nb == 0
    assert len(l2g) == nsteps
    assert np.allclose(l2g, norms[:, 1], rtol=1e-12, atol=0.0)
