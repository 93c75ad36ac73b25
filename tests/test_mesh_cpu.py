"""Mesh generators and the multi-case concatenation, on the CPU oracle."""
import numpy as np

import oracle
from cffc_amd import abi
from cffc_amd import mesh as M


def test_bump_geometry():
    h = 0.08
    m = M.bump_channel(h)
    assert m.nblk == 48 and m.ncell == 192 * 64
    x, y = m.xnode, m.ynode
    # the bump crest (x = 1.5) sits at y = h on the lower wall
    b = 1 * 12 + 0  # any block; check the global wall node at x = 1.5
    c = m.cell_centres()
    assert c[:, 0].min() > 0 and c[:, 0].max() < 3 and c[:, 1].max() < 1
    R = m.info["R"]
    assert abs(R - (0.25 + h * h) / (2 * h)) < 1e-15


def test_concatenated_cases_solve_independently():
    """Two cases in one mesh (disjoint parts) evolve exactly as when solved
    alone under local time stepping (no global coupling)."""
    s = abi.Scheme(local_time_stepping=abi.SCALAR_LOCAL_TIME_STEPPING, cfl_number=0.8)
    parts, W0s = [], []
    for h, mach in ((0.05, 0.4), (0.09, 0.55)):
        m = M.bump_channel(h, nbi=4, nbj=2, ni=8, nj=8)
        W0 = M.uniform_state(m, M.freestream(mach))
        m.set_wo_from_initial(W0)
        parts.append(m)
        W0s.append(W0)
    mc, starts = M.concat(parts)
    oc = oracle.Oracle(mc, s)
    oc.set_W(np.concatenate(W0s))
    oc.run(25)
    Wc = oc.get_W()
    off = 0
    for m, W0 in zip(parts, W0s):
        o = oracle.Oracle(m, s)
        o.set_W(W0)
        o.run(25)
        np.testing.assert_array_equal(o.get_W(), Wc[off:off + m.ncell])
        off += m.ncell
