"""The reference's restart files (CPU).

Fixtures: the restart files of the reference's own Ringleb-flow unit tests
(src_2D/Euler2D/UnitTests/SimulationsData/RinglebFlow and QuadBlockData),
committed gzip-compressed under tests/golden/restart: a 40 x 40 block with
5 ghost layers (state written with 15 digits, zeroth-order cell moments), a
30 x 30 grid block with third-order moments (16 digits), and the quadtree
file of the first.  read_soln / format_soln and read_tree / format_tree must
give back these files byte for byte; meshes written by this package must
read back to the same mesh, topology and state.
"""
import gzip
import os

import numpy as np
import pytest

from cffc_amd import abi, restart
from cffc_amd import mesh as M

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "restart")
FLOW = os.path.join(G, "Current_Ringleb_Flow_40x40_blk000000.soln.gz")
GRID = os.path.join(G, "Master_Ringleb_Grid_blk000000.soln.gz")
TREE = os.path.join(G, "Current_Ringleb_Flow_40x40_quadtree.tree")


@pytest.mark.parametrize("path", [FLOW, GRID])
def test_soln_round_trip_is_byte_exact(path):
    B = restart.read_soln(path)
    with gzip.open(path, "rt") as f:
        text = f.read()
    assert restart.format_soln(B) == text


def test_ringleb_block_contents():
    B = restart.read_soln(FLOW)
    assert (B.NCi, B.NCj, B.nghost, B.ICl, B.ICu) == (50, 50, 5, 5, 44)
    assert B.nsteps == 7500 and abs(B.time - 0.601621006594792) < 1e-15 and B.gas == "AIR"
    assert B.xn.shape == (51, 51) and B.U.shape == (50, 50, 4)
    # a consistent state: positive density and pressure in every cell
    U = B.U.reshape(-1, 4)
    p = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / U[:, 0])
    assert np.all(U[:, 0] > 0) and np.all(p > 0)
    # the cell centroids and areas in the file are those of its nodes
    c = B.cells
    for j, i in [(5, 5), (20, 31), (44, 44), (0, 0)]:
        x = [B.xn[j, i], B.xn[j, i + 1], B.xn[j + 1, i + 1], B.xn[j + 1, i]]
        y = [B.yn[j, i], B.yn[j, i + 1], B.yn[j + 1, i + 1], B.yn[j + 1, i]]
        xc, yc, a = restart._centroid_area(x, y)
        assert abs(xc - c[j, i, 0]) < 1e-12 and abs(yc - c[j, i, 1]) < 1e-12 and abs(a / c[j, i, 2] - 1) < 1e-12


def test_tree_round_trip_is_byte_exact():
    T = restart.read_tree(TREE)
    with open(TREE) as f:
        assert restart.format_tree(T) == f.read()
    assert (T.nri, T.nrj) == (1, 1) and restart.used_blocks(T) == [(0, 0, 0, 0)]


def test_to_mesh_keeps_the_two_inner_ghost_layers():
    B = restart.read_soln(FLOW)
    m, U = restart.to_mesh([B])
    assert m.nblk == 1 and int(m.ni[0]) == 40 and int(m.nj[0]) == 40
    x, y = m.block_nodes(0)
    assert x.shape == (45, 45)
    np.testing.assert_array_equal(x, B.xn[3:48, 3:48])
    np.testing.assert_array_equal(U.reshape(40, 40, 4), B.U[5:45, 5:45])
    assert m.block_bc(0)["N"].size == 44


def test_cartesian_restart_reads_back(tmp_path):
    m = M.cartesian(3, 2, 8, 8)
    W = M.density_perturbation(m, M.riemann_4quadrant(m), seed=4)
    m.set_wo_from_initial(W)
    U = np.stack([W[:, 0], W[:, 0] * W[:, 1], W[:, 0] * W[:, 2],
                  W[:, 3] / 0.4 + 0.5 * W[:, 0] * (W[:, 1] ** 2 + W[:, 2] ** 2)], axis=1)
    prefix = str(tmp_path / "c")
    restart.write_restart(prefix, m, U, nsteps=12, time=0.25)
    blocks, T, m2, U2 = restart.read_restart(prefix)
    assert (T.nri, T.nrj) == (3, 2) and blocks[0].nsteps == 12
    np.testing.assert_array_equal(m2.nbr, m.nbr)
    np.testing.assert_array_equal(m2.bc, m.bc)
    np.testing.assert_allclose(m2.xnode, m.xnode, rtol=0, atol=1e-15)
    np.testing.assert_allclose(U2, U, rtol=1e-15)
    # written files read back byte for byte too
    B = restart.read_soln(prefix + "_blk000004.soln")
    with open(prefix + "_blk000004.soln") as f:
        assert restart.format_soln(B) == f.read()


def test_quadtree_restart_reads_back_topology(tmp_path):
    m = M.quadtree_box(nroot=4, n=8, max_level=2, refine=M.circle_refinement(0.1), jitter=0.1, seed=2)
    W = M.cylindrical_explosion(m, radius=0.1)
    U = np.stack([W[:, 0], 0 * W[:, 0], 0 * W[:, 0], W[:, 3] / 0.4], axis=1)
    prefix = str(tmp_path / "q")
    restart.write_restart(prefix, m, U)
    blocks, T, m2, U2 = restart.read_restart(prefix)
    assert len(blocks) == m.nblk
    np.testing.assert_array_equal(m2.level, m.level)
    np.testing.assert_array_equal(m2.nbr, m.nbr)
    np.testing.assert_array_equal(m2.nbr2, m.nbr2)
    np.testing.assert_allclose(m2.xnode, m.xnode, rtol=0, atol=1e-15)
    np.testing.assert_allclose(U2, U, rtol=1e-15)
    with open(prefix + "_quadtree.tree") as f:
        assert restart.format_tree(restart.read_tree(prefix + "_quadtree.tree")) == f.read()


GRIDFILE = os.path.join(G, "CartesianMesh_3x3.grid.gz")


def test_grid_file_round_trip_is_byte_exact():
    """The reference's own grid data file (HighOrder2D unit tests: 3 x 3
    blocks, 5 ghost layers, third-order cell moments)."""
    Gf = restart.read_grid_file(GRIDFILE)
    with gzip.open(GRIDFILE, "rt") as f:
        assert restart.format_grid_file(Gf) == f.read()
    assert (Gf.grid_type, Gf.nbi, Gf.nbj) == ("Square", 3, 3)


def test_grid_file_mesh_is_consistent():
    """The converted mesh: its ghost nodes are the neighbouring blocks'
    interior nodes to round-off (the reference's grid generator computed
    them separately: -2.8e-17 against 0.0 on one line), and the host
    checks of cffc_create accept it."""
    import ctypes
    Gf = restart.read_grid_file(GRIDFILE)
    m = restart.grid_file_to_mesh(Gf)
    assert m.nblk == 9 and set(m.ni.tolist()) == {10} and set(m.nj.tolist()) == {20}
    for b in range(m.nblk):
        x, y = m.block_nodes(b)
        e = m.nbr[b, abi.NBR_E]
        if e >= 0:
            xe, ye = m.block_nodes(e)
            n = int(m.ni[b])
            # b's east ghost node columns n+3, n+4 = e's interior columns 3, 4
            np.testing.assert_allclose(x[:, n + 3:n + 5], xe[:, 3:5], rtol=0, atol=1e-15)
            np.testing.assert_allclose(y[:, n + 3:n + 5], ye[:, 3:5], rtol=0, atol=1e-15)
        nn = m.nbr[b, abi.NBR_N]
        if nn >= 0:
            xn_, yn_ = m.block_nodes(nn)
            n = int(m.nj[b])
            np.testing.assert_allclose(y[n + 3:n + 5, :], yn_[3:5, :], rtol=0, atol=1e-15)
    from cffc_amd import euler2d as E
    L = E.load()
    L.cffc_check_mesh.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    keep = []
    md = abi.mesh_desc(m, keep)
    counts = (ctypes.c_int64 * 4)()
    assert L.cffc_check_mesh(ctypes.byref(md), counts) == 0, L.cffc_last_error()


def test_axisymmetric_flag_round_trips_and_is_enforced(tmp_path):
    """The Axisymmetric flag each .soln block carries after its grid
    (Euler2DQuad.h:1314 writes it, :1347 reads it) survives a write / read,
    and resume() refuses a scheme that disagrees with it (checked before any
    device work)."""
    m = M.cartesian(2, 1, 4, 4, y0=0.5, y1=1.5)
    W = M.uniform_state(m, M.freestream(0.4))
    U = np.stack([W[:, 0], W[:, 0] * W[:, 1], W[:, 0] * W[:, 2],
                  W[:, 3] / 0.4 + 0.5 * W[:, 0] * (W[:, 1] ** 2 + W[:, 2] ** 2)], axis=1)
    for axi in (0, 1):
        prefix = str(tmp_path / ("axi%d" % axi))
        restart.write_restart(prefix, m, U, axisymmetric=axi)
        blocks, _, m2, U2 = restart.read_restart(prefix)
        assert [B.axisymmetric for B in blocks] == [axi, axi]
        with open(prefix + "_blk000000.soln") as f:
            assert restart.format_soln(restart.read_soln(prefix + "_blk000000.soln")) == f.read()
        with pytest.raises(ValueError):
            restart.resume(prefix, abi.Scheme(axisymmetric=1 - axi))
