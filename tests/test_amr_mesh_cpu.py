"""Host-side checks of quadtree meshes through the C ABI (no GPU needed).

cffc_check_mesh runs cffc_create's metadata, topology and resolution-change
checks without a device.  Its operation counts are compared with counts
derived here from the mesh tables: a coarse block restricts two ghost layers
over every face with finer neighbours and 2 x 2 corner cells from a finer
corner neighbour; a fine block prolongs n/2 + 1 coarse cells along a face
with a coarser neighbour (its span plus the corner over the middle of the
coarse face) and one coarse cell into a coarser corner not so covered; the
flux correction touches every coarse cell along a face with finer neighbours
once (corner cells of two such faces once).
"""
import ctypes

import numpy as np
import pytest

from cffc_amd import abi
from cffc_amd import mesh as M


def _lib():
    from cffc_amd import euler2d as E
    L = E.load()
    L.cffc_check_mesh.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    L.cffc_check_mesh.restype = ctypes.c_int
    return L


def check(mesh):
    L = _lib()
    keep = []
    md = abi.mesh_desc(mesh, keep)
    counts = (ctypes.c_int64 * 4)()
    rc = L.cffc_check_mesh(ctypes.byref(md), counts)
    return rc, list(counts), L.cffc_last_error().decode()


def expected_counts(m):
    lv, nbr = m.level, m.nbr
    n = int(m.ni[0])
    restrict = prolong = 0
    fix_cells = set()
    for b in range(m.nblk):
        covered = set()
        for f in range(4):
            o = nbr[b, f]
            if o < 0:
                continue
            if lv[o] == lv[b] + 1:
                restrict += 2 * n
                for q in range(n):
                    fix_cells.add((b, _edge_cell(f, q, n)))
            elif lv[o] == lv[b] - 1:
                prolong += n // 2 + 1
                sig = 0 if nbr[o, f ^ 1] == b else 1
                covered.add({0: (4, 5), 1: (6, 7), 2: (4, 6), 3: (5, 7)}[f][sig])
        for s in range(4, 8):
            o = nbr[b, s]
            if o < 0:
                continue
            if lv[o] == lv[b] + 1:
                restrict += 4
            elif lv[o] == lv[b] - 1 and s not in covered:
                prolong += 1
    return restrict, prolong, len(fix_cells)


def _edge_cell(f, q, n):
    return {0: (2 + q, n + 1), 1: (2 + q, 2), 2: (n + 1, 2 + q), 3: (2, 2 + q)}[f]


@pytest.mark.parametrize("kw", [dict(nroot=4, n=8, radius=0.1), dict(nroot=5, n=8, radius=0.15),
                                dict(nroot=6, n=16, radius=0.2)])
def test_operation_counts(kw):
    m = M.quadtree_box(nroot=kw["nroot"], n=kw["n"], max_level=2, refine=M.circle_refinement(kw["radius"]),
                       jitter=0.1, seed=1)
    rc, counts, err = check(m)
    assert rc == 0, err
    r, p, f = expected_counts(m)
    assert counts == [r, p, f, 1]


def test_single_level_mesh_has_no_operations():
    m = M.cartesian(3, 2, 8, 8)
    rc, counts, err = check(m)
    assert rc == 0, err
    assert counts == [0, 0, 0, 0]


def _small():
    return M.quadtree_box(nroot=4, n=8, max_level=2, refine=M.circle_refinement(0.1), jitter=0.0)


def _find_finer_face(m):
    for b in range(m.nblk):
        for f in range(4):
            o = m.nbr[b, f]
            if o >= 0 and m.level[o] == m.level[b] + 1:
                return b, f
    raise AssertionError("no resolution change")


def test_rejects_two_level_jump():
    m = _small()
    b, f = _find_finer_face(m)
    m.level = m.level.copy()
    m.level[m.nbr[b, f]] += 1
    rc, _, err = check(m)
    assert rc == abi.E_TOPOLOGY and "levels" in err


def test_rejects_missing_second_fine_neighbour():
    m = _small()
    b, f = _find_finer_face(m)
    m.nbr2 = m.nbr2.copy()
    m.nbr2[b, f] = -1
    rc, _, err = check(m)
    assert rc == abi.E_TOPOLOGY and "finer neighbours" in err


def test_rejects_odd_block_sizes_across_a_change():
    m = M.quadtree_box(nroot=4, n=8, max_level=1, refine=M.circle_refinement(0.1), jitter=0.0)
    # the same tables with 7 x 7-cell blocks (arrays resized): a resolution
    # change needs even cell counts
    n = 7
    m.ni = np.full(m.nblk, n, np.int32)
    m.nj = np.full(m.nblk, n, np.int32)
    nn = (n + 5) * (n + 5)
    g = np.arange(-2, n + 3, dtype=np.float64)
    m.xnode = np.tile(np.tile(g, n + 5), m.nblk)
    m.ynode = np.tile(np.repeat(g, n + 5), m.nblk)
    assert m.xnode.size == m.nblk * nn
    m.bc = np.zeros(m.nblk * 4 * (n + 4), np.int32)
    m.wo = np.zeros((m.bc.size, 4))
    rc, _, err = check(m)
    assert rc == abi.E_TOPOLOGY and "even" in err


def test_rejects_second_neighbour_on_same_level_face():
    m = _small()
    for b in range(m.nblk):
        for f in range(4):
            o = m.nbr[b, f]
            if o >= 0 and m.level[o] == m.level[b]:
                m.nbr2 = m.nbr2.copy()
                m.nbr2[b, f] = o
                rc, _, err = check(m)
                assert rc == abi.E_TOPOLOGY and "second neighbour" in err
                return
    raise AssertionError("no same-level face")
