"""Blocks placed by lattice position (cffc_mesh_desc.lattice), host only.

The library builds a lattice mesh's messages from the blocks' positions:
every ghost cell is copied from the same-level block holding its place,
restricted from the finer one or prolonged from the coarser one, and coarse
faces next to finer cells are corrected.  On a quadtree of equal blocks the
reference's message rules (nbr / nbr2 tables) apply too, so the two
builders must produce the same operations:
  * the same restrictions, prolongations (with the same quarter centroids)
    and flux corrections, record for record;
  * the lattice path's same-level copies are exactly the ghost cells the
    exchange kernel copies from same-level neighbours (faces over the
    interior span, 2 x 2 corners), derived here from the neighbour table.
Merged meshes (merge_quadtree): the merged blocks hold the quadtree's cells
and nodes (checked in merge_quadtree itself), and their operations are the
quadtree's, re-indexed to the merged layout, apart from the copies between
leaves that merging turns into interior cells."""
import ctypes

import numpy as np
import pytest

from cffc_amd import abi
from cffc_amd import mesh as M

RESTRICT = np.dtype([("dst", "<i8"), ("src", "<i8"), ("pitch", "<i4"), ("kind", "<i4")])
PROLONG = np.dtype([("src", "<i8"), ("dst", "<i8", 4), ("xq", "<f8", 4), ("yq", "<f8", 4), ("blk", "<i4"),
                    ("pad", "<i4")])
FIX = np.dtype([("cell", "<i8"), ("n", "<i4"), ("blk", "<i4"), ("own", "<i4", 2), ("fine", "<i4", 2)])


def _lib():
    from cffc_amd import euler2d as E
    L = E.load()
    L.cffc_mesh_ops.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_int64)]
    L.cffc_mesh_ops.restype = ctypes.c_int
    return L


def ops(mesh, kind):
    L = _lib()
    keep = []
    md = abi.mesh_desc(mesh, keep)
    n = ctypes.c_int64(0)
    rc = L.cffc_mesh_ops(ctypes.byref(md), kind, None, 0, ctypes.byref(n))
    assert rc == 0, L.cffc_last_error().decode()
    dt = (RESTRICT, RESTRICT, PROLONG, FIX)[kind]
    out = np.zeros(n.value, dt)
    rc = L.cffc_mesh_ops(ctypes.byref(md), kind, out.ctypes.data, n.value, ctypes.byref(n))
    assert rc == 0
    return out


def layout(mesh):
    """(coff, pitch) of every block: the library's padded layout."""
    coff, out = 0, []
    for b in range(mesh.nblk):
        nci, ncj = int(mesh.ni[b]) + 4, int(mesh.nj[b]) + 4
        pitch = (nci + 15) & ~15
        c0 = ((coff + 15) & ~15) + 14
        out.append((c0, pitch))
        coff = c0 + pitch * ncj
    return out


def same_level_copies(mesh):
    """(dst, src) of every ghost cell the exchange copies from a same-level
    neighbour (faces over the interior span, corners 2 x 2)."""
    lay = layout(mesh)
    pairs = set()
    for b in range(mesh.nblk):
        ni, nj = int(mesh.ni[b]), int(mesh.nj[b])
        cb, pb = lay[b]
        for s in range(8):
            o = int(mesh.nbr[b, s])
            if o < 0 or mesh.level[o] != mesh.level[b]:
                continue
            co, po = lay[o]
            onj, oni = int(mesh.nj[o]), int(mesh.ni[o])
            di = {2: -ni, 3: oni, 4: -ni, 5: oni, 6: -ni, 7: oni}.get(s, 0)
            dj = {0: -nj, 1: onj, 4: -nj, 5: -nj, 6: onj, 7: onj}.get(s, 0)
            irange = {2: (ni + 2, ni + 3), 3: (0, 1), 4: (ni + 2, ni + 3), 5: (0, 1), 6: (ni + 2, ni + 3),
                      7: (0, 1)}.get(s, (2, ni + 1))
            jrange = {0: (nj + 2, nj + 3), 1: (0, 1), 4: (nj + 2, nj + 3), 5: (nj + 2, nj + 3), 6: (0, 1),
                      7: (0, 1)}.get(s, (2, nj + 1))
            for j in range(jrange[0], jrange[1] + 1):
                for i in range(irange[0], irange[1] + 1):
                    pairs.add((cb + j * pb + i, co + (j + dj) * po + (i + di)))
    return pairs


def _rows(a, fields):
    return sorted(tuple(np.asarray(r[f]).ravel().tolist() if np.ndim(r[f]) else [r[f].item()] for f in fields)
                  for r in a)


@pytest.fixture(scope="module")
def quadtree():
    m = M.quadtree_box(nroot=8, n=8, max_level=2, refine=M.circle_refinement(0.3, band=0.02), jitter=0.1, seed=5)
    lat = M.Mesh(**{k: getattr(m, k) for k in ("nblk", "ni", "nj", "xnode", "ynode", "nbr", "bc", "wo", "level",
                                                 "nbr2", "lattice")}, info=dict(m.info, use_lattice=True))
    return m, lat


def test_lattice_builder_equals_the_reference_rules_on_equal_blocks(quadtree):
    m, lat = quadtree
    assert len(ops(m, 1)) == 0   # the neighbour-table path exchanges same-level ghosts itself
    r0, r1 = ops(m, 0), ops(lat, 0)
    assert len(r0) > 0 and _rows(r0, ("dst", "src", "pitch")) == _rows(r1, ("dst", "src", "pitch"))
    p0, p1 = ops(m, 2), ops(lat, 2)
    assert len(p0) > 0 and len(p0) == len(p1)
    k0 = sorted(p0.tolist(), key=lambda r: (r[0], tuple(r[1])))
    k1 = sorted(p1.tolist(), key=lambda r: (r[0], tuple(r[1])))
    for a, b in zip(k0, k1):
        assert a[0] == b[0] and list(a[1]) == list(b[1]) and a[4] == b[4]
        # quarter centroids bit for bit
        assert np.array_equal(np.array(a[2]).view(np.uint64), np.array(b[2]).view(np.uint64))
        assert np.array_equal(np.array(a[3]).view(np.uint64), np.array(b[3]).view(np.uint64))
    f0, f1 = ops(m, 3), ops(lat, 3)
    norm = lambda a: sorted((r["cell"], r["n"], r["blk"], tuple(r["own"][:r["n"]]), tuple(r["fine"][:r["n"]]))
                            for r in a)
    assert len(f0) > 0 and norm(f0) == norm(f1)
    c1 = ops(lat, 1)
    got = set(zip(c1["dst"].tolist(), c1["src"].tolist()))
    assert len(got) == len(c1)
    assert got == same_level_copies(m)


def test_merged_mesh_has_all_three_block_sizes_and_levels():
    m = M.quadtree_box(nroot=16, n=8, max_level=2, refine=M.circle_refinement(0.3, band=0.02), jitter=0.1, seed=5)
    mm, cmap = M.merge_quadtree(m)
    assert mm.ncell == m.ncell and np.array_equal(np.sort(cmap), np.arange(m.ncell))
    assert set(np.unique(mm.ni).tolist()) == {8, 16, 32}
    for lv in range(3):
        assert len(np.unique(mm.ni[mm.level == lv])) >= 2
    # no block and no BC for a ghost cell, split coarse cells, etc. are refused
    assert len(ops(mm, 0)) > 0 and len(ops(mm, 1)) > 0 and len(ops(mm, 2)) > 0 and len(ops(mm, 3)) > 0


def test_merged_operations_are_the_quadtree_operations():
    """Restrictions, prolongations and flux corrections of the merged mesh
    are those of the quadtree, cell for cell, once padded cell indices are
    mapped to (leaf, i, j) places; the quadtree's same-level copies between
    leaves of one merged block disappear, the others remain."""
    m = M.quadtree_box(nroot=8, n=8, max_level=2, refine=M.circle_refinement(0.3, band=0.02), jitter=0.1, seed=5)
    mm, _ = M.merge_quadtree(m)
    n = m.info["n"]
    keys = m.info["keys"]
    lay_q, lay_m = layout(m), layout(mm)
    # place (level, I, J) of a padded cell, from either mesh
    def place_q():
        d = {}
        for b, (lv, x, y) in enumerate(keys):
            c0, p = lay_q[b]
            for j in range(n + 4):
                for i in range(n + 4):
                    d[c0 + j * p + i] = (lv, x * n + i - 2, y * n + j - 2, b, i in (0, 1, n + 2, n + 3) or
                                         j in (0, 1, n + 2, n + 3))
        return d
    def place_m():
        d = {}
        for b in range(mm.nblk):
            c0, p = lay_m[b]
            N = int(mm.ni[b])
            I0, J0 = mm.lattice[b]
            for j in range(N + 4):
                for i in range(N + 4):
                    d[c0 + j * p + i] = (int(mm.level[b]), I0 + i - 2, J0 + j - 2, b)
        return d
    pq, pm = place_q(), place_m()
    # restrictions: (destination place, source place) -- the coarse ghost
    # places a merged block still has
    rq = {(pq[r["dst"]][:3], pq[r["src"]][:3]) for r in ops(m, 0)}
    rm = {(pm[r["dst"]][:3], pm[r["src"]][:3]) for r in ops(mm, 0)}
    assert rm <= rq and len(rm) > 0
    pq2 = {(pq[r["src"]][:3], tuple(pq[d][:3] for d in r["dst"])) for r in ops(m, 2)}
    pm2 = {(pm[r["src"]][:3], tuple(pm[d][:3] for d in r["dst"])) for r in ops(mm, 2)}
    assert pm2 <= pq2 and len(pm2) > 0
    # every restricted or prolonged place of the quadtree is still filled in
    # the merged mesh (ghost places are shared by neighbouring leaves)
    assert {a for a, _ in rq} == {a for a, _ in rm}
    assert {d for _, ds in pq2 for d in ds} == {d for _, ds in pm2 for d in ds}
