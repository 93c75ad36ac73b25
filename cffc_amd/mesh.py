"""Multi-block quadrilateral meshes, topology and initial data.

A Mesh holds what the reference keeps in Grid2D_Quad_Block_HO (nodes and the
BCtypeN/S/E/W arrays, HO_Grid2DQuad.h:264-330), in the solution blocks'
boundary reference states WoN/S/E/W (Euler2DQuad.h:140-205) and in the
same-level neighbour fields of AdaptiveBlock2D_List (AdaptiveBlock.h:820).
Arrays are laid out exactly as cffc_mesh_desc (include/cffc_euler2d.h)
expects: blocks concatenated, i fastest, two ghost layers included.

Generators here are this project's own: a uniform or node-jittered Cartesian
block grid (the reference's Grid_Rectangular_Box, HO_Grid2DQuadMultiBlock.cc:
858-960, and the seeded analogue of Disturb_Interior_Nodes_Without_Update,
HO_Grid2DQuad.cc:12144-12170) and the circular-arc bump channel
(HO_Grid2DQuadMultiBlock.cc:5385-5565).
"""
from dataclasses import dataclass, field

import numpy as np

from . import abi

NG = 2


@dataclass
class Mesh:
    nblk: int
    ni: np.ndarray            # [nblk] int32
    nj: np.ndarray            # [nblk] int32
    xnode: np.ndarray         # concatenated (nj+5)*(ni+5) per block
    ynode: np.ndarray
    nbr: np.ndarray           # [nblk, 8] int32
    bc: np.ndarray            # concatenated W[ncj], E[ncj], S[nci], N[nci]
    wo: np.ndarray            # [len(bc), 4]
    info: dict = field(default_factory=dict)
    level: np.ndarray = None  # [nblk] int32 quadtree level, or None (one level)
    nbr2: np.ndarray = None   # [nblk, 4] int32 second finer face neighbour, or None
    lattice: np.ndarray = None  # [nblk, 2] int32 first interior cell on its level's lattice, or None

    # --- offsets -----------------------------------------------------------
    def node_offsets(self):
        sz = (self.ni + 2 * NG + 1) * (self.nj + 2 * NG + 1)
        return np.concatenate([[0], np.cumsum(sz)]).astype(np.int64)

    def bc_offsets(self):
        sz = 2 * (self.nj + 2 * NG) + 2 * (self.ni + 2 * NG)
        return np.concatenate([[0], np.cumsum(sz)]).astype(np.int64)

    def cell_offsets(self):
        sz = self.ni * self.nj
        return np.concatenate([[0], np.cumsum(sz)]).astype(np.int64)

    @property
    def ncell(self):
        return int(np.sum(self.ni.astype(np.int64) * self.nj))

    def block_nodes(self, b):
        o = self.node_offsets()
        nci, ncj = int(self.ni[b]) + 2 * NG, int(self.nj[b]) + 2 * NG
        x = self.xnode[o[b]:o[b + 1]].reshape(ncj + 1, nci + 1)
        y = self.ynode[o[b]:o[b + 1]].reshape(ncj + 1, nci + 1)
        return x, y

    def block_bc(self, b):
        o = self.bc_offsets()[b]
        nci, ncj = int(self.ni[b]) + 2 * NG, int(self.nj[b]) + 2 * NG
        bc = self.bc
        return {"W": bc[o:o + ncj], "E": bc[o + ncj:o + 2 * ncj],
                "S": bc[o + 2 * ncj:o + 2 * ncj + nci], "N": bc[o + 2 * ncj + nci:o + 2 * ncj + 2 * nci]}

    def cell_centres(self):
        """Approximate centres (average of the 4 nodes) of interior cells, in
        the state-array order; used to place initial data."""
        out = np.empty((self.ncell, 2))
        co = self.cell_offsets()
        for b in range(self.nblk):
            x, y = self.block_nodes(b)
            ni, nj = int(self.ni[b]), int(self.nj[b])
            xs = 0.25 * (x[:-1, :-1] + x[1:, :-1] + x[:-1, 1:] + x[1:, 1:])
            ys = 0.25 * (y[:-1, :-1] + y[1:, :-1] + y[:-1, 1:] + y[1:, 1:])
            out[co[b]:co[b + 1], 0] = xs[NG:NG + nj, NG:NG + ni].ravel()
            out[co[b]:co[b + 1], 1] = ys[NG:NG + nj, NG:NG + ni].ravel()
        return out

    def set_wo_from_initial(self, W):
        """Boundary reference states as the reference's ICs set them
        (Euler2DQuadSingleBlock.cc:2310-2340): WoW[j] = W[ICl][j] over the
        interior span, clamped to the corner cell for ghost rows; likewise
        for E, S, N."""
        W = np.asarray(W).reshape(self.ncell, 4)
        co, bo = self.cell_offsets(), self.bc_offsets()
        for b in range(self.nblk):
            ni, nj = int(self.ni[b]), int(self.nj[b])
            Wb = W[co[b]:co[b + 1]].reshape(nj, ni, 4)
            jj = np.clip(np.arange(-NG, nj + NG), 0, nj - 1)
            ii = np.clip(np.arange(-NG, ni + NG), 0, ni - 1)
            o = bo[b]
            ncj, nci = nj + 2 * NG, ni + 2 * NG
            self.wo[o:o + ncj] = Wb[jj, 0]
            self.wo[o + ncj:o + 2 * ncj] = Wb[jj, ni - 1]
            self.wo[o + 2 * ncj:o + 2 * ncj + nci] = Wb[0, ii]
            self.wo[o + 2 * ncj + nci:o + 2 * ncj + 2 * nci] = Wb[nj - 1, ii]


def _structured_blocks(X, Y, nbi, nbj, ni, nj, bcs):
    """Slice a global node array with 2 ghost node layers on every side
    (shape (NJ+1+2NG, NI+1+2NG)) into nbi x nbj blocks of ni x nj cells."""
    nblk = nbi * nbj
    xs, ys, nbr, bc = [], [], np.full((nblk, 8), -1, np.int32), []
    for jb in range(nbj):
        for ib in range(nbi):
            b = jb * nbi + ib
            i0, j0 = ib * ni, jb * nj
            xs.append(X[j0:j0 + nj + 2 * NG + 1, i0:i0 + ni + 2 * NG + 1].ravel())
            ys.append(Y[j0:j0 + nj + 2 * NG + 1, i0:i0 + ni + 2 * NG + 1].ravel())

            def nb(di, dj):
                a, c = ib + di, jb + dj
                return c * nbi + a if (0 <= a < nbi and 0 <= c < nbj) else -1
            nbr[b] = [nb(0, 1), nb(0, -1), nb(1, 0), nb(-1, 0), nb(1, 1), nb(-1, 1), nb(1, -1), nb(-1, -1)]
            ncj, nci = nj + 2 * NG, ni + 2 * NG
            bc.append(np.full(ncj, bcs["W"] if ib == 0 else abi.BC_NONE, np.int32))
            bc.append(np.full(ncj, bcs["E"] if ib == nbi - 1 else abi.BC_NONE, np.int32))
            bc.append(np.full(nci, bcs["S"] if jb == 0 else abi.BC_NONE, np.int32))
            bc.append(np.full(nci, bcs["N"] if jb == nbj - 1 else abi.BC_NONE, np.int32))
    bc = np.concatenate(bc)
    return Mesh(nblk=nblk, ni=np.full(nblk, ni, np.int32), nj=np.full(nblk, nj, np.int32),
                xnode=np.concatenate(xs), ynode=np.concatenate(ys), nbr=nbr, bc=bc,
                wo=np.zeros((bc.size, 4)),
                info={"nbi": nbi, "nbj": nbj, "ni": ni, "nj": nj})


def cartesian(nbi, nbj, ni, nj, x0=0.0, x1=1.0, y0=0.0, y1=1.0, bcs=None, jitter=0.0, seed=0):
    """nbi x nbj blocks of ni x nj cells on [x0,x1]x[y0,y1].  With jitter > 0
    every node strictly inside the domain moves by r*h at a uniform random
    angle, r ~ U[0, jitter], h the local minimum spacing; seeded."""
    if bcs is None:
        bcs = dict(W=abi.BC_CONSTANT_EXTRAPOLATION, E=abi.BC_CONSTANT_EXTRAPOLATION,
                   S=abi.BC_CONSTANT_EXTRAPOLATION, N=abi.BC_CONSTANT_EXTRAPOLATION)
    NI, NJ = nbi * ni, nbj * nj
    ig = np.arange(-NG, NI + NG + 1, dtype=np.float64)
    jg = np.arange(-NG, NJ + NG + 1, dtype=np.float64)
    xl = x0 + (x1 - x0) * (ig / NI)
    yl = y0 + (y1 - y0) * (jg / NJ)
    X, Y = np.meshgrid(xl, yl)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        h = min((x1 - x0) / NI, (y1 - y0) / NJ)
        inner = (slice(NG + 1, NG + NJ), slice(NG + 1, NG + NI))
        shape = X[inner].shape
        r = rng.uniform(0.0, jitter, shape) * h
        th = rng.uniform(0.0, 2 * np.pi, shape)
        X[inner] += r * np.cos(th)
        Y[inner] += r * np.sin(th)
    m = _structured_blocks(X, Y, nbi, nbj, ni, nj, bcs)
    m.info.update(kind="cartesian", jitter=jitter, seed=seed, box=(x0, x1, y0, y1))
    return m


def bump_channel(h, nbi=12, nbj=4, ni=16, nj=16, length=3.0, height=1.0, x_bump=(1.0, 2.0),
                 inflow=abi.BC_FIXED, outflow=abi.BC_CONSTANT_EXTRAPOLATION):
    """Channel [0,length]x[0,height] with a circular-arc bump of height h on
    the lower wall between x_bump[0] and x_bump[1] (chord 1 by default), the
    arc of radius R = (c^2/4 + h^2)/(2h) as in the reference's bump grid
    (HO_Grid2DQuadMultiBlock.cc:5434-5442).  Nodes are uniform in x and
    stretched linearly between the wall and the top in y.  Boundary
    conditions follow the reference's bump channel (:5556-5565): reflection
    on the lower and upper walls, fixed inflow (west), constant
    extrapolation outflow (east); inflow and outflow select others."""
    NI, NJ = nbi * ni, nbj * nj
    ig = np.arange(-NG, NI + NG + 1, dtype=np.float64)
    jg = np.arange(-NG, NJ + NG + 1, dtype=np.float64)
    xl = length * (ig / NI)
    chord = x_bump[1] - x_bump[0]
    xm = 0.5 * (x_bump[0] + x_bump[1])
    R = (0.25 * chord * chord + h * h) / (2.0 * h)
    inside = (xl > x_bump[0]) & (xl < x_bump[1])
    yb = np.zeros_like(xl)
    yb[inside] = np.sqrt(R * R - (xl[inside] - xm) ** 2) - (R - h)
    X = np.tile(xl, (jg.size, 1))
    Y = yb[None, :] + (height - yb[None, :]) * (jg[:, None] / NJ)
    bcs = dict(W=inflow, E=outflow, S=abi.BC_REFLECTION, N=abi.BC_REFLECTION)
    m = _structured_blocks(X, Y, nbi, nbj, ni, nj, bcs)
    m.info.update(kind="bump", h=h, R=R, length=length, height=height, NI=NI, NJ=NJ)
    return m


# ---- quadtree (AMR) block meshes --------------------------------------------------
# Face slots N, S, E, W as (dx, dy); the two finer blocks of a face in the order
# of the reference's infoX[0], infoX[1] (west then east, south then north).
_FACE_DIR = ((0, 1), (0, -1), (1, 0), (-1, 0))
_CORNER_DIR = ((1, 1), (-1, 1), (1, -1), (-1, -1))   # NE, NW, SE, SW


def _children_on_face(level, x, y, f):
    """The two level+1 squares across face f of square (level, x, y)."""
    if f == 0:
        return (level + 1, 2 * x, 2 * y + 2), (level + 1, 2 * x + 1, 2 * y + 2)
    if f == 1:
        return (level + 1, 2 * x, 2 * y - 1), (level + 1, 2 * x + 1, 2 * y - 1)
    if f == 2:
        return (level + 1, 2 * x + 2, 2 * y), (level + 1, 2 * x + 2, 2 * y + 1)
    return (level + 1, 2 * x - 1, 2 * y), (level + 1, 2 * x - 1, 2 * y + 1)


def _child_at_corner(level, x, y, c):
    dx, dy = _CORNER_DIR[c]
    return (level + 1, 2 * x + (2 if dx > 0 else -1), 2 * y + (2 if dy > 0 else -1))


def _morton(x, y):
    z = 0
    for bit in range(20):
        z |= ((x >> bit) & 1) << (2 * bit) | ((y >> bit) & 1) << (2 * bit + 1)
    return z


def _quad_centroid(x, y):
    """quadCentroid (Math.h:311-350) of vertices SW, SE, NE, NW."""
    a = xs = ys = 0.0
    for p, q in ((3, 0), (0, 1), (1, 2), (2, 3)):
        ai = x[p] * y[q] - x[q] * y[p]
        a += ai
        xs += (x[q] + x[p]) * ai
        ys += (y[q] + y[p]) * ai
    return xs / (3 * a), ys / (3 * a)


def _tables_from_leaves(leaves, nri, nrj):
    """Neighbour tables (nbr [n, 8], nbr2 [n, 4]) of quadtree leaves
    (level, ix, iy) over nri x nrj roots, in the order given: a face's
    neighbour at the same or a coarser level, or its two finer neighbours
    (west/south one in nbr, the other in nbr2); a corner's neighbour is the
    block holding the diagonal cell."""
    keys = set(leaves)
    bid = {k: b for b, k in enumerate(leaves)}

    def inside(lv, x, y):
        return 0 <= x < (nri << lv) and 0 <= y < (nrj << lv)

    def find(lv, x, y):
        for l in range(lv, -1, -1):
            k = (l, x >> (lv - l), y >> (lv - l))
            if k in keys:
                return k
        return None

    n = len(leaves)
    nbr = np.full((n, 8), -1, np.int32)
    nbr2 = np.full((n, 4), -1, np.int32)
    for b, (lv, x, y) in enumerate(leaves):
        for f, (dx, dy) in enumerate(_FACE_DIR):
            if not inside(lv, x + dx, y + dy):
                continue
            o = find(lv, x + dx, y + dy)
            if o is not None:
                nbr[b, f] = bid[o]
            else:
                c0, c1 = _children_on_face(lv, x, y, f)
                nbr[b, f], nbr2[b, f] = bid[c0], bid[c1]
        for c, (dx, dy) in enumerate(_CORNER_DIR):
            if not inside(lv, x + dx, y + dy):
                continue
            o = find(lv, x + dx, y + dy)
            nbr[b, 4 + c] = bid[o] if o is not None else bid[_child_at_corner(lv, x, y, c)]
    return nbr, nbr2


def quadtree_box(nroot=4, n=8, max_level=2, refine=None, x0=0.0, x1=1.0, y0=0.0, y1=1.0, bcs=None,
                 jitter=0.1, seed=0):
    """Square domain [x0,x1]x[y0,y1] split into nroot x nroot root blocks,
    refined as a quadtree (the reference's AMR, QuadTree.h) up to max_level:
    a leaf is split into four when refine(level, xlo, xhi, ylo, yhi) is true,
    then leaves are split until face and corner neighbours differ by at most
    one level.  Every block has n x n cells; blocks are listed in Morton
    order.  With jitter > 0, the nodes strictly inside every block move by
    jitter*h at a uniform random angle (h the block's cell size; the
    reference's Disturb_Interior_Nodes_Without_Update, HO_Grid2DQuad.cc:
    12144-12170, moves them by 0.1 x the node-edge distance, which is h on
    the undisturbed lattice).  Ghost nodes follow the reference's geometry
    messages: a same-level neighbour's nodes; at a finer neighbour, the fine
    nodes at stride 2 (Load_Send_Message_Buffers_ResChange_FineToCoarse,
    AdaptiveBlock2D_MessagePassing.h:1290-1320); at a coarser neighbour, the
    quarters of the coarse cells, cut at the edge midpoints and the cell
    centroid (Load_Send_Message_Buffers_ResChange_CoarseToFine :2518-2570);
    ghost nodes on the domain boundary stay on the lattice."""
    if bcs is None:
        bcs = dict(W=abi.BC_REFLECTION, E=abi.BC_REFLECTION, S=abi.BC_REFLECTION, N=abi.BC_REFLECTION)
    if n % 2 or n < 4:
        raise ValueError("blocks across a resolution change need an even n >= 4")
    leaves = {(0, ix, iy) for ix in range(nroot) for iy in range(nroot)}
    Lx, Ly = x1 - x0, y1 - y0

    def extent(k):
        lv, ix, iy = k
        s = 1.0 / (nroot << lv)
        return x0 + Lx * ix * s, x0 + Lx * (ix + 1) * s, y0 + Ly * iy * s, y0 + Ly * (iy + 1) * s

    def inside(lv, x, y):
        m = nroot << lv
        return 0 <= x < m and 0 <= y < m

    def find(lv, x, y):
        """leaf containing square (lv, x, y) at its level or coarser, else None"""
        for l in range(lv, -1, -1):
            k = (l, x >> (lv - l), y >> (lv - l))
            if k in leaves:
                return k
        return None

    def split(k):
        lv, x, y = k
        leaves.remove(k)
        for cy in range(2):
            for cx in range(2):
                leaves.add((lv + 1, 2 * x + cx, 2 * y + cy))

    if refine is not None:
        for lv in range(max_level):
            for k in sorted(kk for kk in leaves if kk[0] == lv):
                if refine(lv, *extent(k)):
                    split(k)
    changed = True
    while changed:
        changed = False
        for k in sorted(leaves):
            if k not in leaves:
                continue
            lv, x, y = k
            for dx, dy in _FACE_DIR + _CORNER_DIR:
                if not inside(lv, x + dx, y + dy):
                    continue
                o = find(lv, x + dx, y + dy)
                if o is not None and o[0] < lv - 1:
                    split(o)
                    changed = True
    order = sorted(leaves, key=lambda k: _morton(k[1] << (max_level + 1 - k[0]), k[2] << (max_level + 1 - k[0])))
    nblk = len(order)
    level = np.array([k[0] for k in order], np.int32)
    nbr, nbr2 = _tables_from_leaves(order, nroot, nroot)

    # own nodes: lattice (ghost nodes included), interior nodes disturbed
    rng = np.random.default_rng(seed)
    N = n + 2 * NG + 1
    X = np.empty((nblk, N, N))
    Y = np.empty((nblk, N, N))
    # lattice nodes from one integer lattice at the finest level, so that a
    # node shared by several blocks has the same coordinates in all of them
    K = (nroot * n) << max_level
    for b, k in enumerate(order):
        lv, ix, iy = k
        sc = 1 << (max_level - lv)
        hx, hy = Lx / ((nroot * n) << lv), Ly / ((nroot * n) << lv)
        idx = np.arange(-NG, n + NG + 1, dtype=np.int64)
        X[b] = (x0 + Lx * (((ix * n + idx) * sc) / K))[None, :]
        Y[b] = (y0 + Ly * (((iy * n + idx) * sc) / K))[:, None]
        if jitter > 0:
            th = rng.uniform(0.0, 2 * np.pi, (n - 1, n - 1))
            d = jitter * min(hx, hy)
            X[b, NG + 1:NG + n, NG + 1:NG + n] += d * np.cos(th)
            Y[b, NG + 1:NG + n, NG + 1:NG + n] += d * np.sin(th)
    X0, Y0 = X.copy(), Y.copy()   # the blocks' own nodes (sources of the ghost geometry)

    def centroid(b, i, j):
        xs = [X0[b, j, i], X0[b, j, i + 1], X0[b, j + 1, i + 1], X0[b, j + 1, i]]
        ys = [Y0[b, j, i], Y0[b, j, i + 1], Y0[b, j + 1, i + 1], Y0[b, j + 1, i]]
        return _quad_centroid(xs, ys)

    def quarter_nodes(b, i, j, qx, qy):
        """nodes (SW, SE, NW, NE order as (dx, dy) -> xy) of quarter (qx, qy) of cell (i, j)"""
        P = {}
        for a in range(3):
            for c in range(3):
                if a % 2 == 0 and c % 2 == 0:
                    P[a, c] = (X0[b, j + c // 2, i + a // 2], Y0[b, j + c // 2, i + a // 2])
                elif a == 1 and c == 1:
                    P[a, c] = centroid(b, i, j)
                elif a == 1:
                    P[a, c] = (0.5 * (X0[b, j + c // 2, i] + X0[b, j + c // 2, i + 1]),
                               0.5 * (Y0[b, j + c // 2, i] + Y0[b, j + c // 2, i + 1]))
                else:
                    P[a, c] = (0.5 * (X0[b, j, i + a // 2] + X0[b, j + 1, i + a // 2]),
                               0.5 * (Y0[b, j, i + a // 2] + Y0[b, j + 1, i + a // 2]))
        return {(dx, dy): P[qx + dx, qy + dy] for dx in (0, 1) for dy in (0, 1)}

    def put(b, i, j, nodes):
        for (dx, dy), (px, py) in nodes.items():
            ii, jj = i + dx, j + dy
            if NG <= ii <= n + NG and NG <= jj <= n + NG:
                continue      # the block's own nodes are never replaced
            X[b, jj, ii], Y[b, jj, ii] = px, py

    lo, hi = NG, n + NG - 1
    rng_face = {0: ((lo, hi), (hi + 1, hi + 2)), 1: ((lo, hi), (lo - 2, lo - 1)),
                2: ((hi + 1, hi + 2), (lo, hi)), 3: ((lo - 2, lo - 1), (lo, hi))}
    rng_corner = {0: ((hi + 1, hi + 2), (hi + 1, hi + 2)), 1: ((lo - 2, lo - 1), (hi + 1, hi + 2)),
                  2: ((hi + 1, hi + 2), (lo - 2, lo - 1)), 3: ((lo - 2, lo - 1), (lo - 2, lo - 1))}

    def cells(r):
        (ia, ib), (ja, jb) = r
        return [(i, j) for j in range(ja, jb + 1) for i in range(ia, ib + 1)]

    for stage in ("same", "finer", "coarser"):
        for b in range(nblk):
            lv = level[b]
            covered = set()
            regions = [(f, rng_face[f], _FACE_DIR[f]) for f in range(4)]
            regions += [(4 + c, rng_corner[c], _CORNER_DIR[c]) for c in range(4)]
            for f in range(4):
                C = nbr[b, f]
                if C >= 0 and level[C] == lv - 1:
                    sig = 0 if nbr[C, f ^ 1] == b else 1
                    covered.add({0: (4, 5), 1: (6, 7), 2: (4, 6), 3: (5, 7)}[f][sig])
            for slot, r, (dx, dy) in regions:
                o = nbr[b, slot]
                if o < 0:
                    continue
                rel = level[o] - lv
                if slot >= 4 and slot in covered:
                    continue
                if stage == "same" and rel == 0:
                    # the region's nodes, the block's own nodes excepted
                    (ia, ib), (ja, jb) = r
                    ii = np.arange(ia, ib + 2)
                    jj = np.arange(ja, jb + 2)
                    own = ((ii >= NG) & (ii <= n + NG))[None, :] & ((jj >= NG) & (jj <= n + NG))[:, None]
                    src = np.ix_(jj - dy * n, ii - dx * n)
                    dst = np.ix_(jj, ii)
                    X[b][dst] = np.where(own, X[b][dst], X0[o][src])
                    Y[b][dst] = np.where(own, Y[b][dst], Y0[o][src])
                elif stage == "finer" and rel == 1:
                    for i, j in cells(r):
                        if slot < 4:
                            along = (i - lo) if dy != 0 else (j - lo)
                            F = nbr[b, slot] if along < n // 2 else nbr2[b, slot]
                            kk = along % (n // 2)
                            g = (j - hi) if dy > 0 else (lo - j) if dy < 0 else (i - hi) if dx > 0 else (lo - i)
                            if dy != 0:
                                fi = lo + 2 * kk
                                fj = lo + 2 * (g - 1) if dy > 0 else hi - 2 * g + 1
                            else:
                                fj = lo + 2 * kk
                                fi = lo + 2 * (g - 1) if dx > 0 else hi - 2 * g + 1
                        else:
                            F = o
                            gi = (i - hi) if dx > 0 else (lo - i)
                            gj = (j - hi) if dy > 0 else (lo - j)
                            fi = lo + 2 * (gi - 1) if dx > 0 else hi - 2 * gi + 1
                            fj = lo + 2 * (gj - 1) if dy > 0 else hi - 2 * gj + 1
                        put(b, i, j, {(a, c): (X0[F, fj + 2 * c, fi + 2 * a], Y0[F, fj + 2 * c, fi + 2 * a])
                                      for a in (0, 1) for c in (0, 1)})
            if stage != "coarser":
                continue
            for f in range(4):
                C = nbr[b, f]
                if C < 0 or level[C] != lv - 1:
                    continue
                sig = 0 if nbr[C, f ^ 1] == b else 1
                dx, dy = _FACE_DIR[f]
                span = range(lo, hi + 3) if sig == 0 else range(lo - 2, hi + 1)
                for t in span:
                    k = t - lo
                    kf = k // 2
                    for g in (1, 2):
                        if dy != 0:
                            i, j = t, (hi + g if dy > 0 else lo - g)
                            ic, jc = lo + sig * (n // 2) + kf, (lo if dy > 0 else hi)
                            qx, qy = k - 2 * kf, (g - 1 if dy > 0 else 2 - g)
                        else:
                            i, j = (hi + g if dx > 0 else lo - g), t
                            ic, jc = (lo if dx > 0 else hi), lo + sig * (n // 2) + kf
                            qx, qy = (g - 1 if dx > 0 else 2 - g), k - 2 * kf
                        put(b, i, j, quarter_nodes(C, ic, jc, qx, qy))
            for c in range(4):
                D = nbr[b, 4 + c]
                if (4 + c) in covered or D < 0 or level[D] != lv - 1:
                    continue
                dx, dy = _CORNER_DIR[c]
                ic, jc = (lo if dx > 0 else hi), (lo if dy > 0 else hi)
                i0, j0 = (hi + 1 if dx > 0 else lo - 2), (hi + 1 if dy > 0 else lo - 2)
                for qy in (0, 1):
                    for qx in (0, 1):
                        put(b, i0 + qx, j0 + qy, quarter_nodes(D, ic, jc, qx, qy))

    nc = n + 2 * NG
    bc = []
    for b, (lv, x, y) in enumerate(order):
        m = nroot << lv
        bc.append(np.full(nc, bcs["W"] if x == 0 else abi.BC_NONE, np.int32))
        bc.append(np.full(nc, bcs["E"] if x == m - 1 else abi.BC_NONE, np.int32))
        bc.append(np.full(nc, bcs["S"] if y == 0 else abi.BC_NONE, np.int32))
        bc.append(np.full(nc, bcs["N"] if y == m - 1 else abi.BC_NONE, np.int32))
    bc = np.concatenate(bc)
    mesh = Mesh(nblk=nblk, ni=np.full(nblk, n, np.int32), nj=np.full(nblk, n, np.int32),
                xnode=X.reshape(-1), ynode=Y.reshape(-1), nbr=nbr, bc=bc, wo=np.zeros((bc.size, 4)),
                info={"kind": "quadtree", "nroot": nroot, "n": n, "max_level": max_level, "jitter": jitter,
                      "seed": seed, "box": (x0, x1, y0, y1), "keys": order},
                level=level, nbr2=nbr2,
                lattice=np.array([(x * n, y * n) for _, x, y in order], np.int32).reshape(-1, 2))
    return mesh


def merge_quadtree(mesh, depth=2):
    """Blocks of uneven size from a quadtree_box mesh: every aligned group of
    2^k x 2^k leaves of one level (k = depth, then smaller; root blocks
    group the same way over the root grid) becomes one block of (2^k n)^2
    cells covering the same cells with the same nodes.  The result has
    blocks of n^2, (2n)^2, ... (2^depth n)^2 cells at every level and the
    resolution changes of the quadtree between them; the library derives
    its messages from the blocks' lattice positions (cffc_mesh_desc.lattice).
    Returns (merged mesh, cell_map): merged cell k is cell cell_map[k] of
    the quadtree mesh (state-array order)."""
    if mesh.info.get("kind") != "quadtree":
        raise ValueError("merge_quadtree takes a quadtree_box mesh")
    n = mesh.info["n"]
    keys = mesh.info["keys"]
    bid = {k: b for b, k in enumerate(keys)}
    bo = mesh.bc_offsets()
    co = mesh.cell_offsets()
    no = mesh.node_offsets()
    nc0 = n + 2 * NG
    Xs = [(mesh.xnode[no[b]:no[b + 1]].reshape(nc0 + 1, nc0 + 1), mesh.ynode[no[b]:no[b + 1]].reshape(nc0 + 1, nc0 + 1))
          for b in range(mesh.nblk)]
    bcs = [{"W": mesh.bc[bo[b]:bo[b] + nc0], "E": mesh.bc[bo[b] + nc0:bo[b] + 2 * nc0],
            "S": mesh.bc[bo[b] + 2 * nc0:bo[b] + 3 * nc0], "N": mesh.bc[bo[b] + 3 * nc0:bo[b] + 4 * nc0]}
           for b in range(mesh.nblk)]
    used = set()
    groups = []
    for k in range(depth, -1, -1):
        g = 1 << k
        for key in keys:
            lv, x, y = key
            if key in used or x % g or y % g:
                continue
            members = [(lv, x + a, y + c) for c in range(g) for a in range(g)]
            if all(m in bid and m not in used for m in members):
                used.update(members)
                groups.append((lv, x, y, g))
    # blocks in Morton order of their lower-left corner at the finest level
    ml = mesh.info["max_level"]
    groups.sort(key=lambda t: _morton(t[1] << (ml + 1 - t[0]), t[2] << (ml + 1 - t[0])))
    xs, ys, bc, wo, ni, lat, lev, cmap = [], [], [], [], [], [], [], []
    for lv, x, y, g in groups:
        N = g * n
        M = N + 2 * NG + 1
        X = np.empty((M, M))
        Y = np.empty((M, M))
        for cj in range(g):
            for ci in range(g):
                b = bid[(lv, x + ci, y + cj)]
                win = (slice(cj * n, cj * n + n + 2 * NG + 1), slice(ci * n, ci * n + n + 2 * NG + 1))
                X[win] = Xs[b][0]
                Y[win] = Xs[b][1]
        # every leaf's nodes, its ghost nodes included, are the merged
        # block's nodes at the same place: the merged block has the same
        # cells, ghost cells included, as each of its leaves
        for cj in range(g):
            for ci in range(g):
                b = bid[(lv, x + ci, y + cj)]
                win = (slice(cj * n, cj * n + n + 2 * NG + 1), slice(ci * n, ci * n + n + 2 * NG + 1))
                if not (np.array_equal(X[win], Xs[b][0]) and np.array_equal(Y[win], Xs[b][1])):
                    raise ValueError("leaves of a merged block disagree on shared nodes")
        xs.append(X.ravel())
        ys.append(Y.ravel())
        nc = N + 2 * NG
        # boundary condition codes and reference states along each side
        for side in ("W", "E", "S", "N"):
            arr = np.empty(nc, np.int32)
            st = np.empty((nc, 4))
            for t in range(nc):
                cc = min(max((t - NG) // n, 0), g - 1)
                if side in ("W", "E"):
                    b = bid[(lv, x + (0 if side == "W" else g - 1), y + cc)]
                else:
                    b = bid[(lv, x + cc, y + (0 if side == "S" else g - 1))]
                k0 = {"W": 0, "E": n + 2 * NG, "S": 2 * (n + 2 * NG), "N": 3 * (n + 2 * NG)}[side]
                arr[t] = bcs[b][side][t - cc * n]
                st[t] = mesh.wo[bo[b] + k0 + t - cc * n]
            bc.append(arr)
            wo.append(st)
        ni.append(N)
        lat.append((x * n, y * n))
        lev.append(lv)
        # cells: merged (i, j) <- leaf (ci, cj) cell (i mod n, j mod n)
        jj, ii = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
        leaf = np.array([[bid[(lv, x + a, y + c)] for a in range(g)] for c in range(g)])
        b_of = leaf[jj // n, ii // n]
        cmap.append((co[b_of] + (jj % n) * n + (ii % n)).ravel())
    nb = len(groups)
    bc = np.concatenate(bc)
    out = Mesh(nblk=nb, ni=np.array(ni, np.int32), nj=np.array(ni, np.int32), xnode=np.concatenate(xs),
               ynode=np.concatenate(ys), nbr=np.full((nb, 8), -1, np.int32), bc=bc, wo=np.concatenate(wo),
               info={"kind": "merged_quadtree", "n": n, "depth": depth, "max_level": ml,
                     "groups": groups, "from": mesh.info},
               level=np.array(lev, np.int32), lattice=np.array(lat, np.int32).reshape(-1, 2))
    return out, np.concatenate(cmap)


def circle_refinement(radius=0.3, centre=(0.5, 0.5), band=0.0):
    """Refinement test of quadtree_box: split blocks whose square comes within
    band of the circle."""
    cx, cy = centre

    def refine(lv, xa, xb, ya, yb):
        dx = max(xa - cx, 0.0, cx - xb)
        dy = max(ya - cy, 0.0, cy - yb)
        dmin = np.hypot(dx, dy)
        dmax = max(np.hypot(xa - cx, ya - cy), np.hypot(xb - cx, ya - cy), np.hypot(xa - cx, yb - cy),
                   np.hypot(xb - cx, yb - cy))
        return dmin - band <= radius <= dmax + band
    return refine


def cylindrical_explosion(mesh, radius=0.3, centre=(0.5, 0.5), inside=(8.0, 0.0, 0.0, 10.0),
                          outside=(1.0, 0.0, 0.0, 1.0)):
    """Pressure ratio 10 and density ratio 8 across a circle, at rest: the
    states of the reference's IC_CYLINDRICAL_EXPLOSION
    (Euler2DQuadSingleBlock.cc:1978-1990), with the radius as a parameter."""
    c = mesh.cell_centres()
    r = np.hypot(c[:, 0] - centre[0], c[:, 1] - centre[1])
    W = np.empty((mesh.ncell, 4))
    W[r <= radius] = inside
    W[r > radius] = outside
    return W


def concat(meshes):
    """One multi-block mesh holding several disjoint meshes (cases): block
    indices of each part are offset so that neighbour tables stay inside the
    part.  Returns the mesh and the first block of every part."""
    if any(m.lattice is not None and m.info.get("kind") != "quadtree" for m in meshes):
        raise ValueError("meshes placed by lattice position cannot be concatenated")
    starts, nb = [], 0
    nbrs = []
    for m in meshes:
        starts.append(nb)
        t = np.array(m.nbr, dtype=np.int64)
        nbrs.append(np.where(t >= 0, t + nb, -1).astype(np.int32))
        nb += m.nblk
    out = Mesh(nblk=nb, ni=np.concatenate([m.ni for m in meshes]), nj=np.concatenate([m.nj for m in meshes]),
               xnode=np.concatenate([m.xnode for m in meshes]), ynode=np.concatenate([m.ynode for m in meshes]),
               nbr=np.concatenate(nbrs), bc=np.concatenate([m.bc for m in meshes]),
               wo=np.concatenate([m.wo for m in meshes]), info={"kind": "concat", "parts": len(meshes)})
    if any(m.level is not None for m in meshes):
        out.level = np.concatenate([m.level if m.level is not None else np.zeros(m.nblk, np.int32)
                                    for m in meshes]).astype(np.int32)
        n2 = []
        for m, s in zip(meshes, starts):
            t = np.full((m.nblk, 4), -1, np.int64) if m.nbr2 is None else np.array(m.nbr2, dtype=np.int64)
            n2.append(np.where(t >= 0, t + s, -1).astype(np.int32))
        out.nbr2 = np.concatenate(n2)
    return out, starts


def freestream(mach, rho=1.4, p=1.0, gamma=1.4):
    """(rho, u, v, p) of a uniform stream at Mach number mach along +x."""
    a = np.sqrt(gamma * p / rho)
    return (rho, mach * a, 0.0, p)


# ---- initial data -------------------------------------------------------------
RIEMANN_4Q = {  # (rho, u, v, p) by quadrant of [0,1]^2 split at (0.5, 0.5)
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def riemann_4quadrant(mesh, split=(0.5, 0.5), states=RIEMANN_4Q):
    c = mesh.cell_centres()
    east = c[:, 0] > split[0]
    north = c[:, 1] > split[1]
    W = np.empty((mesh.ncell, 4))
    W[east & north] = states["NE"]
    W[~east & north] = states["NW"]
    W[~east & ~north] = states["SW"]
    W[east & ~north] = states["SE"]
    return W


def density_perturbation(mesh, W, amplitude=0.01, seed=0):
    """rho <- rho*(1 + amplitude*xi), xi ~ U[-1,1] i.i.d. per cell, keyed by
    the global cell index through SplitMix64 so any partition of the mesh
    reproduces the same field."""
    n = mesh.ncell
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = idx + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    xi = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0
    W = np.array(W, copy=True)
    W[:, 0] = W[:, 0] * (1.0 + amplitude * xi)
    return W


def uniform_state(mesh, W0):
    return np.tile(np.asarray(W0, dtype=np.float64), (mesh.ncell, 1))
