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


def concat(meshes):
    """One multi-block mesh holding several disjoint meshes (cases): block
    indices of each part are offset so that neighbour tables stay inside the
    part.  Returns the mesh and the first block of every part."""
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
