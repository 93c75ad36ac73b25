"""FAS multigrid on the explicit residual: the host side of
FAS_Multigrid2D_Solver (FASMultigrid2D/FASMultigrid2D.h).

``coarsen`` builds the coarse level of a multi-block mesh the way the
reference's allocate (FASMultigrid2D.h:418-658) does, block by block:

* Half_Mesh_Resolution (HO_Grid2DQuad.cc:12614-12755): every other interior
  node, the same boundary-condition types;
* Update_Exterior_Nodes (HO_Grid2DQuad.cc:4153-4590): ghost nodes outside a
  block edge are the edge node minus the interior offset (a point
  reflection), or, on reflecting walls, the interior offset mirrored about
  the mean of the two adjacent face normals (end nodes point-reflected);
  then the ghost-row passes at the four corners;
* the geometry messages (Exchange_Solution_Information(level, ON),
  FASMultigrid2D.h:2235-2290): ghost nodes inside a neighbour's region are
  that neighbour's coarse nodes (faces and corners).

Update_Corner_Ghost_Nodes and the spline-consistency pass of the reference
are not restated: they move ghost nodes only at block corners where more
than four blocks meet, or on curved splined boundaries next to resolution
changes, neither of which this path meshes.

``MGParams`` mirrors Multigrid_Input_Parameters (FASMultigrid2DInput.h:93-114,
defaults there); ``FASMultigrid`` drives the device solver through the C ABI
(cffc_mg_*).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi
from .mesh import Mesh, NG


@dataclass
class MGParams:
    """Multigrid_Input_Parameters fields on this path (defaults as the
    reference's, FASMultigrid2DInput.h:93-114)."""
    levels: int = 2
    i_cycle: int = abi.MULTIGRID_V_CYCLE
    smooths_on_finest_level: int = 1
    pre_smooths: int = 1
    post_smooths: int = 0
    smooths_on_coarsest_level: int = 3
    defect_correction: int = 1
    prolong_using_injection: int = 0
    first_order_coarse: int = 1
    injection_at_dirichlet: int = 1
    update_stability_switch: int = 1
    max_update_reductions: int = 5
    absolute_tolerance: float = -1.0
    relative_tolerance: float = -1.0
    fmg_cycles: int = 0
    fmg_absolute_tolerance: float = -1.0
    fmg_relative_tolerance: float = -1.0


def mg_desc(p):
    return abi.MGDesc(p.levels, p.i_cycle, p.smooths_on_finest_level, p.pre_smooths, p.post_smooths,
                      p.smooths_on_coarsest_level, p.defect_correction, p.prolong_using_injection,
                      p.first_order_coarse, p.injection_at_dirichlet, p.update_stability_switch,
                      p.max_update_reductions, p.absolute_tolerance, p.relative_tolerance,
                      p.fmg_cycles, 0, p.fmg_absolute_tolerance, p.fmg_relative_tolerance)


_WALLS = (abi.BC_REFLECTION,)


def _normal(xa, ya, xb, yb):
    """unit normal of the edge a -> b, (dy, -dx) / l (nfaceN/S/E/W,
    HO_Grid2DQuad.h:1972-2000: the right of the direction of travel)"""
    dx, dy = xb - xa, yb - ya
    l = (dx * dx + dy * dy) ** 0.5
    return (dy / l, -dx / l)


class _Nodes:
    """One block's node arrays X[j, i], Y[j, i] (ghosts included), with
    scalar (x, y) access."""

    def __init__(self, ni, nj):
        self.ni, self.nj = ni, nj
        self.X = np.zeros((nj + 2 * NG + 1, ni + 2 * NG + 1))
        self.Y = np.zeros_like(self.X)

    def get(self, i, j):
        return (float(self.X[j, i]), float(self.Y[j, i]))

    def put(self, i, j, v):
        self.X[j, i], self.Y[j, i] = v[0], v[1]

    # outward face normals of cell (i, j) from its nodes
    def nW(self, i, j):
        return _normal(self.X[j + 1, i], self.Y[j + 1, i], self.X[j, i], self.Y[j, i])

    def nE(self, i, j):
        return _normal(self.X[j, i + 1], self.Y[j, i + 1], self.X[j + 1, i + 1], self.Y[j + 1, i + 1])

    def nS(self, i, j):
        return _normal(self.X[j, i], self.Y[j, i], self.X[j, i + 1], self.Y[j, i + 1])

    def nN(self, i, j):
        return _normal(self.X[j + 1, i + 1], self.Y[j + 1, i + 1], self.X[j + 1, i], self.Y[j + 1, i])


def _avg(a, b, sign):
    """sign * HALF * (a + b) for normals a, b"""
    return (sign * 0.5 * (a[0] + b[0]), sign * 0.5 * (a[1] + b[1]))


def _ext(base, inner):
    """point reflection: base + (base - inner)"""
    return (base[0] + (base[0] - inner[0]), base[1] + (base[1] - inner[1]))


def _mirror(base, inner, n, sign):
    """base - X_norm + X_tan (sign -1) or base + X_norm - X_tan (sign +1) with
    d = inner - base (sign -1) or base - inner (sign +1)"""
    if sign < 0:
        d = (inner[0] - base[0], inner[1] - base[1])
    else:
        d = (base[0] - inner[0], base[1] - inner[1])
    dn = d[0] * n[0] + d[1] * n[1]
    xn = (dn * n[0], dn * n[1])
    xt = (d[0] - xn[0], d[1] - xn[1])
    if sign < 0:
        return (base[0] - xn[0] + xt[0], base[1] - xn[1] + xt[1])
    return (base[0] + xn[0] - xt[0], base[1] + xn[1] - xt[1])


def _periodic(base, far, far_in):
    """base + (far_in - far): the opposite edge's interior offset"""
    return (base[0] + (far_in[0] - far[0]), base[1] + (far_in[1] - far[1]))


def _exterior_nodes(N, bcW, bcE, bcS, bcN):
    """Update_Exterior_Nodes on one coarse block (cells indexed with ghosts:
    ICl = JCl = NG; nodes INl = JNl = NG)."""
    INl, INu, JNl, JNu = NG, NG + N.ni, NG, NG + N.nj
    ICl, ICu, JCl, JCu = NG, NG + N.ni - 1, NG, NG + N.nj - 1
    G = range(1, NG + 1)
    for j in range(JNl, JNu + 1):
        for bc, side in ((bcW, -1), (bcE, 1)):
            t = bc[j]
            if t != bc[j - 1]:
                continue
            I0 = INl if side < 0 else INu
            for g in G:
                if t == abi.BC_PERIODIC:
                    I1 = INu if side < 0 else INl
                    N.put(I0 + side * g, j, _periodic(N.get(I0, j), N.get(I1, j), N.get(I1 + side * g, j)))
                elif t in _WALLS and JNl < j < JNu:
                    n = (_avg(N.nW(ICl, j), N.nW(ICl, j - 1), -1) if side < 0 else
                         _avg(N.nE(ICu, j), N.nE(ICu, j - 1), 1))
                    N.put(I0 + side * g, j, _mirror(N.get(I0, j), N.get(I0 - side * g, j), n, side))
                else:
                    N.put(I0 + side * g, j, _ext(N.get(I0, j), N.get(I0 - side * g, j)))
    for i in range(INl, INu + 1):
        for bc, side in ((bcS, -1), (bcN, 1)):
            t = bc[i]
            if t != bc[i - 1]:
                continue
            J0 = JNl if side < 0 else JNu
            for g in G:
                if t == abi.BC_PERIODIC:
                    J1 = JNu if side < 0 else JNl
                    N.put(i, J0 + side * g, _periodic(N.get(i, J0), N.get(i, J1), N.get(i, J1 + side * g)))
                elif t in _WALLS and INl < i < INu:
                    n = (_avg(N.nS(i, JCl), N.nS(i - 1, JCl), -1) if side < 0 else
                         _avg(N.nN(i, JCu), N.nN(i - 1, JCu), 1))
                    N.put(i, J0 + side * g, _mirror(N.get(i, J0), N.get(i, J0 - side * g), n, side))
                else:
                    N.put(i, J0 + side * g, _ext(N.get(i, J0), N.get(i, J0 - side * g)))
    # SW/SE then NW/NE corner rows: the W/E rule on the ghost rows, with the
    # BC type of the row (the cell below the node row for the north corners)
    for rows, below in ((range(JNl - NG, JNl + 1), 0), (range(JNu, JNu + NG + 1), 1)):
        for j in rows:
            jb = j - below
            for bc, side in ((bcW, -1), (bcE, 1)):
                t = bc[jb]
                I0 = INl if side < 0 else INu
                ic = ICl if side < 0 else ICu
                nf = N.nW if side < 0 else N.nE
                for g in G:
                    if t == abi.BC_PERIODIC:
                        I1 = INu if side < 0 else INl
                        N.put(I0 + side * g, j, _periodic(N.get(I0, j), N.get(I1, j), N.get(I1 + side * g, j)))
                    elif t in _WALLS:
                        if below == 0:
                            nn = nf(ic, j) if j == JNl - NG else _avg(nf(ic, j), nf(ic, j - 1), 1)
                        else:
                            nn = nf(ic, j - 1) if j == JNu + NG else _avg(nf(ic, j), nf(ic, j - 1), 1)
                        n = (-nn[0], -nn[1]) if side < 0 else nn
                        N.put(I0 + side * g, j, _mirror(N.get(I0, j), N.get(I0 - side * g, j), n, side))
                    else:
                        N.put(I0 + side * g, j, _ext(N.get(I0, j), N.get(I0 - side * g, j)))


_SLOT = {(0, 1): 0, (0, -1): 1, (1, 0): 2, (-1, 0): 3, (1, 1): 4, (-1, 1): 5, (1, -1): 6, (-1, -1): 7}


def coarsen(mesh):
    """The next coarser multigrid level of a single-level multi-block mesh
    (all blocks with even ni, nj)."""
    if getattr(mesh, "level", None) is not None and np.any(np.asarray(mesh.level) != mesh.level[0]):
        raise ValueError("multigrid levels are built for single-level meshes only")
    nb = mesh.nblk
    if np.any(mesh.ni % 2) or np.any(mesh.nj % 2):
        raise ValueError("every block needs even ni and nj to coarsen (FASMultigrid2D.h:470-486)")
    if np.any(mesh.ni < 2 * NG) or np.any(mesh.nj < 2 * NG):
        # Half_Mesh_Resolution refuses blocks of fewer than 2*Nghost cells a
        # side (HO_Grid2DQuad.cc:12618-12621): the coarse block must still
        # fill its neighbours' two ghost layers
        raise ValueError("blocks need at least %d cells a side to coarsen" % (2 * NG))
    cni, cnj = mesh.ni // 2, mesh.nj // 2
    noff = mesh.node_offsets()
    boff = mesh.bc_offsets()
    nodes, bcs = [], []
    for b in range(nb):
        ni, nj = int(mesh.ni[b]), int(mesh.nj[b])
        Xf, Yf = mesh.block_nodes(b)
        N = _Nodes(int(cni[b]), int(cnj[b]))
        sl = (slice(NG, NG + nj + 1, 2), slice(NG, NG + ni + 1, 2))
        N.X[NG:NG + N.nj + 1, NG:NG + N.ni + 1] = Xf[sl]
        N.Y[NG:NG + N.nj + 1, NG:NG + N.ni + 1] = Yf[sl]
        nodes.append(N)
        # BC types: interior rows from the first fine row of each coarse row,
        # ghost rows from the fine ghost rows
        o = boff[b]
        ncj, nci = nj + 2 * NG, ni + 2 * NG
        fW, fE = mesh.bc[o:o + ncj], mesh.bc[o + ncj:o + 2 * ncj]
        fS, fN = mesh.bc[o + 2 * ncj:o + 2 * ncj + nci], mesh.bc[o + 2 * ncj + nci:o + 2 * ncj + 2 * nci]
        wo = mesh.wo[o:o + 2 * ncj + 2 * nci]
        fwo = (wo[:ncj], wo[ncj:2 * ncj], wo[2 * ncj:2 * ncj + nci], wo[2 * ncj + nci:])

        def pick(arr, n_c, n_f):
            idx = np.concatenate([np.arange(NG), NG + 2 * np.arange(n_c), n_f + NG + np.arange(NG)])
            return arr[idx]
        c_bc = (pick(fW, N.nj, nj), pick(fE, N.nj, nj), pick(fS, N.ni, ni), pick(fN, N.ni, ni))
        c_wo = (pick(fwo[0], N.nj, nj), pick(fwo[1], N.nj, nj), pick(fwo[2], N.ni, ni), pick(fwo[3], N.ni, ni))
        bcs.append((c_bc, c_wo))
    for b in range(nb):
        N = nodes[b]
        _exterior_nodes(N, *bcs[b][0])
    # geometry messages: ghost nodes inside a neighbour's region (faces over
    # the interior span of node lines, corners 2 x 2)
    out = [(nodes[b].X.copy(), nodes[b].Y.copy()) for b in range(nb)]
    for b in range(nb):
        N = nodes[b]
        X, Y = out[b]
        INu, JNu = NG + N.ni, NG + N.nj
        spans = {-1: slice(0, NG), 0: None, 1: slice(INu + 1, INu + NG + 1)}
        jspans = {-1: slice(0, NG), 0: None, 1: slice(JNu + 1, JNu + NG + 1)}
        for (di, dj), slot in _SLOT.items():
            o = int(mesh.nbr[b, slot])
            if o < 0:
                continue
            M = nodes[o]
            si = spans[di] if di else slice(NG, INu + 1)
            sj = jspans[dj] if dj else slice(NG, JNu + 1)
            shift_i = M.ni if di < 0 else (-N.ni if di > 0 else 0)
            shift_j = M.nj if dj < 0 else (-N.nj if dj > 0 else 0)
            src_i = slice(si.start + shift_i, si.stop + shift_i)
            src_j = slice(sj.start + shift_j, sj.stop + shift_j)
            X[sj, si] = M.X[src_j, src_i]
            Y[sj, si] = M.Y[src_j, src_i]
    xs = np.concatenate([out[b][0].ravel() for b in range(nb)])
    ys = np.concatenate([out[b][1].ravel() for b in range(nb)])
    bc = np.concatenate([np.concatenate(bcs[b][0]) for b in range(nb)]).astype(np.int32)
    wo = np.concatenate([np.concatenate(bcs[b][1]) for b in range(nb)])
    m = Mesh(nblk=nb, ni=cni.astype(np.int32), nj=cnj.astype(np.int32), xnode=xs, ynode=ys,
             nbr=mesh.nbr.copy(), bc=bc, wo=wo, info=dict(mesh.info, multigrid_parent=mesh.info.get("kind")))
    return m


def hierarchy(mesh, levels):
    """[mesh, coarsen(mesh), ...]: `levels` meshes, finest first."""
    out = [mesh]
    for _ in range(levels - 1):
        out.append(coarsen(out[-1]))
    return out


def _level_view(h, mesh, scheme):
    from .euler2d import Euler2DQuadBlocks
    v = Euler2DQuadBlocks.__new__(Euler2DQuadBlocks)
    v.h = ctypes.c_void_p(h)
    v.mesh = mesh
    v.scheme = scheme
    v.ncell = _lib().cffc_num_cells(v.h)
    v._owned = False
    return v


def _lib():
    from .euler2d import load
    return load()


class FASMultigrid:
    """FAS_Multigrid2D_Solver on the GPU: allocate (meshes built by
    ``hierarchy``), Execute's regular V/W cycles.  Level 0 is the finest;
    ``levels[l]`` are Euler2DQuadBlocks views of the level contexts."""

    def __init__(self, mesh, scheme, params=None, meshes=None):
        from .euler2d import _check
        self.params = params or MGParams()
        self.meshes = meshes or hierarchy(mesh, self.params.levels)
        self.scheme = scheme
        self._keep = []
        arr = (abi.MeshDesc * len(self.meshes))(*[abi.mesh_desc(m, self._keep) for m in self.meshes])
        self._arr = arr
        self._sd = abi.scheme_desc(scheme)
        self._mgd = mg_desc(self.params)
        h = ctypes.c_void_p()
        _check(_lib().cffc_mg_create(ctypes.cast(arr, ctypes.c_void_p), len(self.meshes), ctypes.byref(self._sd),
                                     ctypes.byref(self._mgd), ctypes.byref(h)), "cffc_mg_create")
        self.h = h
        self.levels = [_level_view(_lib().cffc_mg_level(h, l), self.meshes[l], scheme)
                       for l in range(len(self.meshes))]
        self.ncell = self.levels[0].ncell

    def close(self):
        if getattr(self, "h", None):
            for v in self.levels:
                v.h = None
            _lib().cffc_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_W(self, W):
        self.levels[0].set_W(W)

    def get_W(self):
        return self.levels[0].get_W()

    def get_U(self):
        return self.levels[0].get_U()

    def set_residual_smoothing(self, epsilon, iterations):
        from .euler2d import _check
        _check(_lib().cffc_mg_set_residual_smoothing(self.h, epsilon, iterations), "cffc_mg_set_residual_smoothing")

    def initialize(self):
        from .euler2d import _check
        _check(_lib().cffc_mg_initialize(self.h), "cffc_mg_initialize")

    def cycle(self):
        """One FAS cycle; returns 0 or the failing update's i index."""
        from .euler2d import _check
        bad = ctypes.c_int64(0)
        return _check(_lib().cffc_mg_cycle(self.h, ctypes.byref(bad)), "cffc_mg_cycle")

    def set_graph(self, on):
        """execute replays a captured cycle (default) or issues the launches."""
        from .euler2d import _check
        _check(_lib().cffc_mg_set_graph(self.h, 1 if on else 0), "cffc_mg_set_graph")

    def get_P(self, level):
        from .euler2d import _check, _p
        out = np.empty((self.levels[level].ncell, 4))
        _check(_lib().cffc_mg_get_P(self.h, level, _p(out), self.levels[level].ncell), "cffc_mg_get_P")
        return out

    def set_parts(self, first_block, tol, freeze=None):
        """Per-case stopping (cffc_mg_set_parts): part q owns level-0 blocks
        first_block[q] .. first_block[q+1]-1; execute's regular cycles stop
        once every part's own L2 norm has met tol[q], and the limiters
        freeze once every part's norm is at or below freeze[q]."""
        from .euler2d import _check, _p
        fb = np.ascontiguousarray(first_block, dtype=np.int32)
        self._part_tol = np.ascontiguousarray(tol, dtype=np.float64)
        self._part_frz = None if freeze is None else np.ascontiguousarray(freeze, dtype=np.float64)
        npart = fb.size - 1
        if self._part_tol.size != npart or (self._part_frz is not None and self._part_frz.size != npart):
            raise ValueError("one tolerance (and freeze level) per part")
        _check(_lib().cffc_mg_set_parts(self.h, npart, fb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        _p(self._part_tol), None if self._part_frz is None else _p(self._part_frz)),
               "cffc_mg_set_parts")
        self.npart = npart

    def part_status(self):
        """(last L2 norm per part, regular cycles done when each part first
        met its tolerance, -1 before)"""
        from .euler2d import _check, _p
        l2 = np.empty(self.npart)
        done = np.empty(self.npart, dtype=np.int32)
        _check(_lib().cffc_mg_part_status(self.h, _p(l2), done.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "cffc_mg_part_status")
        return l2, done

    def execute(self, max_cycles, freeze_level=0.0):
        """Execute (full-multigrid start-up when fmg_cycles > 0, then the
        regular cycles); returns (rc, cycles, norms[checks, 3]), the norms
        of each check on that check's top level."""
        from .euler2d import _check, _p
        norms = np.zeros((max_cycles + len(self.levels), 3))
        done, chk = ctypes.c_int(0), ctypes.c_int(0)
        rc = _check(_lib().cffc_mg_execute(self.h, max_cycles, freeze_level, _p(norms), ctypes.byref(done),
                                           ctypes.byref(chk)), "cffc_mg_execute")
        return rc, done.value, norms[:chk.value]
