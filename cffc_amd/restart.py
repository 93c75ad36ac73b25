"""The reference's restart files: per-block `<prefix>_blkNNNNNN.soln` and the
quadtree `<prefix>_quadtree.tree`.

Layouts, from the writers and readers that produce and consume them:
  * `.soln` (Write_Restart_Solution / Read_Restart_Solution,
    Euler2DQuadMultiBlock.cc:145-275): a header line "nsteps time  cputime"
    (scientific, 14 digits), the gas type, then the block (operator<<,
    Euler2DQuad.h:1307-1399):
      - its grid (operator<< of Grid2D_Quad_Block_HO, HO_Grid2DQuad.cc:
        13763-14000): NNi INl INu / NNj JNl JNu / Nghost / MaxRecOrder /
        HighOrderBoundary / NCi ICl ICu / NCj JCl JCu, the nodes with ghosts
        (j outer, i inner, " x y", 15 digits), the cells (" I J x y A" and
        their geometric moments, Cell2D_HO operator<<, HO_Cell2D.h:275-283),
        the BC types ("N S" per i, then "E W" per j), twelve boundary-spline
        sections, the spline parameter ranges, the stretching parameters and
        the orthogonality flags;
      - NCi ICl ICu Nghost / NCj JCl JCu / Axisymmetric, U of every cell
        ghosts included (" d dx dy E", 15 digits), WoW/WoE per j, WoS/WoN
        per i, and the high-order variables (a count, then each).
  * `.tree` (Write_QuadTree / Read_QuadTree, AMR.h:169-333; QuadTreeBlock::
    write, QuadTree.h:431-443; AdaptiveBlock2D operator<<, AdaptiveBlock.h:
    681-718): "NRi NRj Ncpu Nblk", then every root (j outer, i inner),
    recursively: " used gblknum cpu blknum ni nj ghost sector level", the
    neighbour counts and infos N, S, E, W, NW, NE, SE, SW one per line, and
    " nchildren", children SW, SE, NW, NE.

read_soln/write_soln reproduce the reference's own files byte for byte
(tests/test_restart.py, on the restart files of the reference's Ringleb-flow
unit tests).  Grid sections this package did not read (boundary splines,
high-order moments) are kept as text.  to_mesh turns blocks with any
number of ghost layers into a cffc_amd.mesh.Mesh with the solver's two.
"""
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

NG = 2


def _e15(v):
    return "%.15e" % v


@dataclass
class SolnBlock:
    header: str                  # "nsteps time  cputime" line as written
    gas: str
    NNi: int
    INl: int
    INu: int
    NNj: int
    JNl: int
    JNu: int
    nghost: int
    max_rec_order: int
    ho_boundary: int
    NCi: int
    ICl: int
    ICu: int
    NCj: int
    JCl: int
    JCu: int
    xn: np.ndarray               # [NNj, NNi] nodes, ghosts included
    yn: np.ndarray
    cells: np.ndarray            # [NCj, NCi, 3] Xc.x, Xc.y, A
    cell_tails: list             # geometric-moment text of every cell (after " I J x y A\n")
    bcN: np.ndarray              # [NCi]
    bcS: np.ndarray
    bcE: np.ndarray              # [NCj]
    bcW: np.ndarray
    grid_tail: str               # splines, ranges, stretching, orthogonality (verbatim)
    axisymmetric: int
    U: np.ndarray                # [NCj, NCi, 4]
    WoW: np.ndarray              # [NCj, 4]
    WoE: np.ndarray
    WoS: np.ndarray              # [NCi, 4]
    WoN: np.ndarray
    tail: str                    # high-order variable count and data (verbatim)
    state_digits: int = 15       # digits after the point of U and Wo (15 as Write_Restart_Solution leaves it)

    @property
    def nsteps(self):
        return int(self.header.split()[0])

    @property
    def time(self):
        return float(self.header.split()[1])


def _ints(line):
    return [int(t) for t in line.split()]


def _read_text(path):
    if path.endswith(".gz"):
        import gzip
        with gzip.open(path, "rt") as f:
            return f.read()
    with open(path) as f:
        return f.read()


def _parse_grid(lines, k):
    """The grid of one block (operator>> of Grid2D_Quad_Block_HO) starting
    at line k: a dict of its fields and the line after it.  The grid ends
    with its orthogonality flags (4 integers)."""
    g = {}
    g["NNi"], g["INl"], g["INu"] = _ints(lines[k])
    g["NNj"], g["JNl"], g["JNu"] = _ints(lines[k + 1])
    g["nghost"] = int(lines[k + 2])
    g["max_rec_order"], g["ho_boundary"] = int(lines[k + 3]), int(lines[k + 4])
    g["NCi"], g["ICl"], g["ICu"] = _ints(lines[k + 5])
    g["NCj"], g["JCl"], g["JCu"] = _ints(lines[k + 6])
    k += 7
    NNi, NNj, NCi, NCj = g["NNi"], g["NNj"], g["NCi"], g["NCj"]
    nn = NNi * NNj
    xy = np.array([[float(t) for t in lines[k + q].split()] for q in range(nn)])
    k += nn
    g["xn"], g["yn"] = xy[:, 0].reshape(NNj, NNi), xy[:, 1].reshape(NNj, NNi)
    ncell = NCi * NCj
    cells = np.empty((ncell, 3))
    tails = []
    for q in range(ncell):
        tok = lines[k].split()
        i, j = int(tok[0]), int(tok[1])
        if (i, j) != (q % NCi, q // NCi):
            raise ValueError("cell record %d is (%d, %d)" % (q, i, j))
        cells[q] = [float(tok[2]), float(tok[3]), float(tok[4])]
        k += 1
        t0 = k
        if q + 1 < ncell:
            nxt = " %d %d " % ((q + 1) % NCi, (q + 1) // NCi)
            while not lines[k].startswith(nxt):
                k += 1
        else:
            while len(lines[k].split()) != 2:
                k += 1
        tails.append("\n".join(lines[t0:k]) + "\n")
    g["cells"], g["cell_tails"] = cells.reshape(NCj, NCi, 3), tails
    bcNS = np.array([_ints(lines[k + q]) for q in range(NCi)], dtype=np.int32)
    k += NCi
    bcEW = np.array([_ints(lines[k + q]) for q in range(NCj)], dtype=np.int32)
    k += NCj
    g["bcN"], g["bcS"], g["bcE"], g["bcW"] = bcNS[:, 0], bcNS[:, 1], bcEW[:, 0], bcEW[:, 1]
    # twelve spline sections (kept verbatim), then the spline parameter
    # ranges (two lines of 4 reals), the stretching parameters (two lines
    # "int real real") and the orthogonality flags (4 integers)
    t0 = k
    while not _grid_end_at(lines, k):
        k += 1
    k += 5
    g["grid_tail"] = "\n".join(lines[t0:k]) + "\n"
    return g, k


def _is_real(t):
    return ("e" in t or "." in t) and t.lstrip("-+").replace(".", "", 1).replace("e", "", 1).replace("-", "").replace(
        "+", "").isdigit()


def _grid_end_at(lines, k):
    """lines[k:k+5] are the ranges, stretching and orthogonality lines that
    close a grid (HO_Grid2DQuad.cc:13989-13996)"""
    if k + 4 >= len(lines):
        raise ValueError("grid section does not end")
    r1, r2, s1, s2, o = (lines[k + q].split() for q in range(5))
    if len(r1) != 4 or len(r2) != 4 or len(s1) != 3 or len(s2) != 3 or len(o) != 4:
        return False
    if not all(_is_real(t) for t in r1 + r2 + s1[1:] + s2[1:]):
        return False
    return all(t.lstrip("-").isdigit() for t in o + [s1[0], s2[0]])


def read_soln(path):
    """Parse one block restart file (gzip-compressed if it ends in .gz)."""
    lines = _read_text(path).split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    header, gas = lines[0], lines[1]
    g, k = _parse_grid(lines, 2)
    NCi, NCj, ICl, ICu, JCl, JCu = g["NCi"], g["NCj"], g["ICl"], g["ICu"], g["JCl"], g["JCu"]
    if lines[k] != "%d %d %d %d" % (NCi, ICl, ICu, g["nghost"]) or _ints(lines[k + 1]) != [NCj, JCl, JCu]:
        raise ValueError("%s: unexpected solution header %r" % (path, lines[k]))
    k += 2
    axisym = int(lines[k])
    k += 1
    ncell = NCi * NCj
    mant = lines[k].split()[0].lstrip("-").split("e")[0]
    digits = len(mant.split(".")[1]) if "." in mant else 15
    U = np.array([[float(t) for t in lines[k + q].split()] for q in range(ncell)]).reshape(NCj, NCi, 4)
    k += ncell
    wo = np.array([[float(t) for t in lines[k + q].split()] for q in range(2 * NCj)])
    k += 2 * NCj
    wo2 = np.array([[float(t) for t in lines[k + q].split()] for q in range(2 * NCi)])
    k += 2 * NCi
    tail = "\n".join(lines[k:]) + "\n"
    return SolnBlock(header=header, gas=gas, axisymmetric=axisym, U=U, WoW=wo[0::2], WoE=wo[1::2],
                     WoS=wo2[0::2], WoN=wo2[1::2], tail=tail, state_digits=digits, **g)


def _format_grid(B):
    out = ["%d %d %d\n" % (B.NNi, B.INl, B.INu), "%d %d %d\n" % (B.NNj, B.JNl, B.JNu),
           "%d\n" % B.nghost, "%d\n" % B.max_rec_order, "%d\n" % B.ho_boundary,
           "%d %d %d\n" % (B.NCi, B.ICl, B.ICu), "%d %d %d\n" % (B.NCj, B.JCl, B.JCu)]
    out += [" %s %s\n" % (_e15(x), _e15(y)) for x, y in zip(B.xn.ravel(), B.yn.ravel())]
    c = B.cells.reshape(-1, 3)
    for q in range(c.shape[0]):
        out.append(" %d %d %s %s %s\n" % (q % B.NCi, q // B.NCi, _e15(c[q, 0]), _e15(c[q, 1]), _e15(c[q, 2])))
        out.append(B.cell_tails[q])
    out += ["%d %d\n" % (n, s) for n, s in zip(B.bcN, B.bcS)]
    out += ["%d %d\n" % (e, w) for e, w in zip(B.bcE, B.bcW)]
    out.append(B.grid_tail)
    return out


def format_soln(B):
    """The text of a block restart file (inverse of read_soln)."""
    out = [B.header + "\n", B.gas + "\n"] + _format_grid(B)
    out += ["%d %d %d %d\n" % (B.NCi, B.ICl, B.ICu, B.nghost), "%d %d %d\n" % (B.NCj, B.JCl, B.JCu),
            "%d\n" % B.axisymmetric]
    fs = " %%.%de" % B.state_digits
    row = fs * 4 + "\n"
    out += [row % tuple(u) for u in B.U.reshape(-1, 4)]
    for w, e in zip(B.WoW, B.WoE):
        out.append(row % tuple(w))
        out.append(row % tuple(e))
    for s_, n_ in zip(B.WoS, B.WoN):
        out.append(row % tuple(s_))
        out.append(row % tuple(n_))
    out.append(B.tail)
    return "".join(out)


def write_soln(path, B):
    with open(path, "w") as f:
        f.write(format_soln(B))


# ---- grid data files (Grid_Type = Read_From_Data_File) -----------------------------
@dataclass
class GridBlock:
    """One block of a grid data file: the grid fields of SolnBlock only."""
    NNi: int
    INl: int
    INu: int
    NNj: int
    JNl: int
    JNu: int
    nghost: int
    max_rec_order: int
    ho_boundary: int
    NCi: int
    ICl: int
    ICu: int
    NCj: int
    JCl: int
    JCu: int
    xn: np.ndarray
    yn: np.ndarray
    cells: np.ndarray
    cell_tails: list
    bcN: np.ndarray
    bcS: np.ndarray
    bcE: np.ndarray
    bcW: np.ndarray
    grid_tail: str


@dataclass
class GridFile:
    grid_type: str
    i_grid: int
    nbi: int
    nbj: int
    blocks: list            # [nbj][nbi] GridBlock (j outer, as written)


def read_grid_file(path):
    """A multi-block grid data file (Write_Multi_Block_Grid_Using_IP /
    Read_Multi_Block_Grid_Using_IP, HO_Grid2DQuadMultiBlock.h:3160-3210):
    the grid type, i_Grid, "NBi NBj", then every block's grid, j outer."""
    lines = _read_text(path).split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    grid_type, i_grid = lines[0], int(lines[1])
    nbi, nbj = _ints(lines[2])
    k = 3
    rows = []
    for _ in range(nbj):
        row = []
        for _ in range(nbi):
            g, k = _parse_grid(lines, k)
            row.append(GridBlock(**g))
        rows.append(row)
    if k != len(lines):
        raise ValueError("%s: %d lines after the last block" % (path, len(lines) - k))
    return GridFile(grid_type, i_grid, nbi, nbj, rows)


def format_grid_file(G):
    out = [G.grid_type + "\n", "%d\n" % G.i_grid, "%d %d\n" % (G.nbi, G.nbj)]
    for row in G.blocks:
        for B in row:
            out += _format_grid(B)
    return "".join(out)


def grid_file_to_mesh(G, wo=None):
    """The solver's Mesh of a grid data file's block array (two inner ghost
    layers kept; same-level neighbours from the array; the reference states
    Wo set to wo, or zero)."""
    from . import mesh as M
    nbi, nbj = G.nbi, G.nbj
    xs, ys, bc, ni, nj = [], [], [], [], []
    nbr = np.full((nbi * nbj, 8), -1, np.int32)
    for jb in range(nbj):
        for ib in range(nbi):
            B = G.blocks[jb][ib]
            g = B.nghost
            if g < NG:
                raise ValueError("a grid block with %d ghost layers: the solver needs %d" % (g, NG))
            d = g - NG
            xs.append(B.xn[d:B.NNj - d, d:B.NNi - d].ravel())
            ys.append(B.yn[d:B.NNj - d, d:B.NNi - d].ravel())
            ci, cj = slice(d, B.NCi - d), slice(d, B.NCj - d)
            bc += [B.bcW[cj], B.bcE[cj], B.bcS[ci], B.bcN[ci]]
            ni.append(B.NCi - 2 * g)
            nj.append(B.NCj - 2 * g)
            b = jb * nbi + ib

            def nb(di, dj):
                a, c = ib + di, jb + dj
                return c * nbi + a if (0 <= a < nbi and 0 <= c < nbj) else -1
            nbr[b] = [nb(0, 1), nb(0, -1), nb(1, 0), nb(-1, 0), nb(1, 1), nb(-1, 1), nb(1, -1), nb(-1, -1)]
    bc = np.concatenate(bc).astype(np.int32)
    wo_arr = np.zeros((bc.size, 4)) if wo is None else np.tile(np.asarray(wo, np.float64), (bc.size, 1))
    return M.Mesh(nblk=nbi * nbj, ni=np.array(ni, np.int32), nj=np.array(nj, np.int32), xnode=np.concatenate(xs),
                  ynode=np.concatenate(ys), nbr=nbr, bc=bc, wo=wo_arr,
                  info={"kind": "grid_file", "grid_type": G.grid_type, "nbi": nbi, "nbj": nbj})


# ---- quadtree files -------------------------------------------------------------
@dataclass
class TreeBlock:
    used: int
    gblknum: int
    info: list                    # cpu, blknum, ni, nj, ghost, sector, level
    neighbours: list              # 8 lists (N, S, E, W, NW, NE, SE, SW) of 7-int infos
    children: list = field(default_factory=list)   # SW, SE, NW, NE when refined


@dataclass
class QuadTreeFile:
    nri: int
    nrj: int
    ncpu: int
    nblk: int
    roots: list                   # [nrj][nri] TreeBlock


def read_tree(path):
    toks = _read_text(path).split()
    pos = [0]

    def nxt():
        v = int(toks[pos[0]])
        pos[0] += 1
        return v

    nri, nrj, ncpu, nblk = nxt(), nxt(), nxt(), nxt()

    def block():
        used, g = nxt(), nxt()
        info = [nxt() for _ in range(7)]
        nb = []
        for _ in range(8):
            n = nxt()
            nb.append([[nxt() for _ in range(7)] for _ in range(n)])
        nch = nxt()
        b = TreeBlock(used, g, info, nb)
        if not used and nch > 0:
            b.children = [block() for _ in range(4)]
        return b

    roots = [[block() for _ in range(nri)] for _ in range(nrj)]
    return QuadTreeFile(nri, nrj, ncpu, nblk, roots)


def format_tree(T):
    out = ["%d %d %d %d\n" % (T.nri, T.nrj, T.ncpu, T.nblk)]

    def info(v):
        return " %d %d %d %d %d %d %d" % tuple(v)

    def block(b):
        s = " %d %d%s" % (b.used, b.gblknum, info(b.info))
        for lst in b.neighbours:
            s += "\n %d" % len(lst) + "".join(" " + info(v) for v in lst)
        out.append(s + "\n")
        out.append(" %d\n" % (4 if (not b.used and b.children) else 0))
        for c in b.children:
            block(c)

    for row in T.roots:
        for b in row:
            block(b)
    return "".join(out)


def write_tree(path, T):
    with open(path, "w") as f:
        f.write(format_tree(T))


def used_blocks(T):
    """[(gblknum, level, ix, iy)] of the used blocks: (ix, iy) the block's
    position on its level's lattice (roots at level 0)."""
    out = []

    def walk(b, lv, ix, iy):
        if b.used:
            out.append((b.gblknum, lv, ix, iy))
        for k, c in enumerate(b.children):
            walk(c, lv + 1, 2 * ix + (k & 1), 2 * iy + (k >> 1))

    for jr, row in enumerate(T.roots):
        for ir, b in enumerate(row):
            walk(b, 0, ir, jr)
    return out


# ---- conversion to and from the solver's mesh and state ------------------------------
def to_mesh(blocks, nbr=None, level=None, nbr2=None):
    """A Mesh (two ghost layers) and the interior conserved state U of a list
    of SolnBlocks; blocks with more ghost layers keep the two innermost.
    nbr (and level, nbr2 across resolution changes) give the topology; a
    single block needs none."""
    from . import mesh as M
    xs, ys, bc, wo, U = [], [], [], [], []
    ni, nj = [], []
    for B in blocks:
        g = B.nghost
        if g < NG:
            raise ValueError("a restart block with %d ghost layers: the solver needs %d" % (g, NG))
        n_i, n_j = B.NCi - 2 * g, B.NCj - 2 * g
        d = g - NG
        xs.append(B.xn[d:B.NNj - d, d:B.NNi - d].ravel())
        ys.append(B.yn[d:B.NNj - d, d:B.NNi - d].ravel())
        ci, cj = slice(d, B.NCi - d), slice(d, B.NCj - d)
        bc += [B.bcW[cj], B.bcE[cj], B.bcS[ci], B.bcN[ci]]
        wo += [B.WoW[cj], B.WoE[cj], B.WoS[ci], B.WoN[ci]]
        U.append(B.U[g:g + n_j, g:g + n_i].reshape(-1, 4))
        ni.append(n_i)
        nj.append(n_j)
    nblk = len(blocks)
    if nbr is None:
        if nblk != 1:
            raise ValueError("several blocks need a neighbour table")
        nbr = np.full((1, 8), -1, np.int32)
    m = M.Mesh(nblk=nblk, ni=np.array(ni, np.int32), nj=np.array(nj, np.int32), xnode=np.concatenate(xs),
               ynode=np.concatenate(ys), nbr=np.asarray(nbr, np.int32), bc=np.concatenate(bc).astype(np.int32),
               wo=np.concatenate(wo).astype(np.float64), info={"kind": "restart"},
               level=None if level is None else np.asarray(level, np.int32),
               nbr2=None if nbr2 is None else np.asarray(nbr2, np.int32))
    return m, np.concatenate(U)


def blocks_from_mesh(mesh, U, nsteps=0, time=0.0, cpu_time=0.0, gas="AIR", axisymmetric=0):
    """SolnBlocks of a Mesh and its interior state U (ghost cells of U are
    written from the nearest interior cell; the reference rebuilds them by
    messages + BCs after reading).  The grids carry no boundary splines and
    zeroth-order geometric moments.  `axisymmetric` is the block's
    Axisymmetric flag (Euler2DQuad.h:1314), written after the grid."""
    U = np.asarray(U, np.float64).reshape(mesh.ncell, 4)
    co = mesh.cell_offsets()
    header = "%d %.14e  %.14e" % (nsteps, time, cpu_time)
    empty_spline = "0\n0\n" * 12
    grid_tail = (empty_spline + "%s %s %s %s\n" % ((_e15(0.0),) * 4) + "%s %s %s %s\n" % ((_e15(0.0),) * 4) +
                 "0 %s %s\n" % (_e15(1.0), _e15(0.0)) + "0 %s %s\n" % (_e15(1.0), _e15(0.0)) + "0 0 0 0\n")
    out = []
    for b in range(mesh.nblk):
        n_i, n_j = int(mesh.ni[b]), int(mesh.nj[b])
        NCi, NCj = n_i + 2 * NG, n_j + 2 * NG
        x, y = mesh.block_nodes(b)
        cells = np.zeros((NCj, NCi, 3))
        for j in range(NCj):
            for i in range(NCi):
                xs = [x[j, i], x[j, i + 1], x[j + 1, i + 1], x[j + 1, i]]
                ys = [y[j, i], y[j, i + 1], y[j + 1, i + 1], y[j + 1, i]]
                cells[j, i] = _centroid_area(xs, ys)
        Ub = U[co[b]:co[b + 1]].reshape(n_j, n_i, 4)
        jj = np.clip(np.arange(NCj) - NG, 0, n_j - 1)
        ii = np.clip(np.arange(NCi) - NG, 0, n_i - 1)
        bcd = mesh.block_bc(b)
        o = mesh.bc_offsets()[b]
        wo = mesh.wo[o:o + 2 * NCj + 2 * NCi]
        out.append(SolnBlock(
            header=header, gas=gas, NNi=NCi + 1, INl=NG, INu=NG + n_i, NNj=NCj + 1, JNl=NG, JNu=NG + n_j,
            nghost=NG, max_rec_order=0, ho_boundary=0, NCi=NCi, ICl=NG, ICu=NG + n_i - 1, NCj=NCj, JCl=NG,
            JCu=NG + n_j - 1, xn=x.copy(), yn=y.copy(), cells=cells,
            cell_tails=["    0\n   0   0                        1\n\n"] * (NCi * NCj),
            bcN=np.asarray(bcd["N"]), bcS=np.asarray(bcd["S"]), bcE=np.asarray(bcd["E"]), bcW=np.asarray(bcd["W"]),
            grid_tail=grid_tail, axisymmetric=int(axisymmetric), U=Ub[jj][:, ii], WoW=wo[:NCj], WoE=wo[NCj:2 * NCj],
            WoS=wo[2 * NCj:2 * NCj + NCi], WoN=wo[2 * NCj + NCi:], tail="0\n"))
    return out


def _centroid_area(x, y):
    """quadCentroid (Math.h:311-350): centroid and area of SW, SE, NE, NW."""
    a = xs = ys = 0.0
    for p, q in ((3, 0), (0, 1), (1, 2), (2, 3)):
        ai = x[p] * y[q] - x[q] * y[p]
        a += ai
        xs += (x[q] + x[p]) * ai
        ys += (y[q] + y[p]) * ai
    return xs / (3 * a), ys / (3 * a), 0.5 * a


def write_restart(prefix, mesh, U, nsteps=0, time=0.0, cpu_time=0.0, gas="AIR", axisymmetric=0):
    """<prefix>_blkNNNNNN.soln for every block and, for meshes built by
    mesh.cartesian or mesh.quadtree_box, <prefix>_quadtree.tree."""
    blocks = blocks_from_mesh(mesh, U, nsteps, time, cpu_time, gas, axisymmetric)
    paths = []
    for b, B in enumerate(blocks):
        p = "%s_blk%06d.soln" % (prefix, b)
        write_soln(p, B)
        paths.append(p)
    T = tree_of_mesh(mesh)
    if T is not None:
        write_tree(prefix + "_quadtree.tree", T)
    return paths


def tree_of_mesh(mesh):
    """The quadtree of a mesh from mesh.cartesian (one level, every block a
    root) or mesh.quadtree_box, with the blocks' neighbour infos."""
    info = mesh.info
    n_i, n_j = int(mesh.ni[0]), int(mesh.nj[0])
    lv = mesh.level if mesh.level is not None else np.zeros(mesh.nblk, np.int32)

    def binfo(b):
        return [0, int(b), int(mesh.ni[b]), int(mesh.nj[b]), NG, -1, int(lv[b])]

    def nbinfos(b):
        out = []
        slots = [0, 1, 2, 3, 5, 4, 6, 7]   # N, S, E, W, NW, NE, SE, SW
        for s in slots:
            lst = []
            o = mesh.nbr[b, s]
            if o >= 0:
                lst.append(binfo(o))
                if s < 4 and mesh.nbr2 is not None and mesh.nbr2[b, s] >= 0:
                    lst.append(binfo(mesh.nbr2[b, s]))
            out.append(lst)
        return out

    if info.get("kind") == "cartesian":
        nbi, nbj = info["nbi"], info["nbj"]
        roots = [[TreeBlock(1, jb * nbi + ib, binfo(jb * nbi + ib), nbinfos(jb * nbi + ib))
                  for ib in range(nbi)] for jb in range(nbj)]
        return QuadTreeFile(nbi, nbj, 1, mesh.nblk, roots)
    if info.get("kind") == "quadtree":
        keys = info["keys"]
        bid = {k: b for b, k in enumerate(keys)}
        R = info["nroot"]

        def node(lv_, x, y, sector):
            k = (lv_, x, y)
            if k in bid:
                b = bid[k]
                i_ = binfo(b)
                i_[5] = sector
                return TreeBlock(1, b, i_, nbinfos(b))
            kids = [node(lv_ + 1, 2 * x + (c & 1), 2 * y + (c >> 1), c) for c in range(4)]
            return TreeBlock(0, -1, [-1, -1, n_i, n_j, NG, sector, lv_], [[] for _ in range(8)], kids)

        roots = [[node(0, ib, jb, -1) for ib in range(R)] for jb in range(R)]
        return QuadTreeFile(R, R, 1, mesh.nblk, roots)
    return None


def read_restart(prefix):
    """Blocks (in global block order), the quadtree file, and the Mesh + U
    of a restart written by write_restart or by the reference (a single
    block, or a one-level grid of roots)."""
    T = read_tree(prefix + "_quadtree.tree") if os.path.exists(prefix + "_quadtree.tree") else None
    if T is None:
        blocks = [read_soln(prefix + "_blk000000.soln")]
        m, U = to_mesh(blocks)
        return blocks, T, m, U
    used = sorted(used_blocks(T))
    blocks = [read_soln("%s_blk%06d.soln" % (prefix, g)) for g, _, _, _ in used]
    # topology from the tree: the quadtree rules of mesh.quadtree_box
    from .mesh import _tables_from_leaves
    leaves = [(lv, x, y) for _, lv, x, y in used]
    nbr, nbr2 = _tables_from_leaves(leaves, T.nri, T.nrj)
    level = np.array([lv for _, lv, _, _ in used], np.int32)
    amr = len(set(level.tolist())) > 1
    m, U = to_mesh(blocks, nbr, level if amr else None, nbr2 if amr else None)
    return blocks, T, m, U


def checkpoint(prefix, solver, nsteps=0, time=0.0, cpu_time=0.0, gas="AIR"):
    """Write_Restart_Solution (+ Write_QuadTree) for a running
    euler2d.Euler2DQuadBlocks: its mesh, current conserved state and the
    scheme's Axisymmetric flag."""
    return write_restart(prefix, solver.mesh, solver.get_U(), nsteps, time, cpu_time, gas,
                         axisymmetric=solver.scheme.axisymmetric)


def resume(prefix, scheme=None):
    """Read_Restart_Solution (+ Read_QuadTree): a new Euler2DQuadBlocks on
    the restart's mesh with its state loaded (U, then W = W(U)), and the
    step count and time of the file.  The blocks' Axisymmetric flag
    (Euler2DQuad.h:1347) sets the scheme's when no scheme is given; a given
    scheme that disagrees with it is refused."""
    from . import abi, euler2d
    blocks, _, m, U = read_restart(prefix)
    axi = blocks[0].axisymmetric
    if any(B.axisymmetric != axi for B in blocks):
        raise ValueError("restart blocks disagree on the Axisymmetric flag")
    if scheme is None:
        scheme = abi.Scheme(axisymmetric=axi)
    elif int(scheme.axisymmetric) != axi:
        raise ValueError("scheme axisymmetric=%d but the restart was written with Axisymmetric=%d"
                         % (scheme.axisymmetric, axi))
    g = euler2d.Euler2DQuadBlocks(m, scheme)
    g.set_U(U)
    return g, blocks[0].nsteps, blocks[0].time
