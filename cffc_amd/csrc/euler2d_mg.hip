// euler2d_mg.hip -- the level-transfer kernels of the FAS multigrid solver
// (FASMultigrid2D/FASMultigrid2D.h), gfx950.
//
// Each multigrid level is its own context (padded SoA layout, BlockMeta per
// block); level l+1 halves level l block by block, so coarse cell (ic, jc)
// of block b (padded indices, 2 ghost layers) covers the fine cells
// (2*(ic-2)+2 + {0,1}, 2*(jc-2)+2 + {0,1}) of block b.  One thread per coarse
// interior cell (grid: x over the block's coarse cells, y over blocks), FP64,
// built with -ffp-contract=off in the reference's operation order.  These are
// memory-bound elementwise passes over a level (tens of bytes per cell) next
// to the stage kernels' hundreds of flops per cell.
#include <hip/hip_runtime.h>

#include "euler2d_device.hpp"
#include "euler2d_kernels.hpp"
#include "euler2d_state.hpp"

namespace cffc {

namespace {

__device__ __forceinline__ int64_t cell(const BlockMeta &m, int i, int j) { return m.coff + (int64_t)j * m.pitch + i; }

__device__ __forceinline__ CS<double> ldU(const double *U, int64_t ncp, int64_t c) {
    CS<double> u = {U[c], U[ncp + c], U[2 * ncp + c], U[3 * ncp + c]};
    return u;
}
__device__ __forceinline__ void stU(double *U, int64_t ncp, int64_t c, const CS<double> &u) {
    U[c] = u.d; U[ncp + c] = u.mx; U[2 * ncp + c] = u.my; U[3 * ncp + c] = u.E;
}

// Restrict_Solution_Blocks :996-1037 (src = U, dst = U) and
// Restrict_Residuals :1053-1095 (src = dUdt[0], dst = P): the area-weighted
// average of the four fine cells, SW, SE, NW, NE in that order
__global__ void mg_restrict_kernel(const BlockMeta *mf, const BlockMeta *mc, const double *src, const double *Af,
                                   int64_t ncpf, double *dst, const double *Ac, int64_t ncpc) {
    const BlockMeta f = mf[blockIdx.y], c = mc[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.ni * c.nj || c.skip) return;
    const int ic = q % c.ni + 2, jc = q / c.ni + 2;
    const int i0 = 2 * (ic - 2) + 2, j0 = 2 * (jc - 2) + 2;
    const int64_t q0 = cell(f, i0, j0), q1 = q0 + 1, q2 = q0 + f.pitch, q3 = q2 + 1;
    const double A0 = Af[q0], A1 = Af[q1], A2 = Af[q2], A3 = Af[q3];
    const int64_t cc = cell(c, ic, jc);
    const Recip<double> rA = recip(Ac[cc]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double *s = src + v * ncpf;
        dst[v * ncpc + cc] = rdiv(s[q0] * A0 + s[q1] * A1 + s[q2] * A2 + s[q3] * A3, rA);
    }
}

// Update_Primitive_Variables :2423-2445
__global__ void mg_primitive_kernel(const BlockMeta *meta, const double *U, double *W, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj || m.skip) return;
    const int64_t c = cell(m, q % m.ni + 2, q / m.ni + 2);
    const PS<double> w = W_of_U(ldU(U, ncp, c));
    W[c] = w.d; W[ncp + c] = w.vx; W[2 * ncp + c] = w.vy; W[3 * ncp + c] = w.p;
}

// Subtract_dUdt_from_P :2331-2352
__global__ void mg_subtract_kernel(const BlockMeta *meta, double *P, const double *R, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj || m.skip) return;
    const int64_t c = cell(m, q % m.ni + 2, q / m.ni + 2);
#pragma unroll
    for (int v = 0; v < 4; ++v) P[v * ncp + c] -= R[v * ncp + c];
}

// CFL_Multigrid :2364-2407
__global__ void mg_cfl_kernel(const BlockMeta *mf, const BlockMeta *mc, const double *dtf, const double *Af,
                              double *dtc, const double *Ac) {
    const BlockMeta f = mf[blockIdx.y], c = mc[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.ni * c.nj || c.skip) return;
    const int ic = q % c.ni + 2, jc = q / c.ni + 2;
    const int i0 = 2 * (ic - 2) + 2, j0 = 2 * (jc - 2) + 2;
    const int64_t sw = cell(f, i0, j0), se = sw + 1, nw = sw + f.pitch, ne = nw + 1;
    const int64_t cc = cell(c, ic, jc);
    const double A = Ac[cc];
    const double dt_NE = qsqrt(qdiv(A, Af[ne])) * dtf[ne];
    const double dt_SE = qsqrt(qdiv(A, Af[se])) * dtf[se];
    const double dt_NW = qsqrt(qdiv(A, Af[nw])) * dtf[nw];
    const double dt_SW = qsqrt(qdiv(A, Af[sw])) * dtf[sw];
    dtc[cc] = dmin(dtc[cc], dmin(dt_NE, dmin(dt_SE, dmin(dt_NW, dt_SW))));
}

// Evaluate_Solution_Changes :2454-2478 over every cell of each block,
// ghosts included (U -= uo); a converged case of a batch is skipped
__global__ void mg_changes_kernel(const BlockMeta *meta, double *U, const double *uo, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.nci * m.ncj || m.skip) return;
    const int64_t c = cell(m, q % m.nci, q / m.nci);
#pragma unroll
    for (int v = 0; v < 4; ++v) U[v * ncp + c] -= uo[v * ncp + c];
}

// uo = U over every cell of each block, ghosts included
// (Store_Current_Solution_in_uo_MG); a kernel rather than a copy node, so
// that a captured cycle holds kernels only
__global__ void mg_copy_kernel(const BlockMeta *meta, double *dst, const double *src, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.nci * m.ncj || m.skip) return;
    const int64_t c = cell(m, q % m.nci, q / m.nci);
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[v * ncp + c] = src[v * ncp + c];
}

// Neumann conditions on the coarse solution change before bilinear
// prolongation (:1640-1669): ghost = interior next to constant-extrapolation
// boundaries; one thread per boundary cell of the interior span
__global__ void mg_neumann_kernel(const BlockMeta *meta, const int32_t *bc, double *U, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    if (m.skip) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t *bW = bc + m.bcoff, *bE = bW + m.ncj, *bS = bE + m.ncj, *bN = bS + m.nci;
    if (t < m.nj) {
        const int j = t + 2;
        // E before W, as the reference loop does; the two never share a cell
        if (bE[j] == 4001) stU(U, ncp, cell(m, m.ni + 2, j), ldU(U, ncp, cell(m, m.ni + 1, j)));
        if (bW[j] == 4001) stU(U, ncp, cell(m, 1, j), ldU(U, ncp, cell(m, 2, j)));
    } else if (t < m.nj + m.ni) {
        const int i = t - m.nj + 2;
        if (bS[i] == 4001) stU(U, ncp, cell(m, i, 1), ldU(U, ncp, cell(m, i, 2)));
        if (bN[i] == 4001) stU(U, ncp, cell(m, i, m.nj + 2), ldU(U, ncp, cell(m, i, m.nj + 1)));
    }
}

// Bilinear_Interpolation, CFD.h:1367-1416 (Zingg and Yarrow)
__device__ bool bilinear(const CS<double> &U1, double x1, double y1, const CS<double> &U2, double x2, double y2,
                         const CS<double> &U3, double x3, double y3, const CS<double> &U4, double x4, double y4,
                         double x0, double y0, CS<double> &U0) {
    const double TOL = 0.0000001;
    const double a0 = x1, a1 = x4 - x1, a2 = x2 - x1, a3 = x1 + x3 - x2 - x4;
    const double b0 = y1, b1 = y4 - y1, b2 = y2 - y1, b3 = y1 + y3 - y2 - y4;
    const double a = a1 * b3 - a3 * b1;
    const double b = a0 * b3 - b3 * x0 + a1 * b2 - a2 * b1 + a3 * y0 - a3 * b0;
    const double c = a0 * b2 - b2 * x0 + a2 * y0 - a2 * b0;
    double zeta0;
    if (a > TOL || a < -TOL) {
        if (b * b - 4.0 * a * c < 0.0) return false;
        zeta0 = (-b + sqrt(b * b - 4.0 * a * c)) / (2.0 * a);
        if (zeta0 < 0.0 || zeta0 > 1.0) {
            zeta0 = (-b - sqrt(b * b - 4.0 * a * c)) / (2.0 * a);
            if (zeta0 < 0.0 || zeta0 > 1.0) return false;
        }
    } else if (b > TOL || b < -TOL) {
        zeta0 = -c / b;
        if (zeta0 < 0.0 || zeta0 > 1.0) return false;
    } else {
        return false;
    }
    if (b2 + b3 * zeta0 < TOL && b2 + b3 * zeta0 > -TOL) return false;
    const double eta0 = (y0 - b0 - b1 * zeta0) / (b2 + b3 * zeta0);
    U0.d = U1.d + (U4.d - U1.d) * zeta0 + (U2.d - U1.d) * eta0 + (U1.d + U3.d - U2.d - U4.d) * zeta0 * eta0;
    U0.mx = U1.mx + (U4.mx - U1.mx) * zeta0 + (U2.mx - U1.mx) * eta0 + (U1.mx + U3.mx - U2.mx - U4.mx) * zeta0 * eta0;
    U0.my = U1.my + (U4.my - U1.my) * zeta0 + (U2.my - U1.my) * eta0 + (U1.my + U3.my - U2.my - U4.my) * zeta0 * eta0;
    U0.E = U1.E + (U4.E - U1.E) * zeta0 + (U2.E - U1.E) * eta0 + (U1.E + U3.E - U2.E - U4.E) * zeta0 * eta0;
    return true;
}

__device__ __forceinline__ bool unphysical(const CS<double> &U) { return U.d <= 0.0 || U.E <= 0.0 || U_e(U) <= 0.0; }

// interpolation quadrilaterals per fine cell (SW, SE, NE, NW), by the offset
// of the quadrilateral's SW coarse cell (:1718-2211)
__constant__ int8_t kQuad[4][4][2] = {
    {{-1, -1}, {-1, 0}, {0, -1}, {0, 0}},
    {{0, -1}, {0, 0}, {-1, -1}, {-1, 0}},
    {{0, 0}, {0, -1}, {-1, 0}, {-1, -1}},
    {{-1, 0}, {0, 0}, {-1, -1}, {0, -1}},
};

// Prolong_and_Update_Solution_Blocks :1627-2213.  Uc: the coarse solution
// change (ghosts included).  A coarse cell next to a fixed boundary injects
// into its four fine cells; otherwise each fine cell interpolates from the
// first quadrilateral that contains it, and a cell none contains restarts
// the coarse cell with injection from the SW fine cell (cells already
// updated are updated again, as in the reference).  Updates that leave the
// state unphysical are halved up to max_red times, then dropped.
//
// assign (Prolong_Solution_Blocks :1200-1625, full multigrid): Uc is the
// coarse solution; the fine cells take it (same quadrilateral order and
// injection fall-back), with neither Dirichlet injection nor reduction.
__global__ void mg_prolong_kernel(const BlockMeta *mc, const BlockMeta *mf, const int32_t *bc, const double *Uc,
                                  const double *gc, int64_t ncpc, double *Uf, const double *gf, int64_t ncpf,
                                  int inj_dirichlet, int inj_always, int stability, int max_red, int assign) {
    const BlockMeta c = mc[blockIdx.y], f = mf[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.ni * c.nj || c.skip) return;
    const int ic = q % c.ni + 2, jc = q / c.ni + 2;
    const int i0 = 2 * (ic - 2) + 2, j0 = 2 * (jc - 2) + 2;
    const int32_t *bW = bc + c.bcoff, *bE = bW + c.ncj, *bS = bE + c.ncj, *bN = bS + c.nci;
    const CS<double> Ucc = ldU(Uc, ncpc, cell(c, ic, jc));
    const int fdi[4] = {0, 1, 1, 0}, fdj[4] = {0, 0, 1, 1};
    const bool dirichlet = (ic == 2 && bW[jc] == 11000) || (ic == c.ni + 1 && bE[jc] == 11000) ||
                           (jc == 2 && bS[ic] == 11000) || (jc == c.nj + 1 && bN[ic] == 11000);
    if (!assign && inj_dirichlet && dirichlet) {
        for (int k = 0; k < 4; ++k) {
            const int64_t fc = cell(f, i0 + fdi[k], j0 + fdj[k]);
            CS<double> u = ldU(Uf, ncpf, fc);
            u.d += Ucc.d; u.mx += Ucc.mx; u.my += Ucc.my; u.E += Ucc.E;
            stU(Uf, ncpf, fc, u);
        }
        return;
    }
    const double *xc = gc + GEO_XC * ncpc, *yc = gc + GEO_YC * ncpc;
    bool inject = inj_always != 0;
    for (int k = 0; k < 4; ++k) {
        const int64_t fc = cell(f, i0 + fdi[k], j0 + fdj[k]);
        CS<double> dU;
        if (!inject) {
            bool ok = false;
            for (int t = 0; t < 4 && !ok; ++t) {
                const int a = ic + kQuad[k][t][0], e = jc + kQuad[k][t][1];
                const int64_t p1 = cell(c, a, e), p2 = p1 + c.pitch, p3 = p2 + 1, p4 = p1 + 1;
                ok = bilinear(ldU(Uc, ncpc, p1), xc[p1], yc[p1], ldU(Uc, ncpc, p2), xc[p2], yc[p2], ldU(Uc, ncpc, p3),
                              xc[p3], yc[p3], ldU(Uc, ncpc, p4), xc[p4], yc[p4], gf[GEO_XC * ncpf + fc],
                              gf[GEO_YC * ncpf + fc], dU);
            }
            if (!ok) {
                inject = true;
                k = -1;   // restart the coarse cell with injection
                continue;
            }
        } else {
            dU = Ucc;
        }
        if (assign) {
            stU(Uf, ncpf, fc, dU);
            continue;
        }
        CS<double> u = ldU(Uf, ncpf, fc);
        if (stability) {
            CS<double> un = {u.d + dU.d, u.mx + dU.mx, u.my + dU.my, u.E + dU.E};
            if (unphysical(un)) {
                int n = 1;
                while (n <= max_red) {
                    dU.d *= 0.5; dU.mx *= 0.5; dU.my *= 0.5; dU.E *= 0.5;
                    CS<double> um = {u.d + dU.d, u.mx + dU.mx, u.my + dU.my, u.E + dU.E};
                    if (!unphysical(um)) break;
                    n++;
                }
                if (n > max_red) { dU.d = 0.0; dU.mx = 0.0; dU.my = 0.0; dU.E = 0.0; }
            }
        }
        u.d += dU.d; u.mx += dU.mx; u.my += dU.my; u.E += dU.E;
        stU(Uf, ncpf, fc, u);
    }
}

}  // namespace
}  // namespace cffc

using namespace cffc;

extern "C" hipError_t cffc_launch_mg_restrict(const BlockMeta *mf, const BlockMeta *mc, int nblk, int max_int_c,
                                   const double *src, const double *Af, int64_t ncpf, double *dst, const double *Ac,
                                   int64_t ncpc, hipStream_t st) {
    dim3 grid((max_int_c + 255) / 256, nblk);
    hipLaunchKernelGGL(mg_restrict_kernel, grid, dim3(256), 0, st, mf, mc, src, Af, ncpf, dst, Ac, ncpc);
    return hipGetLastError();
}
extern "C" hipError_t cffc_launch_mg_primitive(const BlockMeta *meta, int nblk, int max_int, const double *U, double *W,
                                    int64_t ncp, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(mg_primitive_kernel, grid, dim3(256), 0, st, meta, U, W, ncp);
    return hipGetLastError();
}
extern "C" hipError_t cffc_launch_mg_subtract(const BlockMeta *meta, int nblk, int max_int, double *P, const double *R,
                                   int64_t ncp, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(mg_subtract_kernel, grid, dim3(256), 0, st, meta, P, R, ncp);
    return hipGetLastError();
}
extern "C" hipError_t cffc_launch_mg_cfl(const BlockMeta *mf, const BlockMeta *mc, int nblk, int max_int_c, const double *dtf,
                              const double *Af, double *dtc, const double *Ac, hipStream_t st) {
    dim3 grid((max_int_c + 255) / 256, nblk);
    hipLaunchKernelGGL(mg_cfl_kernel, grid, dim3(256), 0, st, mf, mc, dtf, Af, dtc, Ac);
    return hipGetLastError();
}
extern "C" hipError_t cffc_launch_mg_changes(const BlockMeta *meta, int nblk, int max_cells, double *U,
                                             const double *uo, int64_t ncp, hipStream_t st) {
    hipLaunchKernelGGL(mg_changes_kernel, dim3((max_cells + 255) / 256, nblk), dim3(256), 0, st, meta, U, uo, ncp);
    return hipGetLastError();
}
extern "C" hipError_t cffc_launch_mg_copy(const BlockMeta *meta, int nblk, int max_cells, double *dst,
                                          const double *src, int64_t ncp, hipStream_t st) {
    hipLaunchKernelGGL(mg_copy_kernel, dim3((max_cells + 255) / 256, nblk), dim3(256), 0, st, meta, dst, src, ncp);
    return hipGetLastError();
}
extern "C" hipError_t cffc_launch_mg_prolong(const BlockMeta *mc, const BlockMeta *mf, int nblk, int max_int_c, int max_edge_c,
                                  const int32_t *bc, double *Uc, const double *gc, int64_t ncpc, double *Uf,
                                  const double *gf, int64_t ncpf, int neumann, int inj_dirichlet, int inj_always,
                                  int stability, int max_red, int assign, hipStream_t st) {
    if (neumann)
        hipLaunchKernelGGL(mg_neumann_kernel, dim3((max_edge_c + 255) / 256, nblk), dim3(256), 0, st, mc, bc, Uc, ncpc);
    hipLaunchKernelGGL(mg_prolong_kernel, dim3((max_int_c + 255) / 256, nblk), dim3(256), 0, st, mc, mf, bc, Uc, gc,
                       ncpc, Uf, gf, ncpf, inj_dirichlet, inj_always, stability, max_red, assign);
    return hipGetLastError();
}
