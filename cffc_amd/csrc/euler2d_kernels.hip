// euler2d_kernels.hip -- FP64 gfx950 kernels of the 2D Euler explicit path.
//
// Built with -ffp-contract=off so that every FP64 operation rounds exactly as
// in the reference's strict-IEEE CPU build.  Contents:
//   geometry_kernel    cell centroid/area and face geometry from the nodes
//                      (HO_Grid2DQuad.h:1403-1410, 1782-1980; Math.h:311-350)
//   exchange_bc_kernel Send_All_Messages (same-level faces and corners,
//                      AdaptiveBlock2D_MessagePassing.h:32-1183, 5680-6282,
//                      7883-8020) fused with BCs (Euler2DQuadSingleBlock.cc:
//                      2354-3010), one workgroup per block
//   cfl_kernel         CFL + global minimum (SingleBlock.cc:3019-3080)
//   update_kernel      Update_Solution_Multistage_Explicit (SingleBlock.cc:5300)
//   norm kernels       L1/L2/Max_Norm_Residual (SingleBlock.cc:3111-3167)
//   stage kernels      the fused residual kernel (euler2d_stage.hpp), FP64
#include <hip/hip_runtime.h>

#include "euler2d_device.hpp"
#include "euler2d_kernels.hpp"
#include "euler2d_stage.hpp"
#include "euler2d_state.hpp"

namespace cffc {

// ---------------------------------------------------------------------------
// geometry: one thread per padded cell, all blocks (blockIdx.y = block)
__global__ void geometry_kernel(const BlockMeta *meta, const double *xn, const double *yn, double *geom,
                                int64_t ncp, unsigned int *ndegenerate) {
    const BlockMeta m = meta[blockIdx.y];
    const int ncell = m.nci * m.ncj;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const int i = q % m.nci, j = q / m.nci;
    const int nn = m.nci + 1;
    auto X = [&](int a, int b) { return xn[m.noff + (int64_t)b * nn + a]; };
    auto Y = [&](int a, int b) { return yn[m.noff + (int64_t)b * nn + a]; };
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    // quadCentroid with SW, SE, NE, NW
    const double X0x = X(i, j), X0y = Y(i, j), X1x = X(i + 1, j), X1y = Y(i + 1, j);
    const double X2x = X(i + 1, j + 1), X2y = Y(i + 1, j + 1), X3x = X(i, j + 1), X3y = Y(i, j + 1);
    double ai, atmp = 0, xtmp = 0, ytmp = 0;
    ai = X3x * X0y - X0x * X3y; atmp += ai; xtmp += (X0x + X3x) * ai; ytmp += (X0y + X3y) * ai;
    ai = X0x * X1y - X1x * X0y; atmp += ai; xtmp += (X1x + X0x) * ai; ytmp += (X1y + X0y) * ai;
    ai = X1x * X2y - X2x * X1y; atmp += ai; xtmp += (X2x + X1x) * ai; ytmp += (X2y + X1y) * ai;
    ai = X2x * X3y - X3x * X2y; atmp += ai; xtmp += (X3x + X2x) * ai; ytmp += (X3y + X2y) * ai;
    geom[GEO_A * ncp + c] = 0.5 * atmp;
    geom[GEO_XC * ncp + c] = atmp != 0 ? xtmp / (3 * atmp) : 0.0;
    geom[GEO_YC * ncp + c] = atmp != 0 ? ytmp / (3 * atmp) : 0.0;
    // face lengths: E = |N(i+1,j)-N(i+1,j+1)|, W = |N(i,j)-N(i,j+1)|,
    //               N = |N(i,j+1)-N(i+1,j+1)|, S = |N(i,j)-N(i+1,j)|
    const double lE = sqrt((X1x - X2x) * (X1x - X2x) + (X1y - X2y) * (X1y - X2y));
    const double lW = sqrt((X0x - X3x) * (X0x - X3x) + (X0y - X3y) * (X0y - X3y));
    const double lN = sqrt((X3x - X2x) * (X3x - X2x) + (X3y - X2y) * (X3y - X2y));
    const double lS = sqrt((X0x - X1x) * (X0x - X1x) + (X0y - X1y) * (X0y - X1y));
    const double perim = lE + lN + lW + lS;
    geom[GEO_LE * ncp + c] = lE;
    geom[GEO_LN * ncp + c] = lN;
    geom[GEO_XFE * ncp + c] = 0.5 * (X1x + X2x);
    geom[GEO_YFE * ncp + c] = 0.5 * (X1y + X2y);
    geom[GEO_XFN * ncp + c] = 0.5 * (X3x + X2x);
    geom[GEO_YFN * ncp + c] = 0.5 * (X3y + X2y);
    // nfaceE = (dy, -dx)/|d| with d = N(i+1,j+1) - N(i+1,j)
    if (lE > 0.000000001 * perim) {
        const double dx = X2x - X1x, dy = X2y - X1y, l = sqrt(dx * dx + dy * dy);
        geom[GEO_NEX * ncp + c] = dy / l;
        geom[GEO_NEY * ncp + c] = -dx / l;
    } else {
        geom[GEO_NEX * ncp + c] = 1.0;
        geom[GEO_NEY * ncp + c] = 0.0;
        atomicAdd(ndegenerate, 1u);
    }
    // nfaceN = (a.y-b.y, -(a.x-b.x))/|a-b| with a = N(i,j+1), b = N(i+1,j+1)
    if (lN > 0.000000001 * perim) {
        const double dx = X3x - X2x, dy = X3y - X2y, l = sqrt(dx * dx + dy * dy);
        geom[GEO_NNX * ncp + c] = dy / l;
        geom[GEO_NNY * ncp + c] = -dx / l;
    } else {
        geom[GEO_NNX * ncp + c] = 1.0;
        geom[GEO_NNY * ncp + c] = 0.0;
        atomicAdd(ndegenerate, 1u);
    }
    // a W or S face that is degenerate seen from this cell but not from its
    // neighbour would break nfaceW(i) = -nfaceE(i-1); count it too
    if (!(lW > 0.000000001 * perim) || !(lS > 0.000000001 * perim)) atomicAdd(ndegenerate, 1u);
}

// FP32 geometry of the mixed path: one thread per padded cell.  Positions are
// taken relative to the centroid of the block's first interior cell, so the
// FP32 differences the reconstruction forms keep their precision.
__global__ void geometry_f32_kernel(const BlockMeta *meta, const double *geom, float *gf, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.nci * m.ncj) return;
    const int i = q % m.nci, j = q / m.nci;
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    const int64_t c0 = m.coff + 2 * (int64_t)m.pitch + 2;
    const double x0 = geom[GEO_XC * ncp + c0], y0 = geom[GEO_YC * ncp + c0];
    gf[GF_XC * ncp + c] = (float)(geom[GEO_XC * ncp + c] - x0);
    gf[GF_YC * ncp + c] = (float)(geom[GEO_YC * ncp + c] - y0);
    gf[GF_XFE * ncp + c] = (float)(geom[GEO_XFE * ncp + c] - x0);
    gf[GF_YFE * ncp + c] = (float)(geom[GEO_YFE * ncp + c] - y0);
    gf[GF_XFN * ncp + c] = (float)(geom[GEO_XFN * ncp + c] - x0);
    gf[GF_YFN * ncp + c] = (float)(geom[GEO_YFN * ncp + c] - y0);
    gf[GF_NEX * ncp + c] = (float)geom[GEO_NEX * ncp + c];
    gf[GF_NEY * ncp + c] = (float)geom[GEO_NEY * ncp + c];
    gf[GF_NNX * ncp + c] = (float)geom[GEO_NNX * ncp + c];
    gf[GF_NNY * ncp + c] = (float)geom[GEO_NNY * ncp + c];
    gf[GF_LE * ncp + c] = (float)geom[GEO_LE * ncp + c];
    gf[GF_LN * ncp + c] = (float)geom[GEO_LN * ncp + c];
}

// ---------------------------------------------------------------------------
// state scatter / gather between host AoS interior order and device SoA padded
__global__ void scatter_state_kernel(const BlockMeta *meta, const double *aos, double *dst, int64_t ncp, int kind,
                                     double *Ucons) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj) return;
    const int i = q % m.ni + 2, j = q / m.ni + 2;
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    const double *a = aos + 4 * (m.ioff + q);
    if (kind == 0) {  // W given: W, and U = U(W)
        PS<double> W = {a[0], a[1], a[2], a[3]};
        CS<double> U = U_of_W(W);
        dst[c] = W.d; dst[ncp + c] = W.vx; dst[2 * ncp + c] = W.vy; dst[3 * ncp + c] = W.p;
        Ucons[c] = U.d; Ucons[ncp + c] = U.mx; Ucons[2 * ncp + c] = U.my; Ucons[3 * ncp + c] = U.E;
    } else if (kind == 1) {  // U given: U, and W = W(U)
        CS<double> U = {a[0], a[1], a[2], a[3]};
        PS<double> W = W_of_U(U);
        dst[c] = W.d; dst[ncp + c] = W.vx; dst[2 * ncp + c] = W.vy; dst[3 * ncp + c] = W.p;
        Ucons[c] = U.d; Ucons[ncp + c] = U.mx; Ucons[2 * ncp + c] = U.my; Ucons[3 * ncp + c] = U.E;
    }
}

__global__ void gather_state_kernel(const BlockMeta *meta, const double *src, double *aos, int64_t ncp,
                                    int nvar) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj) return;
    const int i = q % m.ni + 2, j = q / m.ni + 2;
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    double *a = aos + (int64_t)nvar * (m.ioff + q);
    for (int v = 0; v < nvar; ++v) a[v] = src[v * ncp + c];
}

// ---------------------------------------------------------------------------
// ghost exchange + physical BCs.  One workgroup per block.
// W is stored as ST (FP64, or FP32 for the fused mixed path); the BC states
// are always formed in FP64.
template <typename ST> __device__ __forceinline__ PS<double> ldW(const ST *W, int64_t ncp, int64_t c) {
    PS<double> r = {(double)W[c], (double)W[ncp + c], (double)W[2 * ncp + c], (double)W[3 * ncp + c]};
    return r;
}
template <typename ST> __device__ __forceinline__ void stW(ST *W, int64_t ncp, int64_t c, const PS<double> &r) {
    W[c] = (ST)r.d; W[ncp + c] = (ST)r.vx; W[2 * ncp + c] = (ST)r.vy; W[3 * ncp + c] = (ST)r.p;
}

template <typename ST>
__global__ __launch_bounds__(256) void exchange_bc_kernel(ExchangeParams p) {
    if (p.stop_flag && *p.stop_flag) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && p.reset_slot) *p.reset_slot = 0x412E848000000000ull;  // MILLION
    const int b = p.blocks ? p.blocks[blockIdx.x] : (int)blockIdx.x;
    const BlockMeta m = p.meta[b];
    if (m.skip) return;   // uniform over the workgroup
    const int64_t ncp = p.ncp;
    ST *W;
    if constexpr (sizeof(ST) == 4) W = p.Wf;
    else W = p.W;
    const int ICl = 2, ICu = m.ni + 1, JCl = 2, JCu = m.nj + 1;
    const int tid = threadIdx.x, NT = blockDim.x;
    // phase 0: same-level neighbours, faces (2 layers over the interior span)
    // and corners (2 x 2); U is transferred and W = W(U) recomputed on stage 1
    // (Euler2DQuad.h:1643-1647); later stages copy W, which equals W(U).
    // The eight neighbours' tables are read once into LDS and the ghost
    // cells of all eight segments form one index space, so that every
    // thread's loads are independent of the other segments' stores.
    __shared__ int32_t sNb[8];
    __shared__ BlockMeta sMeta[8];
    if (!p.bcs_only) {
        if (tid < 8) {
            const int nb = p.nbr[8 * b + tid];
            sNb[tid] = nb;
            if (nb >= 0) sMeta[tid] = p.meta[nb];
        }
        __syncthreads();
        int start[9];
        start[0] = 0;
#pragma unroll
        for (int s = 0; s < 8; ++s)
            start[s + 1] = start[s] + (sNb[s] < 0 ? 0 : s < 2 ? 2 * m.ni : s < 4 ? 2 * m.nj : 4);
        for (int k = tid; k < start[8]; k += NT) {
            int s = 0;
            while (k >= start[s + 1]) ++s;
            const int q = k - start[s];
            const BlockMeta &S = sMeta[s];
            int i0, j0, w, di, dj;
            switch (s) {
            case 0: i0 = ICl; j0 = JCu + 1; w = m.ni; di = 2 - ICl; dj = 2 - (JCu + 1); break;
            case 1: i0 = ICl; j0 = JCl - 2; w = m.ni; di = 2 - ICl; dj = (S.nj + 1) - (JCl - 1); break;
            case 2: i0 = ICu + 1; j0 = JCl; w = 2; di = 2 - (ICu + 1); dj = 2 - JCl; break;
            case 3: i0 = ICl - 2; j0 = JCl; w = 2; di = (S.ni + 1) - (ICl - 1); dj = 2 - JCl; break;
            case 4: i0 = ICu + 1; j0 = JCu + 1; w = 2; di = 2 - (ICu + 1); dj = 2 - (JCu + 1); break;
            case 5: i0 = ICl - 2; j0 = JCu + 1; w = 2; di = (S.ni + 1) - (ICl - 1); dj = 2 - (JCu + 1); break;
            case 6: i0 = ICu + 1; j0 = JCl - 2; w = 2; di = 2 - (ICu + 1); dj = (S.nj + 1) - (JCl - 1); break;
            default: i0 = ICl - 2; j0 = JCl - 2; w = 2; di = (S.ni + 1) - (ICl - 1); dj = (S.nj + 1) - (JCl - 1); break;
            }
            const int i = i0 + q % w, j = j0 + q / w;
            const int64_t cd = m.coff + (int64_t)j * m.pitch + i;
            const int64_t cs = S.coff + (int64_t)(j + dj) * S.pitch + (i + di);
            PS<double> Wg;
            if (p.from_U) {
                const CS<double> U = {p.U[cs], p.U[ncp + cs], p.U[2 * ncp + cs], p.U[3 * ncp + cs]};
                Wg = W_of_U(U);
                if (p.store_U) {
                    p.U[cd] = U.d; p.U[ncp + cd] = U.mx; p.U[2 * ncp + cd] = U.my; p.U[3 * ncp + cd] = U.E;
                }
            } else {
                Wg = ldW(W, ncp, cs);
            }
            stW(W, ncp, cd, Wg);
        }
    }
    if (p.no_bcs) return;   // uniform over the workgroup
    __syncthreads();
    // With double-buffered W (the fused path) the next stage's buffer also
    // receives the BC ghost states: the coarse-to-fine messages of the next
    // stage run before its BCs and read these ghosts, which in the
    // reference hold the states of the previous BCs call
    // (Euler2DQuadSolvers.cc:807-826).
    ST *W2;
    if constexpr (sizeof(ST) == 4) W2 = p.Wf2;
    else W2 = p.W2;
    auto stWb = [&](int64_t c, const PS<double> &r) {
        stW(W, ncp, c, r);
        if (W2) stW(W2, ncp, c, r);
    };
    const int32_t *bcW = p.bc + m.bcoff, *bcE = bcW + m.ncj, *bcS = bcE + m.ncj, *bcN = bcS + m.nci;
    const double *woW = p.wo + 4 * m.bcoff, *woE = woW + 4 * m.ncj, *woS = woE + 4 * m.ncj, *woN = woS + 4 * m.nci;
    auto cidx = [&](int i, int j) { return m.coff + (int64_t)j * m.pitch + i; };
    auto wo = [&](const double *a, int k) { PS<double> r = {a[4 * k], a[4 * k + 1], a[4 * k + 2], a[4 * k + 3]}; return r; };
    // ghost U next to the ghost W (store_U): U(W) of a state the BC sets, or
    // the source cell's U where the BC copies a cell (the reference's BCs
    // keep U and W of every ghost cell)
    auto putU = [&](int64_t c, const PS<double> &w) {
        if (!p.store_U) return;
        const CS<double> u = U_of_W(w);
        p.U[c] = u.d; p.U[ncp + c] = u.mx; p.U[2 * ncp + c] = u.my; p.U[3 * ncp + c] = u.E;
    };
    auto cpyU = [&](int64_t cd, int64_t cs) {
        if (!p.store_U) return;
        for (int v = 0; v < 4; ++v) p.U[v * ncp + cd] = p.U[v * ncp + cs];
    };
    // phase 1: West and East boundaries, j over all rows (corner rows when the
    // South/North BC of the ghost column is BC_NONE), SingleBlock.cc:2362-2676
    for (int q = tid; q < 2 * m.ncj; q += NT) {
        const int j = q >> 1;
        const bool west = (q & 1) == 0;
        const int bcv = west ? bcW[j] : bcE[j];
        if (bcv == 0) continue;
        const int icol = west ? ICl - 1 : ICu + 1;
        const bool apply = (j >= JCl && j <= JCu) || (j < JCl && bcS[icol] == 0) || (j > JCu && bcN[icol] == 0);
        if (!apply) continue;
        const int in = west ? ICl : ICu;       // adjacent interior column
        const int sg = west ? -1 : 1;          // ghost direction
        // normal of the boundary face: nfaceW(ICl) = -nfaceE(ICl-1), nfaceE(ICu)
        const int64_t cf = west ? cidx(ICl - 1, j) : cidx(ICu, j);
        const double nx = west ? -p.geom[GEO_NEX * ncp + cf] : p.geom[GEO_NEX * ncp + cf];
        const double ny = west ? -p.geom[GEO_NEY * ncp + cf] : p.geom[GEO_NEY * ncp + cf];
        const int64_t g1c = cidx(in + sg, j), g2c = cidx(in + 2 * sg, j);
        switch (bcv) {
        case 11000: {  // FIXED
            const PS<double> w0 = wo(west ? woW : woE, j);
            stWb(g1c, w0); stWb(g2c, w0);
            putU(g1c, w0); putU(g2c, w0);
        } break;
        case 7001: {  // REFLECTION: ghost g mirrors cell in-(g-1)*sg
            const PS<double> r1 = reflect(ldW(W, ncp, cidx(in, j)), nx, ny);
            const PS<double> r2 = reflect(ldW(W, ncp, cidx(in - sg, j)), nx, ny);
            stWb(g1c, r1); stWb(g2c, r2);
            putU(g1c, r1); putU(g2c, r2);
        } break;
        case 6000: {  // PERIODIC: W ghosts take ICu-g+1 (E ghosts take ICl+g-1)
            const int64_t s1 = cidx(west ? ICu : ICl, j), s2 = cidx(west ? ICu - 1 : ICl + 1, j);
            stWb(g1c, ldW(W, ncp, s1));
            stWb(g2c, ldW(W, ncp, s2));
            cpyU(g1c, s1); cpyU(g2c, s2);
        } break;
        case 8001: {  // CHARACTERISTIC: first layer from BC_Characteristic_Pressure, second copied
            const PS<double> g1 = bc_characteristic_pressure(ldW(W, ncp, cidx(in, j)), wo(west ? woW : woE, j), nx, ny);
            stWb(g1c, g1); stWb(g2c, g1);
            putU(g1c, g1); cpyU(g2c, g1c);
        } break;
        default: {  // CONSTANT_EXTRAPOLATION (and the reference's default)
            const PS<double> w0 = ldW(W, ncp, cidx(in, j));
            stWb(g1c, w0); stWb(g2c, w0);
            cpyU(g1c, cidx(in, j)); cpyU(g2c, cidx(in, j));
        } break;
        }
    }
    __syncthreads();
    // phase 2: South and North boundaries, i over all columns, :2678-2989
    for (int q = tid; q < 2 * m.nci; q += NT) {
        const int i = q >> 1;
        const bool south = (q & 1) == 0;
        const int bcv = south ? bcS[i] : bcN[i];
        if (bcv == 0) continue;
        const int in = south ? JCl : JCu;
        const int sg = south ? -1 : 1;
        const int64_t cf = south ? cidx(i, JCl - 1) : cidx(i, JCu);
        const double nx = south ? -p.geom[GEO_NNX * ncp + cf] : p.geom[GEO_NNX * ncp + cf];
        const double ny = south ? -p.geom[GEO_NNY * ncp + cf] : p.geom[GEO_NNY * ncp + cf];
        const int64_t g1c = cidx(i, in + sg), g2c = cidx(i, in + 2 * sg);
        switch (bcv) {
        case 11000: {
            const PS<double> w0 = wo(south ? woS : woN, i);
            stWb(g1c, w0); stWb(g2c, w0);
            putU(g1c, w0); putU(g2c, w0);
        } break;
        case 7001: {
            const PS<double> r1 = reflect(ldW(W, ncp, cidx(i, in)), nx, ny);
            const PS<double> r2 = reflect(ldW(W, ncp, cidx(i, in - sg)), nx, ny);
            stWb(g1c, r1); stWb(g2c, r2);
            putU(g1c, r1); putU(g2c, r2);
        } break;
        case 6000: {
            const int64_t s1 = cidx(i, south ? JCu : JCl), s2 = cidx(i, south ? JCu - 1 : JCl + 1);
            stWb(g1c, ldW(W, ncp, s1));
            stWb(g2c, ldW(W, ncp, s2));
            cpyU(g1c, s1); cpyU(g2c, s2);
        } break;
        case 8001: {
            const PS<double> g1 = bc_characteristic_pressure(ldW(W, ncp, cidx(i, in)), wo(south ? woS : woN, i), nx, ny);
            stWb(g1c, g1); stWb(g2c, g1);
            putU(g1c, g1); cpyU(g2c, g1c);
        } break;
        default: {
            const PS<double> w0 = ldW(W, ncp, cidx(i, in));
            stWb(g1c, w0); stWb(g2c, w0);
            cpyU(g1c, cidx(i, in)); cpyU(g2c, cidx(i, in));
        } break;
        }
    }
}

// ---------------------------------------------------------------------------
// CFL over the interior cells of each tile; global minimum with atomicMin on
// the bit pattern (positive doubles order like their bits), unless dtmin is
// null (local time steps only).
__global__ __launch_bounds__(256) void cfl_kernel(const BlockMeta *meta, const TileDesc *tiles, int ti, int tj,
                                                  const double *W, const double *geom, int64_t ncp, double *dt,
                                                  unsigned long long *dtmin) {
    __shared__ double sm[4];
    const TileDesc t = tiles[blockIdx.x];
    const BlockMeta m = meta[t.b];
    if (m.skip) return;   // uniform over the workgroup
    double dmn = MILLION;
    for (int q = threadIdx.x; q < ti * tj; q += blockDim.x) {
        const int ip = t.i0 + 2 + q % ti, jp = t.j0 + 2 + q / ti;
        if (ip > m.ni + 1 || jp > m.nj + 1) continue;
        const int64_t c = m.coff + (int64_t)jp * m.pitch + ip;
        const double *G = geom;
        const double A = G[GEO_A * ncp + c];
        const double lW = G[GEO_LE * ncp + c - 1], lE = G[GEO_LE * ncp + c];
        const double lS = G[GEO_LN * ncp + c - m.pitch], lN = G[GEO_LN * ncp + c];
        const double d_i = 2.0 * qdiv(A, lE + lW);
        const double d_j = 2.0 * qdiv(A, lN + lS);
        const PS<double> Wc = ldW(W, ncp, c);
        const double nWx = -G[GEO_NEX * ncp + c - 1], nWy = -G[GEO_NEY * ncp + c - 1];
        const double nSx = -G[GEO_NNX * ncp + c - m.pitch], nSy = -G[GEO_NNY * ncp + c - m.pitch];
        const double v_i = 0.5 * (Wc.vx * (G[GEO_NEX * ncp + c] - nWx) + Wc.vy * (G[GEO_NEY * ncp + c] - nWy));
        const double v_j = 0.5 * (Wc.vx * (G[GEO_NNX * ncp + c] - nSx) + Wc.vy * (G[GEO_NNY * ncp + c] - nSy));
        const double a = W_a(Wc);
        const double d = dmin(qdiv(d_i, a + fabs(v_i)), qdiv(d_j, a + fabs(v_j)));
        dt[c] = d;
        dmn = dmin(dmn, d);
    }
    for (int off = 32; off > 0; off >>= 1) dmn = dmin(dmn, __shfl_xor(dmn, off));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = dmn;
    __syncthreads();
    if (threadIdx.x == 0 && dtmin) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) dmn = dmin(dmn, sm[w]);
        dmn = dmin(sm[0], dmn);
        atomicMin(dtmin, dbits(dmn));
    }
}

// Set_Global_TimeStep over the interior cells
__global__ void set_global_dt_kernel(const BlockMeta *meta, double *dt, const unsigned long long *dtmin) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj) return;
    const int i = q % m.ni + 2, j = q / m.ni + 2;
    dt[m.coff + (int64_t)j * m.pitch + i] = bitsd(*dtmin);
}

// ---------------------------------------------------------------------------
// Update_Solution_Multistage_Explicit on its own (granular path), in place.
__global__ void update_kernel(const BlockMeta *meta, double *W, double *U, const double *Uo, double *dU_k,
                              double *dt, int64_t ncp, double omega, int lts, unsigned long long *err) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj || m.skip) return;
    const int i = q % m.ni + 2, j = q / m.ni + 2;
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    double d[4] = {dU_k[c], dU_k[ncp + c], dU_k[2 * ncp + c], dU_k[3 * ncp + c]};
    const double U0[4] = {Uo[c], Uo[ncp + c], Uo[2 * ncp + c], Uo[3 * ncp + c]};
    CS<double> Un = {U0[0] + omega * d[0], U0[1] + omega * d[1], U0[2] + omega * d[2], U0[3] + omega * d[3]};
    bool bad = (Un.d <= 0.0 || Un.E <= 0.0 || U_e(Un) <= 0.0);
    if (bad && lts) {
        double f = 1.0;
        double dtl = dt[c];
        for (int n = 1; n <= 10; ++n) {
            f = 0.5 * f;
            dtl = f * dtl;
            for (int v = 0; v < 4; ++v) d[v] = f * d[v];
            Un.d = U0[0] + omega * d[0]; Un.mx = U0[1] + omega * d[1];
            Un.my = U0[2] + omega * d[2]; Un.E = U0[3] + omega * d[3];
            if (Un.d > 0.0 && Un.E > 0.0 && U_e(Un) > 0.0) break;
        }
        dt[c] = dtl;
        for (int v = 0; v < 4; ++v) dU_k[v * ncp + c] = d[v];
        bad = (Un.d <= 0.0 || Un.E <= 0.0 || U_e(Un) <= 0.0);
    }
    U[c] = Un.d; U[ncp + c] = Un.mx; U[2 * ncp + c] = Un.my; U[3 * ncp + c] = Un.E;
    if (bad) {
        atomicMin(err, ((unsigned long long)blockIdx.y << 32) | ((unsigned long long)j << 16) | (unsigned long long)i);
        return;
    }
    stW(W, ncp, c, W_of_U(Un));
}

// Uo = U over the interior (the stage-1 copy of dUdt_Multistage_Explicit)
__global__ void copy_interior_kernel(const BlockMeta *meta, const double *src, double *dst, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj || m.skip) return;
    const int i = q % m.ni + 2, j = q / m.ni + 2;
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    for (int v = 0; v < 4; ++v) dst[v * ncp + c] = src[v * ncp + c];
}

// residual norms: per-tile partials, then one workgroup sums them in order
__global__ __launch_bounds__(256) void norm_partial_kernel(const BlockMeta *meta, const TileDesc *tiles, int ti,
                                                           int tj, const double *dU, int64_t ncp, int var,
                                                           double *part) {
    __shared__ double sm[4][3];
    const TileDesc t = tiles[blockIdx.x];
    const BlockMeta m = meta[t.b];
    if (m.skip) return;   // a converged case keeps its last partials
    double a = 0.0, b = 0.0, mx = 0.0;
    for (int q = threadIdx.x; q < ti * tj; q += blockDim.x) {
        const int ip = t.i0 + 2 + q % ti, jp = t.j0 + 2 + q / ti;
        if (ip > m.ni + 1 || jp > m.nj + 1) continue;
        const double v = dU[(int64_t)(var - 1) * ncp + m.coff + (int64_t)jp * m.pitch + ip];
        a += fabs(v);
        b += v * v;
        mx = dmax(mx, fabs(v));
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off); b += __shfl_xor(b, off); mx = dmax(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) { sm[threadIdx.x >> 6][0] = a; sm[threadIdx.x >> 6][1] = b; sm[threadIdx.x >> 6][2] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double A = sm[0][0], B = sm[0][1], C = sm[0][2];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { A += sm[w][0]; B += sm[w][1]; C = dmax(C, sm[w][2]); }
        part[3 * blockIdx.x] = A; part[3 * blockIdx.x + 1] = B; part[3 * blockIdx.x + 2] = C;
    }
}

__global__ __launch_bounds__(256) void norm_final_kernel(const double *part, int n, double *out) {
    __shared__ double sm[4][3];
    double a = 0.0, b = 0.0, mx = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) { a += part[3 * q]; b += part[3 * q + 1]; mx = dmax(mx, part[3 * q + 2]); }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off); b += __shfl_xor(b, off); mx = dmax(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) { sm[threadIdx.x >> 6][0] = a; sm[threadIdx.x >> 6][1] = b; sm[threadIdx.x >> 6][2] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double A = sm[0][0], B = sm[0][1], C = sm[0][2];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { A += sm[w][0]; B += sm[w][1]; C = dmax(C, sm[w][2]); }
        out[0] = A; out[1] = sqrt(B); out[2] = C;
    }
}

// The L2 norm of each part's cells from the per-tile partials of
// norm_partial_kernel (sum of squares at 3q + 1): one wavefront per part, the
// lanes over the part's tiles from its first, so a part's norm is the same
// alone and in a batch.
__global__ __launch_bounds__(256) void part_norm_kernel(const double *part, const int32_t *tile0, int npart,
                                                        double *out) {
    const int pt = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (pt >= npart) return;
    double b = 0.0;
    for (int q = tile0[pt] + lane; q < tile0[pt + 1]; q += 64) b += part[3 * q + 1];
    for (int off = 32; off > 0; off >>= 1) b += __shfl_xor(b, off);
    if (lane == 0) out[pt] = sqrt(b);
}

// One workgroup closes a fused time step.  Partials are summed in tile order
// with a fixed thread assignment, so the norms are reproducible run to run.
// Skipped when a stage of this step (or an earlier step) raised the stop
// flag: a failing step is not counted, as in the reference's loop.
// Without parts the 16 waves stride over all tiles; with parts each wave
// reduces whole parts (lanes stride over the part's tiles), so a case gives
// the same norms whether it runs alone or in a batch.
constexpr int FINISH_NT = 1024;
__global__ __launch_bounds__(FINISH_NT) void step_finish_kernel(FinishParams p) {
    constexpr int NW = FINISH_NT / 64;
    __shared__ double sm[NW][4];
    __shared__ int sdone[NW];
    if (*p.stop_flag) return;
    const int n = p.ntiles, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = *p.step_ctr;
    double a = 0.0, b = 0.0, mx = 0.0, dmn = MILLION;
    int ndone = 0;
    if (p.npart == 0) {
        // partials of the tiles, or of chunks of tiles (chunk_partial_kernel)
        const double *src = p.nchunk > 0 ? p.chunk_part : p.norm_part;
        const int ns = p.nchunk > 0 ? p.nchunk : n;
        for (int q = threadIdx.x; q < ns; q += FINISH_NT) {
            a += src[q];
            b += src[ns + q];
            mx = dmax(mx, src[2 * ns + q]);
            dmn = dmin(dmn, src[3 * ns + q]);
        }
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off); b += __shfl_xor(b, off);
            mx = dmax(mx, __shfl_xor(mx, off)); dmn = dmin(dmn, __shfl_xor(dmn, off));
        }
    } else {
        for (int pt = w; pt < p.npart; pt += NW) {
            double pa = 0.0, pb = 0.0, pm = 0.0, pd = MILLION;
            for (int q = p.part_tile0[pt] + lane; q < p.part_tile0[pt + 1]; q += 64) {
                pa += p.norm_part[q];
                pb += p.norm_part[n + q];
                pm = dmax(pm, p.norm_part[2 * n + q]);
                pd = dmin(pd, p.norm_part[3 * n + q]);
            }
            for (int off = 32; off > 0; off >>= 1) {
                pa += __shfl_xor(pa, off); pb += __shfl_xor(pb, off);
                pm = dmax(pm, __shfl_xor(pm, off)); pd = dmin(pd, __shfl_xor(pd, off));
            }
            a += pa; b += pb; mx = dmax(mx, pm); dmn = dmin(dmn, pd);
            if (lane == 0) {
                const double l2p = sqrt(pb);
                int st = p.part_state[pt];
                if (st != PART_DONE) {
                    if (p.part_l2) p.part_l2[(int64_t)s * p.npart + pt] = l2p;
                    p.part_last[pt] = l2p;
                    if (p.tol > 0.0 && l2p <= p.tol) {
                        st = PART_DONE;
                        p.part_steps[pt] = p.step_offset + s + 1;
                    } else if (st == PART_RUNNING && p.freeze_level > 0.0 && l2p <= p.freeze_level) {
                        st = PART_FROZEN;
                    } else if (st == PART_RUNNING && p.stall_window > 0) {
                        // a new minimum must be 1% below the best so far, so
                        // that a cycling norm's small dips do not count
                        const int gs = p.step_offset + s;
                        if (l2p < 0.99 * p.part_best[pt]) {
                            p.part_best[pt] = l2p;
                            p.part_best_step[pt] = gs;
                        } else if (gs - p.part_best_step[pt] >= p.stall_window) {
                            st = PART_FROZEN;
                        }
                    }
                    p.part_state[pt] = st;
                } else if (p.part_l2) {
                    p.part_l2[(int64_t)s * p.npart + pt] = 0.0;
                }
                ndone += st == PART_DONE;
            }
        }
    }
    if (lane == 0) { sm[w][0] = a; sm[w][1] = b; sm[w][2] = mx; sm[w][3] = dmn; sdone[w] = ndone; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double A = sm[0][0], B = sm[0][1], C = sm[0][2], D = sm[0][3];
        int nd = sdone[0];
        for (int k = 1; k < NW; ++k) {
            A += sm[k][0]; B += sm[k][1]; C = dmax(C, sm[k][2]); D = dmin(D, sm[k][3]); nd += sdone[k];
        }
        const double l2n = sqrt(B);
        p.norm_out[3 * s + 0] = A;
        p.norm_out[3 * s + 1] = l2n;
        p.norm_out[3 * s + 2] = C;
        p.dt_out[s] = bitsd(*p.dtmin_in);
        *p.dtmin_out = dbits(D);
        if (p.step_clock) p.step_clock[s] = wall_clock64();
        *p.step_ctr = s + 1;
        if (p.npart == 0 ? (p.tol > 0.0 && l2n <= p.tol) : nd == p.npart) *p.stop_flag = 1;
    }
}

// FP32 -> FP64 copy (the mixed path's residual accumulator into dUdt[0])
__global__ void widen_kernel(const float *src, double *dst, int64_t n) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (double)src[q];
}

// interior W = W(U) (FP64 arithmetic), stored as ST
template <typename ST>
__global__ void w_from_u_kernel(const BlockMeta *meta, const double *U, ST *W, int64_t ncp) {
    const BlockMeta m = meta[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.ni * m.nj) return;
    const int i = q % m.ni + 2, j = q / m.ni + 2;
    const int64_t c = m.coff + (int64_t)j * m.pitch + i;
    const CS<double> Uc = {U[c], U[ncp + c], U[2 * ncp + c], U[3 * ncp + c]};
    stW(W, ncp, c, W_of_U(Uc));
}

// Residual_Smoothing, Euler2DQuadSingleBlock.cc:4521-4564, one workgroup per
// block.  T (the scratch residual dUdt[k+1]) takes S = dUdt[k] on the cells
// one away from the block edge; Gauss-Seidel sweeps in the reference's
// lexicographic order (j outer, i inner) run as anti-diagonal wavefronts: a
// cell's new value needs the W and S neighbours of the same sweep (the
// previous diagonal) and the E and N neighbours of the previous sweep (the
// next diagonal), so each diagonal is one parallel step and the results are
// the sequential ones bit for bit.  Blocks of up to 32 x 32 cells sweep in
// LDS, larger ones in place.
constexpr int RS_NT = 256, RS_LDS = 32;
__global__ __launch_bounds__(RS_NT) void residual_smoothing_kernel(const BlockMeta *meta, double *S, double *T,
                                                                   int64_t ncp, double eps, int iters,
                                                                   const int32_t *part_state,
                                                                   const int32_t *stop_flag) {
    __shared__ double sT[4][RS_LDS][RS_LDS], sR[4][RS_LDS][RS_LDS];
    if (stop_flag && *stop_flag) return;
    const BlockMeta m = meta[blockIdx.x];
    if (part_state && part_state[m.spare] == PART_DONE) return;
    const int tid = threadIdx.x;
    const int I = m.ni - 2, J = m.nj - 2;  // smoothed cells: interior minus its edge ring
    if (I <= 0 || J <= 0) return;
    auto cidx = [&](int i, int j) { return m.coff + (int64_t)(j + 2) * m.pitch + (i + 2); };  // interior (i, j)
    const double den = 1.0 + 4.0 * eps;
    const bool lds = m.ni <= RS_LDS && m.nj <= RS_LDS;
    // T = S on the smoothed cells (interior coordinates 1..ni-2)
    for (int q = tid; q < I * J; q += RS_NT) {
        const int i = 1 + q % I, j = 1 + q / I;
        const int64_t c = cidx(i, j);
        for (int v = 0; v < 4; ++v) T[v * ncp + c] = S[v * ncp + c];
    }
    if (lds) {
        for (int q = tid; q < m.ni * m.nj; q += RS_NT) {
            const int i = q % m.ni, j = q / m.ni;
            const int64_t c = cidx(i, j);
            for (int v = 0; v < 4; ++v) {
                sR[v][j][i] = S[v * ncp + c];
                sT[v][j][i] = (i >= 1 && i <= I && j >= 1 && j <= J) ? S[v * ncp + c] : T[v * ncp + c];
            }
        }
    }
    __syncthreads();
    for (int n = 0; n < iters; ++n) {
        for (int d = 0; d <= I + J - 2; ++d) {
            const int lo = d - (J - 1) > 0 ? d - (J - 1) : 0, hi = d < I - 1 ? d : I - 1;
            for (int a = lo + tid; a <= hi; a += RS_NT) {
                const int i = 1 + a, j = 1 + (d - a);
                if (lds) {
                    for (int v = 0; v < 4; ++v)
                        sT[v][j][i] = (sR[v][j][i] + eps * (sT[v][j - 1][i] + sT[v][j][i - 1] + sT[v][j][i + 1] +
                                                            sT[v][j + 1][i])) / den;
                } else {
                    const int64_t c = cidx(i, j), cs = cidx(i, j - 1), cw = cidx(i - 1, j), ce = cidx(i + 1, j),
                                  cn = cidx(i, j + 1);
                    for (int v = 0; v < 4; ++v) {
                        const int64_t o = v * ncp;
                        T[o + c] = (S[o + c] + eps * (T[o + cs] + T[o + cw] + T[o + ce] + T[o + cn])) / den;
                    }
                }
            }
            __syncthreads();
        }
    }
    // S = T on the smoothed cells
    for (int q = tid; q < I * J; q += RS_NT) {
        const int i = 1 + q % I, j = 1 + q / I;
        const int64_t c = cidx(i, j);
        for (int v = 0; v < 4; ++v) {
            const double x = lds ? sT[v][j][i] : T[v * ncp + c];
            if (lds) T[v * ncp + c] = x;
            S[v * ncp + c] = x;
        }
    }
}

// device wall clock (constant rate, hipDeviceAttributeWallClockRate kHz)
__global__ void stamp_kernel(unsigned long long *dst) {
    if (threadIdx.x == 0) *dst = wall_clock64();
}

// First level of the step reduction for large meshes: workgroup k sums the
// tile partials of chunk k (a fixed contiguous range, so the order is fixed).
constexpr int CHUNK_NT = 256;
__global__ __launch_bounds__(CHUNK_NT) void chunk_partial_kernel(const double *part, int n, double *out, int nchunk,
                                                                 const int32_t *stop_flag) {
    __shared__ double sm[CHUNK_NT / 64][4];
    if (*stop_flag) return;
    const int k = blockIdx.x;
    const int q0 = (int)((int64_t)n * k / nchunk), q1 = (int)((int64_t)n * (k + 1) / nchunk);
    double a = 0.0, b = 0.0, mx = 0.0, dmn = MILLION;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += CHUNK_NT) {
        a += part[q]; b += part[n + q]; mx = dmax(mx, part[2 * n + q]); dmn = dmin(dmn, part[3 * n + q]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off); b += __shfl_xor(b, off);
        mx = dmax(mx, __shfl_xor(mx, off)); dmn = dmin(dmn, __shfl_xor(dmn, off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[w][0] = a; sm[w][1] = b; sm[w][2] = mx; sm[w][3] = dmn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < CHUNK_NT / 64; ++j) {
            sm[0][0] += sm[j][0]; sm[0][1] += sm[j][1]; sm[0][2] = dmax(sm[0][2], sm[j][2]);
            sm[0][3] = dmin(sm[0][3], sm[j][3]);
        }
        out[k] = sm[0][0]; out[nchunk + k] = sm[0][1]; out[2 * nchunk + k] = sm[0][2]; out[3 * nchunk + k] = sm[0][3];
    }
}

// small device-side scalar writes (graph-safe, no memset nodes)
__global__ void set_u64_kernel(unsigned long long *dst, unsigned long long v0, unsigned long long v1, int n) {
    if (threadIdx.x == 0) { dst[0] = v0; if (n > 1) dst[1] = v1; }
}

// ---------------------------------------------------------------------------
// launchers
}  // namespace cffc

using namespace cffc;

// The stage kernels' FP64 arithmetic beside the compiler's own, per operand
// pair: rdiv(a, recip(b)) and a / b, qsqrt(|a|) and sqrt(|a|), vmin64 and
// fmin, vmax64 and fmax (out: 8 values per pair).
__global__ void arith_check_kernel(const double *a, const double *b, int64_t n, double *out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = a[i], y = b[i];
    double *o = out + 8 * i;
    o[0] = rdiv(x, recip(y));
    o[1] = x / y;
    o[2] = qsqrt(fabs(x));
    o[3] = sqrt(fabs(x));
    o[4] = vmin64(x, y);
    o[5] = fmin(x, y);
    o[6] = vmax64(x, y);
    o[7] = fmax(x, y);
}

extern "C" {

hipError_t cffc_launch_geometry(const BlockMeta *meta, int nblk, int max_cells, const double *xn, const double *yn,
                                double *geom, int64_t ncp, unsigned int *ndeg, hipStream_t st) {
    dim3 grid((max_cells + 255) / 256, nblk);
    hipLaunchKernelGGL(geometry_kernel, grid, dim3(256), 0, st, meta, xn, yn, geom, ncp, ndeg);
    return hipGetLastError();
}
hipError_t cffc_launch_geometry_f32(const BlockMeta *meta, int nblk, int max_cells, const double *geom, float *gf,
                                    int64_t ncp, hipStream_t st) {
    dim3 grid((max_cells + 255) / 256, nblk);
    hipLaunchKernelGGL(geometry_f32_kernel, grid, dim3(256), 0, st, meta, geom, gf, ncp);
    return hipGetLastError();
}
hipError_t cffc_launch_scatter(const BlockMeta *meta, int nblk, int max_int, const double *aos, double *W, int64_t ncp,
                               int kind, double *U, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(scatter_state_kernel, grid, dim3(256), 0, st, meta, aos, W, ncp, kind, U);
    return hipGetLastError();
}
hipError_t cffc_launch_gather(const BlockMeta *meta, int nblk, int max_int, const double *src, double *aos,
                              int64_t ncp, int nvar, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(gather_state_kernel, grid, dim3(256), 0, st, meta, src, aos, ncp, nvar);
    return hipGetLastError();
}
hipError_t cffc_launch_exchange(const ExchangeParams *p, int nblk, hipStream_t st) {
    const int nt = (p->nthreads == 64 || p->nthreads == 128) ? p->nthreads : 256;
    if (p->Wf) hipLaunchKernelGGL(exchange_bc_kernel<float>, dim3(nblk), dim3(nt), 0, st, *p);
    else hipLaunchKernelGGL(exchange_bc_kernel<double>, dim3(nblk), dim3(nt), 0, st, *p);
    return hipGetLastError();
}
hipError_t cffc_launch_arith_check(const double *a, const double *b, int64_t n, double *out, hipStream_t st) {
    hipLaunchKernelGGL(arith_check_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, n, out);
    return hipGetLastError();
}
hipError_t cffc_launch_part_norms(const double *part, const int32_t *tile0, int npart, double *out,
                                  hipStream_t st) {
    hipLaunchKernelGGL(part_norm_kernel, dim3((unsigned)((npart + 3) / 4)), dim3(256), 0, st, part, tile0, npart,
                       out);
    return hipGetLastError();
}
hipError_t cffc_launch_widen(const float *src, double *dst, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}
hipError_t cffc_launch_cfl(const BlockMeta *meta, const TileDesc *tiles, int ntiles, int ti, int tj, const double *W,
                           const double *geom, int64_t ncp, double *dt, unsigned long long *dtmin, hipStream_t st) {
    // one thread per tile cell, in whole wavefronts up to 256
    const int nt = ti * tj <= 64 ? 64 : ti * tj <= 128 ? 128 : 256;
    hipLaunchKernelGGL(cfl_kernel, dim3(ntiles), dim3(nt), 0, st, meta, tiles, ti, tj, W, geom, ncp, dt, dtmin);
    return hipGetLastError();
}
hipError_t cffc_launch_set_global_dt(const BlockMeta *meta, int nblk, int max_int, double *dt,
                                     const unsigned long long *dtmin, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(set_global_dt_kernel, grid, dim3(256), 0, st, meta, dt, dtmin);
    return hipGetLastError();
}
hipError_t cffc_launch_update(const BlockMeta *meta, int nblk, int max_int, double *W, double *U, const double *Uo,
                              double *dU, double *dt, int64_t ncp, double omega, int lts, unsigned long long *err,
                              hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(update_kernel, grid, dim3(256), 0, st, meta, W, U, Uo, dU, dt, ncp, omega, lts, err);
    return hipGetLastError();
}
hipError_t cffc_launch_copy_interior(const BlockMeta *meta, int nblk, int max_int, const double *src, double *dst,
                                     int64_t ncp, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    hipLaunchKernelGGL(copy_interior_kernel, grid, dim3(256), 0, st, meta, src, dst, ncp);
    return hipGetLastError();
}
hipError_t cffc_launch_norms(const BlockMeta *meta, const TileDesc *tiles, int ntiles, int ti, int tj,
                             const double *dU, int64_t ncp, int var, double *part, double *out, hipStream_t st) {
    hipLaunchKernelGGL(norm_partial_kernel, dim3(ntiles), dim3(256), 0, st, meta, tiles, ti, tj, dU, ncp, var, part);
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(256), 0, st, part, ntiles, out);
    return hipGetLastError();
}
hipError_t cffc_launch_set_u64(unsigned long long *dst, unsigned long long v0, unsigned long long v1, int n,
                               hipStream_t st) {
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(64), 0, st, dst, v0, v1, n);
    return hipGetLastError();
}

hipError_t cffc_launch_w_from_u(const BlockMeta *meta, int nblk, int max_int, const double *U, double *W, float *Wf,
                                int64_t ncp, hipStream_t st) {
    dim3 grid((max_int + 255) / 256, nblk);
    if (Wf) hipLaunchKernelGGL(w_from_u_kernel<float>, grid, dim3(256), 0, st, meta, U, Wf, ncp);
    else hipLaunchKernelGGL(w_from_u_kernel<double>, grid, dim3(256), 0, st, meta, U, W, ncp);
    return hipGetLastError();
}
hipError_t cffc_launch_residual_smoothing(const BlockMeta *meta, int nblk, double *S, double *T, int64_t ncp,
                                          double eps, int iters, const int32_t *part_state, const int32_t *stop_flag,
                                          hipStream_t st) {
    hipLaunchKernelGGL(residual_smoothing_kernel, dim3(nblk), dim3(RS_NT), 0, st, meta, S, T, ncp, eps, iters,
                       part_state, stop_flag);
    return hipGetLastError();
}
hipError_t cffc_launch_update_tile_fp64(const StageParams *p, int ti, int tj, hipStream_t st) {
    hipLaunchKernelGGL((update_tile_kernel<double>), dim3(p->tile_list ? p->nlist : p->ntiles), dim3(UPD_NT), 0, st, *p, ti,
                       tj);
    return hipGetLastError();
}
hipError_t cffc_launch_stamp(unsigned long long *dst, hipStream_t st) {
    hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(64), 0, st, dst);
    return hipGetLastError();
}
hipError_t cffc_launch_finish(const FinishParams *p, double *chunk_part, int nchunk, hipStream_t st) {
    FinishParams q = *p;
    q.chunk_part = nullptr;
    q.nchunk = 0;
    if (p->npart == 0 && nchunk > 0) {
        hipLaunchKernelGGL(chunk_partial_kernel, dim3(nchunk), dim3(CHUNK_NT), 0, st, p->norm_part, p->ntiles,
                           chunk_part, nchunk, p->stop_flag);
        q.chunk_part = chunk_part;
        q.nchunk = nchunk;
    }
    hipLaunchKernelGGL(step_finish_kernel, dim3(1), dim3(FINISH_NT), 0, st, q);
    return hipGetLastError();
}

hipError_t cffc_launch_stage_fp64(const StageParams *p, int ti, int tj, int fused, hipStream_t st) {
#define L(TI, TJ)                                                                                          \
    if (ti == TI && tj == TJ) {                                                                            \
        if (fused) hipLaunchKernelGGL((stage_kernel<double, TI, TJ, true>), dim3(p->ntiles), dim3((TileThreads<TI, TJ>::value)), 0, st, *p); \
        else hipLaunchKernelGGL((stage_kernel<double, TI, TJ, false>), dim3(p->ntiles), dim3((TileThreads<TI, TJ>::value)), 0, st, *p); \
        return hipGetLastError();                                                                          \
    }
    CFFC_TILE_LIST(L)
    CFFC_SMALL_TILE_LIST(L)
#undef L
    return hipErrorInvalidValue;
}

}  // extern "C"
