// euler2d_stage.hpp -- the fused per-stage residual kernel for gfx950.
//
// One workgroup owns a TI x TJ tile of interior cells of one block and, for
// that tile, performs what the reference does per block and stage in
//   Linear_Reconstruction_LeastSquares   Euler2DQuadSingleBlock.cc:3678-4087
//   Limiter_Venkatakrishnan              CFD.cc:728-752
//   dUdt_Multistage_Explicit             Euler2DQuad.cc:1235-1727
//   Update_Solution_Multistage_Explicit  Euler2DQuadSingleBlock.cc:5300-5507
//   CFL (next step) and the norms        Euler2DQuadSingleBlock.cc:3019-3167
// in four LDS-staged phases:
//   1. load W and the cell centroids of the (TI+4) x (TJ+4) window;
//   2. least-squares gradient + limiter for the (TI+2) x (TJ+2) cells whose
//      reconstructions the tile's faces need (the tile plus one ring);
//   3. Roe fluxes of the (TI+1) x TJ i-faces and TI x (TJ+1) j-faces;
//   4. per cell: accumulate +W -E +S -N (the reference's order), update.
// FT is the arithmetic of phases 2-3 (double: FP64 path, bitwise with the
// reference when compiled with -ffp-contract=off; float: mixed path).  The
// accumulation, update, time step and norms are always FP64.
#pragma once

#include "euler2d_device.hpp"
#include "euler2d_state.hpp"

namespace cffc {

struct StageParams {
    const BlockMeta *meta;
    const TileDesc *tiles;
    int32_t ntiles;
    int64_t ncp;            // padded cells of all blocks
    const double *geom;     // GEO_COUNT x ncp
    const float *geomf;     // GF_COUNT x ncp (mixed path)
    const int32_t *bc;      // concatenated BCtypeW/E/S/N
    const double *wo;       // reference states of the bc entries (AoS, 4)
    const double *Win;      // W^(s), 4 x ncp, ghosts filled
    double *Wout;           // W^(s+1) interior (fused) -- may alias Win when !fused
    const float *Winf;      // fused mixed path: W^(s) and W^(s+1) stored as FP32
    float *Woutf;
    double *U;              // conserved U^n (fused path)
    double *S;              // residual accumulator dUdt[k]
    float *Sf;              // fused mixed path: the accumulator in FP32 (null: S)
    const double *S1, *S2;  // dUdt[1], dUdt[2] (granular RK4 stage 4)
    double *dt;             // per-cell time step (padded)
    const unsigned long long *dtmin_in;  // global dt (bit pattern) of this step
    double *phi;            // stored limiter values (freeze), 4 x ncp
    double cfl, omega;
    int32_t mode;           // StageMode (fused) / granular init mode
    int32_t phi_mode;       // PhiMode
    int32_t limiter;        // CFFC_LIMITER_*
    int32_t lts;            // scalar local time stepping
    int32_t last_stage;     // fused: produce norms and the next time step
    unsigned long long *err;   // first failing cell key (atomicMin)
    double *norm_part;      // [4][ntiles]: L1, L2^2, max of the last stage's dUdt (rho), next dt
    int32_t *stop_flag;     // nonzero: converged or failed, skip work
    const int32_t *part_state;  // per-part PartState (batched steady runs), or null
    int32_t done_copy;      // parts that are done copy W forward (odd stage count)
    const int32_t *nbr;     // 8 same-level neighbours per block (push exchange)
    int32_t push;           // fused: also write W^(s+1) into the neighbours' ghost cells
    int32_t defer_update;   // fused: store the residual only; update_tile_kernel follows
    int32_t skip_S;         // fused: no stage after this one reads S (multistage smoothing), so it is not stored
    // resolution changes
    const uint8_t *tile_defer;  // fused: per tile, nonzero = store the residual in Sd (flux correction follows)
    // granular: per-case limiter freezing of a multigrid batch (null: the
    // phi_mode of the launch); the part of every block and whether each
    // part's limiters are frozen
    const int32_t *part_of_block;
    const int32_t *part_frozen;
    double *Sd;             // residual of deferred tiles; update_tile_kernel reads it
    const int32_t *tile_list;   // update_tile_kernel: the tiles to update (null: all)
    int32_t nlist;
    double *eflux;          // saved block-edge fluxes (AMR_FLUX_* faces)
    const uint8_t *ustore;  // per padded cell: U^(s+1) is read by a coarser neighbour's restriction
    double *Ustage;         // fused: U^(s+1) of those cells
    // multigrid levels (granular FP64 kernel only)
    int32_t hlle;           // HLLE flux instead of Roe
    int32_t residual;       // steady residual (F*l)/A, no CFL*dt factor (dUdt_Residual_Evaluation)
    const double *P;        // forcing term added after the fluxes: +(CFL*dt)*P, or +P with residual
    int32_t axisym;         // axisymmetric source (CFL*dt)*S(W, Xc) after the east face
};

// Push form of the same-level exchange: the cell (ip, jp) of block b that a
// neighbour's ghost cell copies (the exchange kernel's phase 0 read in the
// other direction) is written there by the tile that computes it, so the
// next stage finds those ghosts filled without a separate exchange.  Needs
// a symmetric neighbour table (checked by the host).  Slots N, S, E, W, NE,
// NW, SE, SW; the neighbour's ghost index follows from its slot facing b.
template <typename ST>
__device__ __forceinline__ void push_ghosts(const StageParams &p, ST *Wo, const BlockMeta &m, int b, int ip, int jp,
                                            const PS<double> &Wn) {
    const int ICu = m.ni + 1, JCu = m.nj + 1;
    const bool nW = ip <= 3, nE = ip >= ICu - 1, nS = jp <= 3, nN = jp >= JCu - 1;
    if (!(nW || nE || nS || nN)) return;
    const int32_t *nb8 = p.nbr + 8 * b;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const bool hit = s == 0 ? nN : s == 1 ? nS : s == 2 ? nE : s == 3 ? nW
                       : s == 4 ? (nN && nE) : s == 5 ? (nN && nW) : s == 6 ? (nS && nE) : (nS && nW);
        if (!hit) continue;
        const int nb = nb8[s];
        if (nb < 0) continue;
        const BlockMeta S = p.meta[nb];
        const bool up = s == 0 || s == 4 || s == 5, down = s == 1 || s == 6 || s == 7;
        const bool east = s == 2 || s == 4 || s == 6, west = s == 3 || s == 5 || s == 7;
        const int ig = east ? ip - m.ni : west ? ip + S.ni : ip;
        const int jg = up ? jp - m.nj : down ? jp + S.nj : jp;
        const int64_t cg = S.coff + (int64_t)jg * S.pitch + ig;
        Wo[cg] = (ST)Wn.d; Wo[p.ncp + cg] = (ST)Wn.vx; Wo[2 * p.ncp + cg] = (ST)Wn.vy; Wo[3 * p.ncp + cg] = (ST)Wn.p;
    }
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup k to XCD
// k mod 8), each XCD with its own L2.  Tiles are listed block by block, so
// giving XCD x the x-th contiguous eighth of the list keeps neighbouring
// tiles, which share halo rows and face geometry, in one L2.  A bijection of
// [0, n) for any n.
__device__ __forceinline__ int xcd_tile(int k, int n) {
    const int q = n >> 3, r = n & 7, x = k & 7, idx = k >> 3;
    return x * q + (x < r ? x : r) + idx;
}

// granular init modes (dUdt_Multistage_Explicit called on its own)
enum GranularInit { GINIT_ZERO = 0, GINIT_KEEP = 1, GINIT_RK4 = 2 };

// Stencil of Linear_Reconstruction_LeastSquares (SingleBlock.cc:3691-3906)
// as a mask over the neighbours in the reference's listing order
// (i-1,j-1) (i,j-1) (i+1,j-1) (i-1,j) (i+1,j) (i-1,j+1) (i,j+1) (i+1,j+1);
// every reduced stencil of the reference lists its points in that order.
__device__ __forceinline__ bool bc_extrap(int bc) {
    return bc == 6000 || bc == 4001 || bc == 4002 || bc == 8001;
}
__device__ __forceinline__ unsigned stencil_mask(int i, int j, int ICu, int JCu, int bcw, int bce, int bcs,
                                                 int bcn) {
    const int ICl = 2, JCl = 2;
    if (i == 0 || i == ICu + 2 || j == 0 || j == JCu + 2) return 0u;
    if (i == 1 && bcw != 0) {
        if (j == 1 || j == JCu + 1) return 0u;
        if (bc_extrap(bcw)) return j == JCl ? 0xF8u : j == JCu ? 0x1Fu : 0xFFu;
        return j == JCl ? 0xD0u : j == JCu ? 0x16u : 0xD6u;
    }
    if (i == ICu + 1 && bce != 0) {
        if (j == 1 || j == JCu + 1) return 0u;
        if (bc_extrap(bce)) return j == JCl ? 0xF8u : j == JCu ? 0x1Fu : 0xFFu;
        return j == JCl ? 0x68u : j == JCu ? 0x0Bu : 0x6Bu;
    }
    if (j == 1 && bcs != 0) {
        if (i == 1 || i == ICu + 1) return 0u;
        if (bc_extrap(bcs)) return i == ICl ? 0xD6u : i == ICu ? 0x6Bu : 0xFFu;
        return i == ICl ? 0xD0u : i == ICu ? 0x68u : 0xF8u;
    }
    if (j == JCu + 1 && bcn != 0) {
        if (i == 1 || i == ICu + 1) return 0u;
        if (bc_extrap(bcn)) return i == ICl ? 0xD6u : i == ICu ? 0x6Bu : 0xFFu;
        return i == ICl ? 0x16u : i == ICu ? 0x0Bu : 0x1Fu;
    }
    return 0xFFu;
}

__device__ __forceinline__ bool face_bc(int bc) { return bc == 7001 || bc == 8001; }

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long u) { return __longlong_as_double((long long)u); }

// W + (phi^dWdx)*dX.x + (phi^dWdy)*dX.y  (Euler2DQuad.cc:1338-1340)
template <typename FT>
__device__ __forceinline__ PS<FT> face_state(const PS<FT> &W, const PS<FT> &px, const PS<FT> &py, FT dx, FT dy) {
    PS<FT> r;
    r.d = W.d + px.d * dx + py.d * dy;
    r.vx = W.vx + px.vx * dx + py.vx * dy;
    r.vy = W.vy + px.vy * dx + py.vy * dy;
    r.p = W.p + px.p * dx + py.p * dy;
    return r;
}

// Geometry as the two paths read it: the FP64 path reads the FP64 planes
// (GEO_*); the mixed path reads FP32 planes (GF_*) with coordinates relative
// to the block origin, which halves the geometry traffic.
template <typename FT> struct CoordT { typedef double type; };
template <> struct CoordT<float> { typedef float type; };
template <typename FT>
__device__ __forceinline__ typename CoordT<FT>::type rdgeo(const StageParams &p, int gd, int gf, int64_t i) {
    if constexpr (Fast<FT>::on) return p.geomf[gf * p.ncp + i];
    else return p.geom[gd * p.ncp + i];
}

// Least-squares gradient and Venkatakrishnan limiter of one cell
// (Linear_Reconstruction_LeastSquares, SingleBlock.cc:3678-4087).  FULL: the
// 8-point stencil, straight-line; otherwise the reduced stencil in mask.
// The window arrays are indexed [row][col] with the cell at (jl, il).
template <typename FT, bool FULL, int LJ, int LI, typename XT>
__device__ __forceinline__ void reconstruct(const StageParams &p, int phim, int64_t c, int64_t ncp, int pitch, int il,
                                            int jl, unsigned mask, const FT (&sW)[4][LJ][LI], const XT (&sX)[LJ][LI],
                                            const XT (&sY)[LJ][LI], PS<FT> &gx, PS<FT> &gy, PS<FT> &ph) {
    const PS<FT> W0 = {sW[0][jl][il], sW[1][jl][il], sW[2][jl][il], sW[3][jl][il]};
    gx = {FT(0), FT(0), FT(0), FT(0)};
    gy = gx;
    ph = gx;
    if (!FULL && mask == 0u) return;
    PS<FT> Dx = {FT(0), FT(0), FT(0), FT(0)}, Dy = Dx;
    FT DxDx = FT(0), DxDy = FT(0), DyDy = FT(0);
    const XT xc0 = sX[jl][il], yc0 = sY[jl][il];
    int n_pts = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int kk = s < 4 ? s : s + 1;  // skip the centre of the 3x3 window
        const int a = il + (kk % 3) - 1, b = jl + (kk / 3) - 1;
        const FT dX = (FT)(sX[b][a] - xc0), dY = (FT)(sY[b][a] - yc0);
        const PS<FT> DU = {sW[0][b][a] - W0.d, sW[1][b][a] - W0.vx, sW[2][b][a] - W0.vy, sW[3][b][a] - W0.p};
        if (FULL || (mask & (1u << s))) {
            Dx.d += DU.d * dX; Dx.vx += DU.vx * dX; Dx.vy += DU.vy * dX; Dx.p += DU.p * dX;
            Dy.d += DU.d * dY; Dy.vx += DU.vx * dY; Dy.vy += DU.vy * dY; Dy.p += DU.p * dY;
            DxDx += dX * dX; DxDy += dX * dY; DyDy += dY * dY;
            ++n_pts;
        }
    }
    // averages over the stencil: x/8 and x*0.125 round identically, so the
    // full stencil multiplies; reduced stencils divide (reference: /n_pts)
    if constexpr (Fast<FT>::on) {
        // mixed path: the common 1/n_pts of all sums cancels in the solve
    } else if (FULL) {
        const FT rn = FT(0.125);
        Dx.d *= rn; Dx.vx *= rn; Dx.vy *= rn; Dx.p *= rn;
        Dy.d *= rn; Dy.vx *= rn; Dy.vy *= rn; Dy.p *= rn;
        DxDx *= rn; DxDy *= rn; DyDy *= rn;
    } else {
        const Recip<FT> rn = recip((FT)n_pts);
        Dx.d = rdiv(Dx.d, rn); Dx.vx = rdiv(Dx.vx, rn); Dx.vy = rdiv(Dx.vy, rn); Dx.p = rdiv(Dx.p, rn);
        Dy.d = rdiv(Dy.d, rn); Dy.vx = rdiv(Dy.vx, rn); Dy.vy = rdiv(Dy.vy, rn); Dy.p = rdiv(Dy.p, rn);
        DxDx = rdiv(DxDx, rn); DxDy = rdiv(DxDy, rn); DyDy = rdiv(DyDy, rn);
    }
    const FT det = DxDx * DyDy - DxDy * DxDy;
    if constexpr (Fast<FT>::on) {
        const FT rd = __builtin_amdgcn_rcpf(det);
        gx.d = (Dx.d * DyDy - Dy.d * DxDy) * rd; gx.vx = (Dx.vx * DyDy - Dy.vx * DxDy) * rd;
        gx.vy = (Dx.vy * DyDy - Dy.vy * DxDy) * rd; gx.p = (Dx.p * DyDy - Dy.p * DxDy) * rd;
        gy.d = (Dy.d * DxDx - Dx.d * DxDy) * rd; gy.vx = (Dy.vx * DxDx - Dx.vx * DxDy) * rd;
        gy.vy = (Dy.vy * DxDx - Dx.vy * DxDy) * rd; gy.p = (Dy.p * DxDx - Dx.p * DxDy) * rd;
    } else {
        // eight correctly rounded quotients over one shared reciprocal of det
        const Recip<FT> rd = recip(det);
        gx.d = rdiv(Dx.d * DyDy - Dy.d * DxDy, rd); gx.vx = rdiv(Dx.vx * DyDy - Dy.vx * DxDy, rd);
        gx.vy = rdiv(Dx.vy * DyDy - Dy.vy * DxDy, rd); gx.p = rdiv(Dx.p * DyDy - Dy.p * DxDy, rd);
        gy.d = rdiv(Dy.d * DxDx - Dx.d * DxDy, rd); gy.vx = rdiv(Dy.vx * DxDx - Dx.vx * DxDy, rd);
        gy.vy = rdiv(Dy.vy * DxDx - Dx.vy * DxDy, rd); gy.p = rdiv(Dy.p * DxDx - Dx.p * DxDy, rd);
    }
    if (phim == PHI_LOAD) {
        ph.d = (FT)p.phi[c]; ph.vx = (FT)p.phi[ncp + c]; ph.vy = (FT)p.phi[2 * ncp + c]; ph.p = (FT)p.phi[3 * ncp + c];
        return;
    }
    if (p.limiter == -1) { ph.d = ph.vx = ph.vy = ph.p = FT(1); return; }
    if (p.limiter == 0) { ph.d = ph.vx = ph.vy = ph.p = FT(0); return; }
    // face-midpoint offsets: E, W (= E of i-1), N, S (= N of j-1)
    const FT dxE = (FT)(rdgeo<FT>(p, GEO_XFE, GF_XFE, c) - xc0), dyE = (FT)(rdgeo<FT>(p, GEO_YFE, GF_YFE, c) - yc0);
    const FT dxW = (FT)(rdgeo<FT>(p, GEO_XFE, GF_XFE, c - 1) - xc0);
    const FT dyW = (FT)(rdgeo<FT>(p, GEO_YFE, GF_YFE, c - 1) - yc0);
    const FT dxN = (FT)(rdgeo<FT>(p, GEO_XFN, GF_XFN, c) - xc0), dyN = (FT)(rdgeo<FT>(p, GEO_YFN, GF_YFN, c) - yc0);
    const FT dxS = (FT)(rdgeo<FT>(p, GEO_XFN, GF_XFN, c - pitch) - xc0);
    const FT dyS = (FT)(rdgeo<FT>(p, GEO_YFN, GF_YFN, c - pitch) - yc0);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const FT u0 = n == 0 ? W0.d : n == 1 ? W0.vx : n == 2 ? W0.vy : W0.p;
        FT umin = u0, umax = u0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int kk = s < 4 ? s : s + 1;
            const FT v = sW[n][jl + (kk / 3) - 1][il + (kk % 3) - 1];
            if (FULL || (mask & (1u << s))) {
                if constexpr (Fast<FT>::on) { umin = fminf(umin, v); umax = fmaxf(umax, v); }
                // fmin/fmax (one instruction) equal the reference's
                // compare-and-select min/max here: they differ only in the sign
                // of a zero result, and the limiter uses u_max/u_min only
                // through u_max - u_0 and u_min - u_0, whose zero sign can only
                // reach y = +-0, where f(y) = (y^2 + 2y)/(y^2 + y + 2) is +0
                else { umin = vmin64(umin, v); umax = vmax64(umax, v); }
            }
        }
        const FT gxn = n == 0 ? gx.d : n == 1 ? gx.vx : n == 2 ? gx.vy : gx.p;
        const FT gyn = n == 0 ? gy.d : n == 1 ? gy.vx : n == 2 ? gy.vy : gy.p;
        FT phi = FT(1);
        if constexpr (Fast<FT>::on) {
            // mixed path: min(1, f(y)) is non-decreasing in y >= 0 (f < 1
            // rising on [0, 2), f >= 1 beyond), so the minimum over the four
            // quadrature points is f at the smallest y, capped at y = 2
            // (f(2) = 1); dead-band points do not lower it
            const FT up = umax - u0, dn = umin - u0;
            FT y = FT(2);
            const FT dq[4] = {gxn * dxE + gyn * dyE, gxn * dxW + gyn * dyW, gxn * dxN + gyn * dyN,
                              gxn * dxS + gyn * dyS};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const FT du = dq[q];
                const FT yq = (du > FT(0) ? up : dn) * __builtin_amdgcn_rcpf(du);
                y = fabsf(du) > Consts<FT>::TOLER ? fminf(y, yq) : y;
            }
            phi = (y * y + FT(2) * y) * __builtin_amdgcn_rcpf(y * y + y + FT(2));
        } else {
            phi = venkat_fold(phi, u0 + gxn * dxE + gyn * dyE, u0, umin, umax);
            phi = venkat_fold(phi, u0 + gxn * dxW + gyn * dyW, u0, umin, umax);
            phi = venkat_fold(phi, u0 + gxn * dxN + gyn * dyN, u0, umin, umax);
            phi = venkat_fold(phi, u0 + gxn * dxS + gyn * dyS, u0, umin, umax);
        }
        if (n == 0) ph.d = phi; else if (n == 1) ph.vx = phi; else if (n == 2) ph.vy = phi; else ph.p = phi;
    }
}

// The update's rejection test U.d <= 0 || U.E <= 0 || U.e() <= 0
// (Euler2DQuadSingleBlock.cc:5367-5371).  e() costs two divisions; when
// E*d exceeds 0.5*|dv|^2 by a relative margin of 1e-6 (orders of magnitude
// above rounding) e() is certainly positive and is not evaluated, so the
// outcome is the reference's exactly.
template <typename FT> __device__ __forceinline__ bool positivity_fails(const CS<double> &U) {
    if (U.d <= 0.0 || U.E <= 0.0) return true;
    const double k2 = 0.5 * (U.mx * U.mx + U.my * U.my);
    if (U.E * U.d > k2 * (1.0 + 1e-6)) return false;
    if constexpr (Fast<FT>::on) return U.E * U.d <= k2;
    else return U_e(U) <= 0.0;
}

// The fused update of one cell from its accumulated residual dU
// (Update_Solution_Multistage_Explicit, Euler2DQuadSingleBlock.cc:5300-5507):
// U^(s+1) = U^n + omega dU with the positivity test and the scalar
// local-time-step halving, W^(s+1) (and its pushed ghost copies), the
// accumulator, and at the last stage U^(n+1), the norm terms and the next
// step's CFL time step (SingleBlock.cc:3036-3053).
// lW, lE, lS, lN: the cell's face lengths as the stage read them (FP64
// planes in the FP64 path, FP32 planes in the mixed path); U0: U^n of the
// cell; Sacc: the FP64 accumulator's value as the caller loaded it, if
// have_Sacc.
template <typename FT>
__device__ __forceinline__ void fused_update(const StageParams &p, const BlockMeta &m, const TileDesc &t, int64_t c,
                                             int ip, int jp, double A, double dtc, double lW, double lE, double lS,
                                             double lN, const double (&U0)[4], bool have_Sacc,
                                             const double (&Sacc)[4], double (&dU)[4], double &l1, double &l2,
                                             double &mx, double &dtn) {
    const double om = p.omega;
    CS<double> Un;
    Un.d = U0[0] + om * dU[0]; Un.mx = U0[1] + om * dU[1];
    Un.my = U0[2] + om * dU[2]; Un.E = U0[3] + om * dU[3];
    bool bad = positivity_fails<FT>(Un);
    if (bad && p.lts) {
        // scalar local time stepping: halve dt and dUdt up to 10 times
        double f = 1.0, dtl = dtc;
        for (int n = 1; n <= 10; ++n) {
            f = 0.5 * f;
            dtl = f * dtl;
#pragma unroll
            for (int v = 0; v < 4; ++v) dU[v] = f * dU[v];
            Un.d = U0[0] + om * dU[0]; Un.mx = U0[1] + om * dU[1];
            Un.my = U0[2] + om * dU[2]; Un.E = U0[3] + om * dU[3];
            if (!positivity_fails<FT>(Un)) break;
        }
        p.dt[c] = dtl;
        bad = positivity_fails<FT>(Un);
    }
    if (bad) {
        const unsigned long long key =
            ((unsigned long long)t.b << 32) | ((unsigned long long)jp << 16) | (unsigned long long)ip;
        atomicMin(p.err, key);
        // the reference returns from the failing stage; later kernels skip
        if (p.stop_flag) atomicExch(p.stop_flag, 1);
    }
    PS<double> Wn;
    if constexpr (Fast<FT>::on) {
        const double r = 1.0 / Un.d;
        Wn.d = Un.d; Wn.vx = Un.mx * r; Wn.vy = Un.my * r;
        Wn.p = gm1<double>() * (Un.E - 0.5 * (Un.mx * Wn.vx + Un.my * Wn.vy));
    } else {
        Wn = W_of_U(Un);
    }
    if constexpr (Fast<FT>::on) {
        p.Woutf[c] = (float)Wn.d; p.Woutf[p.ncp + c] = (float)Wn.vx;
        p.Woutf[2 * p.ncp + c] = (float)Wn.vy; p.Woutf[3 * p.ncp + c] = (float)Wn.p;
        if (p.push) push_ghosts(p, p.Woutf, m, t.b, ip, jp, Wn);
    } else {
        p.Wout[c] = Wn.d; p.Wout[p.ncp + c] = Wn.vx; p.Wout[2 * p.ncp + c] = Wn.vy; p.Wout[3 * p.ncp + c] = Wn.p;
        if (p.push) push_ghosts(p, p.Wout, m, t.b, ip, jp, Wn);
    }
    if ((m.amr & AMR_USTORE) && p.Ustage && p.ustore[c]) {
        // a coarser neighbour restricts U^(s+1) of this cell into its ghosts
        p.Ustage[c] = Un.d; p.Ustage[p.ncp + c] = Un.mx; p.Ustage[2 * p.ncp + c] = Un.my; p.Ustage[3 * p.ncp + c] = Un.E;
    }
    if (p.skip_S) {
        // the next stage starts from zero and the last stage stores dUdt[0]
    } else if (Fast<FT>::on && p.Sf) {
        // mixed path: the RK accumulator holds FP32-flux residuals, so it
        // is kept in FP32 (half the bytes); the stage's own dU stays FP64
        if (p.mode == MODE_ZERO_ACC2) {
#pragma unroll
            for (int v = 0; v < 4; ++v) p.Sf[v * p.ncp + c] = (float)((double)p.Sf[v * p.ncp + c] + 2.0 * dU[v]);
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) p.Sf[v * p.ncp + c] = (float)dU[v];
        }
    } else if (p.mode == MODE_ZERO_ACC2) {
#pragma unroll
        for (int v = 0; v < 4; ++v) p.S[v * p.ncp + c] = (have_Sacc ? Sacc[v] : p.S[v * p.ncp + c]) + 2.0 * dU[v];
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) p.S[v * p.ncp + c] = dU[v];
    }
    if (p.last_stage) {
        p.U[c] = Un.d; p.U[p.ncp + c] = Un.mx; p.U[2 * p.ncp + c] = Un.my; p.U[3 * p.ncp + c] = Un.E;
        l1 = fabs(dU[0]);
        l2 = dU[0] * dU[0];
        mx = l1;
        // CFL of the next step, SingleBlock.cc:3036-3053, in FP64 from the
        // geometry planes the stage reads (the mixed path's FP32 normals and
        // lengths are still in cache from its flux phase)
        const double d_i = 2.0 * qdiv(A, lE + lW);
        const double d_j = 2.0 * qdiv(A, lN + lS);
        const double nEx = rdgeo<FT>(p, GEO_NEX, GF_NEX, c), nEy = rdgeo<FT>(p, GEO_NEY, GF_NEY, c);
        const double nNx = rdgeo<FT>(p, GEO_NNX, GF_NNX, c), nNy = rdgeo<FT>(p, GEO_NNY, GF_NNY, c);
        const double nWx = -(double)rdgeo<FT>(p, GEO_NEX, GF_NEX, c - 1);
        const double nWy = -(double)rdgeo<FT>(p, GEO_NEY, GF_NEY, c - 1);
        const double nSx = -(double)rdgeo<FT>(p, GEO_NNX, GF_NNX, c - m.pitch);
        const double nSy = -(double)rdgeo<FT>(p, GEO_NNY, GF_NNY, c - m.pitch);
        const double v_i = 0.5 * (Wn.vx * (nEx - nWx) + Wn.vy * (nEy - nWy));
        const double v_j = 0.5 * (Wn.vx * (nNx - nSx) + Wn.vy * (nNy - nSy));
        const double a = W_a(Wn);
        dtn = dmin(qdiv(d_i, a + fabs(v_i)), qdiv(d_j, a + fabs(v_j)));
        if (p.lts) p.dt[c] = dtn;
    }
}

// Threads per workgroup for a TI x TJ tile: one per tile cell plus one per
// halo ("ring") cell whose reconstruction the tile's boundary faces need,
// rounded up to whole wavefronts.  For 16 x 16: 256 + 64 = 320 (5 waves),
// so every thread reconstructs exactly one cell.
template <int TI, int TJ> struct TileThreads {
    static constexpr int tile = TI * TJ, ring = 2 * (TI + TJ);
    // tiles of 256 cells or more: one thread per tile cell, and the ring
    // cells' reconstructions and ring-face fluxes in a second pass
    static constexpr bool ring_pass = tile >= 256;
    static constexpr int value = ring_pass ? ((tile + 63) / 64) * 64 : ((tile + ring + 63) / 64) * 64;
};

// Resident wavefronts per SIMD asked of the compiler: the mixed ring-pass
// tiles fit six in 80 VGPRs without spilling (0.5% faster at 32 x 8 than the
// five that 82 VGPRs allow); the FP64 kernels are left to the compiler.
template <typename FT, int TI, int TJ> struct StageWaves {
    static constexpr int value = (Fast<FT>::on && TileThreads<TI, TJ>::ring_pass) ? 6 : 1;
};

template <typename FT, int TI, int TJ, bool FUSED>
__global__ __launch_bounds__((TileThreads<TI, TJ>::value))
__attribute__((amdgpu_waves_per_eu(StageWaves<FT, TI, TJ>::value))) void stage_kernel(StageParams p) {
    constexpr int NT = TileThreads<TI, TJ>::value;
    constexpr int NTILE = TI * TJ;
    constexpr int LI = TI + 4, LJ = TJ + 4;
    constexpr int RI = TI + 2, RJ = TJ + 2;
    __shared__ FT sW[4][LJ][LI];
    typedef typename CoordT<FT>::type XT;
    __shared__ XT sX[LJ][LI], sY[LJ][LI];
    // phi*gradient (phases 2-3) and face fluxes (phases 3-4) share one LDS
    // buffer: the fluxes are stored only after every thread has read the
    // gradients, which keeps the workgroup's LDS small enough for more
    // resident workgroups per CU
    constexpr int NP = 2 * 4 * RJ * RI, NF = 4 * TJ * (TI + 1) + 4 * (TJ + 1) * TI;
    __shared__ FT sPF[NP > NF ? NP : NF];
    FT(*sPx)[RJ][RI] = reinterpret_cast<FT(*)[RJ][RI]>(sPF);
    FT(*sPy)[RJ][RI] = reinterpret_cast<FT(*)[RJ][RI]>(sPF + 4 * RJ * RI);
    // ring pass: the W and S ring-face fluxes, outside the shared buffer, so
    // that they can be stored as soon as they are computed
    constexpr bool RP = TileThreads<TI, TJ>::ring_pass;
    __shared__ FT sRing[RP ? 4 * (TI + TJ) : 1];
    FT(*sFi)[TJ][TI + 1] = reinterpret_cast<FT(*)[TJ][TI + 1]>(sPF);
    FT(*sFj)[TJ + 1][TI] = reinterpret_cast<FT(*)[TJ + 1][TI]>(sPF + 4 * TJ * (TI + 1));
    __shared__ double sRed[NT / 64][4];

    if (p.stop_flag && *p.stop_flag) return;
    const int tix = xcd_tile(blockIdx.x, p.ntiles);
    const TileDesc t = p.tiles[tix];
    const BlockMeta m = p.meta[t.b];
    if (!FUSED && m.skip) return;   // a converged case of a multigrid batch
    const int tid = threadIdx.x;
    const int64_t ncp = p.ncp;
    int phim = p.phi_mode;
    if (FUSED && p.part_state) {
        // batched steady run: a converged part is left alone, a frozen part
        // loads its limiters, a running part stores them at its last stage
        const int ps = p.part_state[t.part];
        if (ps == PART_DONE) {
            if (p.done_copy) {
                for (int q = tid; q < TI * TJ; q += NT) {
                    const int ip = t.i0 + 2 + q % TI, jp = t.j0 + 2 + q / TI;
                    if (ip <= m.ni + 1 && jp <= m.nj + 1) {
                        const int64_t c = m.coff + (int64_t)jp * m.pitch + ip;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            if constexpr (Fast<FT>::on) p.Woutf[v * ncp + c] = p.Winf[v * ncp + c];
                            else p.Wout[v * ncp + c] = p.Win[v * ncp + c];
                        }
                    }
                }
            }
            if (p.last_stage && tid == 0) {
                p.norm_part[tix] = 0.0;
                p.norm_part[p.ntiles + tix] = 0.0;
                p.norm_part[2 * p.ntiles + tix] = 0.0;
                p.norm_part[3 * p.ntiles + tix] = MILLION;
            }
            return;
        }
        phim = ps == PART_FROZEN ? PHI_LOAD : p.last_stage ? PHI_COMPUTE_STORE : PHI_COMPUTE;
    }
    if (!FUSED && p.part_frozen) phim = p.part_frozen[p.part_of_block[t.b]] ? PHI_LOAD : PHI_COMPUTE_STORE;
    const double *G = p.geom;
    const int ICu = m.ni + 1, JCu = m.nj + 1;
    const int pi0 = t.i0, pj0 = t.j0;   // padded index of the LDS window origin
    // the update waits (in update_tile_kernel) for residual smoothing, or
    // for the flux correction of a tile next to a finer neighbour
    const bool defer = FUSED && (p.defer_update || (p.tile_defer && p.tile_defer[tix]));
    const int32_t *bcW = p.bc + m.bcoff, *bcE = bcW + m.ncj, *bcS = bcE + m.ncj, *bcN = bcS + m.nci;
    // role of this thread: a tile cell (tc, tr), or a ring cell of side
    // 0 = W (column -1), 1 = E (column TI), 2 = S (row -1), 3 = N (row TJ)
    const bool is_tile = tid < NTILE;
    const int tc = tid % TI, tr = tid / TI;
    int ring_side = -1, ring_k = 0;
    if (!is_tile) {
        int r = tid - NTILE;
        if (r < TJ) { ring_side = 0; ring_k = r; }
        else if ((r -= TJ) < TJ) { ring_side = 1; ring_k = r; }
        else if ((r -= TJ) < TI) { ring_side = 2; ring_k = r; }
        else if ((r -= TI) < TI) { ring_side = 3; ring_k = r; }
    }

    // ---- phase 1: W and centroids of the window ---------------------------
    if constexpr (!RP) {
        // tiles with ring threads: every thread issues the loads of all its
        // window cells before its first LDS store, without branches.  A
        // thread past the window's end repeats the window's last cell, and
        // a window cell outside the block takes the nearest cell of the
        // block: such slots are never read (every reconstructed cell's
        // stencil lies inside its block), and a slot written twice gets the
        // same value twice.
        constexpr int NWIN = LI * LJ, NIT = (NWIN + NT - 1) / NT;
        FT wv[NIT][4];
        XT xw[NIT], yw[NIT];
        int slot[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int q = tid + k * NT, qq = q < NWIN ? q : NWIN - 1;
            slot[k] = qq;
            const int ip = pi0 + qq % LI, jp = pj0 + qq / LI;
            const int ipc = ip < m.nci ? ip : m.nci - 1, jpc = jp < m.ncj ? jp : m.ncj - 1;
            const int64_t c = m.coff + (int64_t)jpc * m.pitch + ipc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if constexpr (FUSED && Fast<FT>::on) wv[k][v] = p.Winf[v * ncp + c];
                else wv[k][v] = (FT)p.Win[v * ncp + c];
            }
            xw[k] = rdgeo<FT>(p, GEO_XC, GF_XC, c);
            yw[k] = rdgeo<FT>(p, GEO_YC, GF_YC, c);
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int il = slot[k] % LI, jl = slot[k] / LI;
#pragma unroll
            for (int v = 0; v < 4; ++v) sW[v][jl][il] = wv[k][v];
            sX[jl][il] = xw[k];
            sY[jl][il] = yw[k];
        }
    } else {
        for (int q = tid; q < LI * LJ; q += NT) {
            const int il = q % LI, jl = q / LI;
            const int ip = pi0 + il, jp = pj0 + jl;
            if (ip < m.nci && jp < m.ncj) {
                const int64_t c = m.coff + (int64_t)jp * m.pitch + ip;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if constexpr (FUSED && Fast<FT>::on) sW[v][jl][il] = p.Winf[v * ncp + c];
                    else sW[v][jl][il] = (FT)p.Win[v * ncp + c];
                }
                sX[jl][il] = rdgeo<FT>(p, GEO_XC, GF_XC, c);
                sY[jl][il] = rdgeo<FT>(p, GEO_YC, GF_YC, c);
            }
        }
    }
    __syncthreads();

    // ---- phase 2: least-squares reconstruction + limiter -------------------
    // One cell per thread: tile threads take their own cell, ring threads the
    // halo cell next to the tile.  Cells of blocks without physical
    // boundaries, and cells away from the ring next to a physical boundary,
    // use the full 8-point stencil; only the ring cells of boundary blocks
    // read the BC arrays (reduced stencils).
    for (int role = tid; role < (RP ? NTILE + 2 * (TI + TJ) : tid + 1); role += NT) {
        // role: a tile cell, then the W, E, S, N ring cells
        int ir = -1, jr = -1;
        if (RP) {
            int r = role - NTILE;
            if (r < 0) { ir = role % TI + 1; jr = role / TI + 1; }
            else if (r < TJ) { ir = 0; jr = r + 1; }
            else if ((r -= TJ) < TJ) { ir = TI + 1; jr = r + 1; }
            else if ((r -= TJ) < TI) { ir = r + 1; jr = 0; }
            else { r -= TI; ir = r + 1; jr = TJ + 1; }
        } else if (is_tile) { ir = tc + 1; jr = tr + 1; }
        else if (ring_side == 0) { ir = 0; jr = ring_k + 1; }
        else if (ring_side == 1) { ir = TI + 1; jr = ring_k + 1; }
        else if (ring_side == 2) { ir = ring_k + 1; jr = 0; }
        else if (ring_side == 3) { ir = ring_k + 1; jr = TJ + 1; }
        const int il = ir + 1, jl = jr + 1;
        const int ip = pi0 + il, jp = pj0 + jl;
        if (ir >= 0 && ip <= m.nci - 2 && jp <= m.ncj - 2) {
            const int64_t c = m.coff + (int64_t)jp * m.pitch + ip;
            unsigned mask = 0xFFu;
            if (m.has_bc && (ip <= 1 || ip >= ICu + 1 || jp <= 1 || jp >= JCu + 1))
                mask = stencil_mask(ip, jp, ICu, JCu, bcW[jp], bcE[jp], bcS[ip], bcN[ip]);
            PS<FT> gx, gy, ph;
            if (mask == 0xFFu) reconstruct<FT, true>(p, phim, c, ncp, m.pitch, il, jl, 0xFFu, sW, sX, sY, gx, gy, ph);
            else reconstruct<FT, false>(p, phim, c, ncp, m.pitch, il, jl, mask, sW, sX, sY, gx, gy, ph);
            if (phim == PHI_COMPUTE_STORE) {
                // each stored cell is owned by exactly one tile: its own
                // cells, plus the block's outer ring on the block edge
                const bool own_i = (ip >= pi0 + 2 && ip < pi0 + 2 + TI) || ip == 1 || ip == m.nci - 2;
                const bool own_j = (jp >= pj0 + 2 && jp < pj0 + 2 + TJ) || jp == 1 || jp == m.ncj - 2;
                if (own_i && own_j) {
                    p.phi[c] = ph.d; p.phi[ncp + c] = ph.vx; p.phi[2 * ncp + c] = ph.vy; p.phi[3 * ncp + c] = ph.p;
                }
            }
            // phi^dWdx, phi^dWdy (zero stencil: both products are 0*0)
            sPx[0][jr][ir] = ph.d * gx.d; sPx[1][jr][ir] = ph.vx * gx.vx;
            sPx[2][jr][ir] = ph.vy * gx.vy; sPx[3][jr][ir] = ph.p * gx.p;
            sPy[0][jr][ir] = ph.d * gy.d; sPy[1][jr][ir] = ph.vx * gy.vx;
            sPy[2][jr][ir] = ph.vy * gy.vy; sPy[3][jr][ir] = ph.p * gy.p;
        }
    }

    // Faces of phase 3: a tile thread (tc,tr) takes its E face (i-face
    // tc+1) and N face (j-face tr+1); W-ring thread k the i-face 0 of row k,
    // S-ring thread k the j-face 0 of column k (E/N-ring threads have none);
    // in the ring pass the first TI + TJ threads also take the W ring face
    // of row k = tid or the S ring face of column k = tid - TJ.
    // A face's normal and midpoint are read from the cell on its low side.
    struct FaceGeo { XT nx, ny, xf, yf; };
    auto face_cell = [&](bool idir, int f, int g, int64_t &cL) -> bool {
        const int ipL = idir ? pi0 + 1 + f : pi0 + 2 + g;
        const int jpL = idir ? pj0 + 2 + g : pj0 + 1 + f;
        cL = m.coff + (int64_t)jpL * m.pitch + ipL;
        return jpL <= JCu && ipL <= ICu;
    };
    auto face_geo = [&](bool idir, int64_t cL) -> FaceGeo {
        FaceGeo q;
        q.nx = idir ? rdgeo<FT>(p, GEO_NEX, GF_NEX, cL) : rdgeo<FT>(p, GEO_NNX, GF_NNX, cL);
        q.ny = idir ? rdgeo<FT>(p, GEO_NEY, GF_NEY, cL) : rdgeo<FT>(p, GEO_NNY, GF_NNY, cL);
        q.xf = idir ? rdgeo<FT>(p, GEO_XFE, GF_XFE, cL) : rdgeo<FT>(p, GEO_XFN, GF_XFN, cL);
        q.yf = idir ? rdgeo<FT>(p, GEO_YFE, GF_YFE, cL) : rdgeo<FT>(p, GEO_YFN, GF_YFN, cL);
        return q;
    };
    const bool ring_face = RP && tid < TI + TJ;
    const bool ring_w = tid < TJ;
    const int ring_fk = ring_w ? tid : tid - TJ;
    int64_t cF1 = 0, cF2 = 0, cFR = 0;
    bool v1 = false, v2 = false, vR = false;
    if (is_tile) {
        v1 = face_cell(true, tc + 1, tr, cF1);
        v2 = face_cell(false, tr + 1, tc, cF2);
    } else if (ring_side == 0) {
        v1 = face_cell(true, 0, ring_k, cF1);
    } else if (ring_side == 2) {
        v2 = face_cell(false, 0, ring_k, cF2);
    }
    if (ring_face) vR = face_cell(ring_w, 0, ring_fk, cFR);
    // issued before the barrier: the loads are in flight while the
    // workgroup's last reconstructions finish
    FaceGeo g1{}, g2{}, gR{};
    if (v1) g1 = face_geo(true, cF1);
    if (v2) g2 = face_geo(false, cF2);
    if (vR) gR = face_geo(ring_w, cFR);
    __syncthreads();

    // ---- phase 3: Roe fluxes ------------------------------------------------
    auto face = [&](bool idir, int f, int g, const FaceGeo &geo, CS<FT> &Fout) -> bool {
        // idir: i-face f in row g; else j-face f in column g.  L/R recon coords:
        int ipL, jpL, rl_i, rl_j, rr_i, rr_j;
        if (idir) {
            ipL = pi0 + 1 + f; jpL = pj0 + 2 + g;
            rl_i = f; rl_j = g + 1; rr_i = f + 1; rr_j = g + 1;
        } else {
            ipL = pi0 + 2 + g; jpL = pj0 + 1 + f;
            rl_i = g + 1; rl_j = f; rr_i = g + 1; rr_j = f + 1;
        }
        if (jpL > JCu || ipL > ICu) return false;
        const XT xf = geo.xf, yf = geo.yf;
        const FT nx = (FT)geo.nx, ny = (FT)geo.ny;
        auto load_state = [&](int ri, int rj, PS<FT> &W) {
            const int il = ri + 1, jl = rj + 1;
            W.d = sW[0][jl][il]; W.vx = sW[1][jl][il]; W.vy = sW[2][jl][il]; W.p = sW[3][jl][il];
            const PS<FT> px = {sPx[0][rj][ri], sPx[1][rj][ri], sPx[2][rj][ri], sPx[3][rj][ri]};
            const PS<FT> py = {sPy[0][rj][ri], sPy[1][rj][ri], sPy[2][rj][ri], sPy[3][rj][ri]};
            W = face_state(W, px, py, (FT)(xf - sX[jl][il]), (FT)(yf - sY[jl][il]));
        };
        PS<FT> Wl, Wr;
        int bcL = 0, bcR = 0, slotL, slotR;
        if (idir) {
            if (m.has_bc) {
                bcL = (ipL == 1) ? bcW[jpL] : 0;
                bcR = (ipL == ICu) ? bcE[jpL] : 0;
            }
            slotL = jpL; slotR = m.ncj + jpL;
        } else {
            if (m.has_bc) {
                bcL = (jpL == 1) ? bcS[ipL] : 0;
                bcR = (jpL == JCu) ? bcN[ipL] : 0;
            }
            slotL = 2 * m.ncj + ipL; slotR = 2 * m.ncj + m.nci + ipL;
        }
        if (face_bc(bcL)) {
            // boundary face on the low side: the outside state comes from the
            // inside face state, normal nfaceW(i+1) = -nfaceE(i) (resp. S, N)
            load_state(rr_i, rr_j, Wr);
            if (bcL == 7001) {
                Wl = reflect(Wr, -nx, -ny);
            } else {
                const double *wo = p.wo + 4 * (m.bcoff + slotL);
                const PS<FT> Wo = {(FT)wo[0], (FT)wo[1], (FT)wo[2], (FT)wo[3]};
                Wl = bc_characteristic_pressure(Wr, Wo, -nx, -ny);
            }
        } else if (face_bc(bcR)) {
            load_state(rl_i, rl_j, Wl);
            if (bcR == 7001) {
                Wr = reflect(Wl, nx, ny);
            } else {
                const double *wo = p.wo + 4 * (m.bcoff + slotR);
                const PS<FT> Wo = {(FT)wo[0], (FT)wo[1], (FT)wo[2], (FT)wo[3]};
                Wr = bc_characteristic_pressure(Wl, Wo, nx, ny);
            }
        } else {
            load_state(rl_i, rl_j, Wl);
            load_state(rr_i, rr_j, Wr);
        }
        if constexpr (!FUSED && !Fast<FT>::on) {
            Fout = p.hlle ? flux_hlle_n(Wl, Wr, nx, ny) : flux_roe_n(Wl, Wr, nx, ny);
        } else {
            Fout = flux_roe_n(Wl, Wr, nx, ny);
        }
        return true;
    };
    // up to two faces per thread: (idir, f, g) of each
    CS<FT> F1, F2;
    bool h1 = false, h2 = false;
    if (vR) {
        // ring pass: W ring face (row tid) or S ring face (column tid - TJ),
        // stored at once (first, so no flux is held while computing it)
        CS<FT> Fr;
        face(ring_w, 0, ring_fk, gR, Fr);
        FT *dst = ring_w ? sRing + ring_fk : sRing + 4 * TJ + ring_fk;
        const int st = ring_w ? TJ : TI;
        dst[0] = Fr.d; dst[st] = Fr.mx; dst[2 * st] = Fr.my; dst[3 * st] = Fr.E;
    }
    if (v1) h1 = face(true, is_tile ? tc + 1 : 0, is_tile ? tr : ring_k, g1, F1);
    if (v2) h2 = face(false, is_tile ? tr + 1 : 0, is_tile ? tc : ring_k, g2, F2);

    // phase-4 operands of the thread's cell: area, face lengths, time step,
    // U^n and the FP64 accumulator where the stage reads it.  Only this
    // thread writes U, S and dt of its cell in this launch.
    const int ip = pi0 + 2 + tc, jp = pj0 + 2 + tr;
    const bool valid = is_tile && (ip <= ICu) && (jp <= JCu);
    const int64_t c4 = m.coff + (int64_t)jp * m.pitch + ip;
    double A4 = 0.0, lW4 = 0.0, lE4 = 0.0, lS4 = 0.0, lN4 = 0.0, dt4 = 0.0;
    double U4[4] = {0.0, 0.0, 0.0, 0.0}, S4[4] = {0.0, 0.0, 0.0, 0.0};
    const bool fastS = Fast<FT>::on && p.Sf;
    const bool readS = FUSED && !fastS &&
                       (p.mode == MODE_FROM_S || (p.mode == MODE_ZERO_ACC2 && !defer && !p.skip_S));
    auto load4 = [&]() {
        A4 = G[GEO_A * ncp + c4];
        lW4 = rdgeo<FT>(p, GEO_LE, GF_LE, c4 - 1);
        lE4 = rdgeo<FT>(p, GEO_LE, GF_LE, c4);
        lS4 = rdgeo<FT>(p, GEO_LN, GF_LN, c4 - m.pitch);
        lN4 = rdgeo<FT>(p, GEO_LN, GF_LN, c4);
        dt4 = p.lts ? p.dt[c4] : bitsd(*p.dtmin_in);
        if (FUSED && !defer) {
#pragma unroll
            for (int v = 0; v < 4; ++v) U4[v] = p.U[v * ncp + c4];
        }
        if (readS) {
#pragma unroll
            for (int v = 0; v < 4; ++v) S4[v] = p.S[v * ncp + c4];
        }
    };
    // issued before the two barriers of the flux exchange, which they overlap
    if (valid) load4();
    __syncthreads();  // all gradient reads done: the flux arrays may now overwrite them
    if (h1) {
        const int f = is_tile ? tc + 1 : 0, g = is_tile ? tr : ring_k;
        sFi[0][g][f] = F1.d; sFi[1][g][f] = F1.mx; sFi[2][g][f] = F1.my; sFi[3][g][f] = F1.E;
    }
    if (h2) {
        const int f = is_tile ? tr + 1 : 0, g = is_tile ? tc : ring_k;
        sFj[0][f][g] = F2.d; sFj[1][f][g] = F2.mx; sFj[2][f][g] = F2.my; sFj[3][f][g] = F2.E;
    }
    __syncthreads();

    // ---- phase 4: accumulate (+W -E +S -N) and update -------------------------
    double l1 = 0.0, l2 = 0.0, mx = 0.0, dtn = MILLION;
    if (valid) {
        const int64_t c = c4;
        const double A = A4, lW = lW4, lE = lE4, lS = lS4, lN = lN4, dtc = dt4;
        const double cdt = p.cfl * dtc;
        double dU[4];
        if (FUSED) {
            if (p.mode == MODE_FROM_S && fastS) {
#pragma unroll
                for (int v = 0; v < 4; ++v) dU[v] = (double)p.Sf[v * ncp + c];
            } else if (p.mode == MODE_FROM_S) {
#pragma unroll
                for (int v = 0; v < 4; ++v) dU[v] = S4[v];
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) dU[v] = 0.0;
            }
        } else {
            if (p.mode == GINIT_KEEP) {
#pragma unroll
                for (int v = 0; v < 4; ++v) dU[v] = p.S[v * ncp + c];
            } else if (p.mode == GINIT_RK4) {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    dU[v] = p.S[v * ncp + c] + 2.0 * p.S1[v * ncp + c] + 2.0 * p.S2[v * ncp + c];
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) dU[v] = 0.0;
            }
        }
        // W and S face fluxes of this cell (from the ring array on the tile's
        // W column / S row in the ring pass)
        auto fW = [&](int v) -> FT {
            if constexpr (RP) return tc == 0 ? sRing[v * TJ + tr] : sFi[v][tr][tc];
            else return sFi[v][tr][tc];
        };
        auto fS = [&](int v) -> FT {
            if constexpr (RP) return tr == 0 ? sRing[4 * TJ + v * TI + tc] : sFj[v][tr][tc];
            else return sFj[v][tr][tc];
        };
        // axisymmetric source terms (Euler2DQuad.cc:1511-1516, 1016-1020):
        // S(W, Xc) of the cell's W^(s) and centroid radius
        double Sx[4] = {0.0, 0.0, 0.0, 0.0};
        const bool axi = p.axisym != 0;
        if (axi) {
            // the stage's W as the tile loaded it (the fused mixed path
            // keeps W^(s) in FP32 only)
            PS<double> Wc;
            if constexpr (FUSED && Fast<FT>::on) {
                Wc = {(double)p.Winf[c], (double)p.Winf[ncp + c], (double)p.Winf[2 * ncp + c],
                      (double)p.Winf[3 * ncp + c]};
            } else {
                Wc = {p.Win[c], p.Win[ncp + c], p.Win[2 * ncp + c], p.Win[3 * ncp + c]};
            }
            const CS<double> s = axisym_source(Wc, G[GEO_YC * ncp + c]);
            Sx[0] = s.d; Sx[1] = s.mx; Sx[2] = s.my; Sx[3] = s.E;
        }
        if constexpr (Fast<FT>::on) {
            // mixed path: net face flux in FP32, one FP64 scaling per cell
            const double s = cdt / A;
            const FT flW = (FT)lW, flE = (FT)lE, flS = (FT)lS, flN = (FT)lN;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const FT net = (fW(v) * flW - sFi[v][tr][tc + 1] * flE) +
                               (fS(v) * flS - sFj[v][tr + 1][tc] * flN);
                dU[v] += s * (double)net;
                if (axi) dU[v] += cdt * Sx[v];
            }
        } else {
            // ((CFL*dt)*F*l)/A per face, +W -E +S -N (Euler2DQuad.cc:1502-1509)
            bool res = false;
            if constexpr (!FUSED) res = p.residual != 0;
            // sixteen correctly rounded quotients over one reciprocal of A
            const Recip<double> rA = recip(A);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double FW = (double)fW(v), FE = (double)sFi[v][tr][tc + 1];
                const double FS = (double)fS(v), FN = (double)sFj[v][tr + 1][tc];
                double d = dU[v];
                if (res) {
                    // (Flux*l)/A of dUdt_Residual_Evaluation, Euler2DQuad.cc:1001-1006
                    d += rdiv(FW * lW, rA);
                    d -= rdiv(FE * lE, rA);
                    if (axi) d += Sx[v];
                    d += rdiv(FS * lS, rA);
                    d -= rdiv(FN * lN, rA);
                } else {
                    d += rdiv(cdt * FW * lW, rA);
                    d -= rdiv(cdt * FE * lE, rA);
                    if (axi) d += cdt * Sx[v];
                    d += rdiv(cdt * FS * lS, rA);
                    d -= rdiv(cdt * FN * lN, rA);
                }
                dU[v] = d;
            }
            if constexpr (!FUSED) {
                // FAS forcing term (FASMultigrid2D.h:2736-2739, 2840-2847)
                if (p.P) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) dU[v] += res ? p.P[v * ncp + c] : cdt * p.P[v * ncp + c];
                }
            }
        }
        if (m.amr & AMR_FLUX_ANY) {
            // block-edge fluxes FluxN/S/E/W = (+/-) Flux*lface (Euler2DQuad.cc:
            // 1518-1523, 1709-1715) for the resolution-change corrections
            double *eb = p.eflux + 4 * m.eoff;
            if ((m.amr & AMR_FLUX_N) && jp == JCu) {
                double *e = eb + 4 * (ip - 2);
#pragma unroll
                for (int v = 0; v < 4; ++v) e[v] = (double)sFj[v][tr + 1][tc] * lN;
            }
            if ((m.amr & AMR_FLUX_S) && jp == 2) {
                double *e = eb + 4 * (m.ni + ip - 2);
#pragma unroll
                for (int v = 0; v < 4; ++v) e[v] = (-(double)fS(v)) * lS;
            }
            if ((m.amr & AMR_FLUX_E) && ip == ICu) {
                double *e = eb + 4 * (2 * m.ni + jp - 2);
#pragma unroll
                for (int v = 0; v < 4; ++v) e[v] = (double)sFi[v][tr][tc + 1] * lE;
            }
            if ((m.amr & AMR_FLUX_W) && ip == 2) {
                double *e = eb + 4 * (2 * m.ni + m.nj + jp - 2);
#pragma unroll
                for (int v = 0; v < 4; ++v) e[v] = (-(double)fW(v)) * lW;
            }
        }
        if (!FUSED || defer) {
            // granular residual, or a correction / smoothing follows: store dUdt
            double *Sx = FUSED ? p.Sd : p.S;
#pragma unroll
            for (int v = 0; v < 4; ++v) Sx[v * ncp + c] = dU[v];
        } else {
            fused_update<FT>(p, m, t, c, ip, jp, A, dtc, lW, lE, lS, lN, U4, readS, S4, dU, l1, l2, mx, dtn);
        }
    }
    if (!(FUSED) || !p.last_stage || defer) return;

    // ---- last stage: workgroup partials of the norms and the next dt --------
    // wavefront reductions with DPP/permute shuffles, then across the 4 waves
    for (int off = 32; off > 0; off >>= 1) {
        l1 += __shfl_xor(l1, off);
        l2 += __shfl_xor(l2, off);
        mx = dmax(mx, __shfl_xor(mx, off));
        dtn = dmin(dtn, __shfl_xor(dtn, off));
    }
    const int wv = tid >> 6;
    if ((tid & 63) == 0) { sRed[wv][0] = l1; sRed[wv][1] = l2; sRed[wv][2] = mx; sRed[wv][3] = dtn; }
    __syncthreads();
    if (tid == 0) {
        double a = sRed[0][0], b = sRed[0][1], cmx = sRed[0][2], dmn = sRed[0][3];
        for (int w = 1; w < NT / 64; ++w) {
            a += sRed[w][0]; b += sRed[w][1]; cmx = dmax(cmx, sRed[w][2]); dmn = dmin(dmn, sRed[w][3]);
        }
        // per-tile partials (SoA), reduced in tile order by step_finish_kernel;
        // the kernel boundary orders these stores before that kernel's loads
        // (no device-scope fence or atomic per workgroup)
        p.norm_part[tix] = a;
        p.norm_part[p.ntiles + tix] = b;
        p.norm_part[2 * p.ntiles + tix] = cmx;
        p.norm_part[3 * p.ntiles + tix] = dmn;
    }
}

// Update of a tile from a stored residual: the second half of a fused stage
// when implicit residual smoothing runs between the residual and the update
// (Euler2DQuadSolvers.cc:855-877).  Same tile order, part handling and
// last-stage partials as the stage kernel.
constexpr int UPD_NT = 256;
template <typename FT>
__global__ __launch_bounds__(UPD_NT) void update_tile_kernel(StageParams p, int TI, int TJ) {
    __shared__ double sRed[UPD_NT / 64][4];
    if (p.stop_flag && *p.stop_flag) return;
    const int tix = p.tile_list ? p.tile_list[blockIdx.x] : xcd_tile(blockIdx.x, p.ntiles);
    const TileDesc t = p.tiles[tix];
    const BlockMeta m = p.meta[t.b];
    const int tid = threadIdx.x;
    const int64_t ncp = p.ncp;
    if (p.part_state && p.part_state[t.part] == PART_DONE) {
        if (p.done_copy) {
            for (int q = tid; q < TI * TJ; q += UPD_NT) {
                const int ip = t.i0 + 2 + q % TI, jp = t.j0 + 2 + q / TI;
                if (ip <= m.ni + 1 && jp <= m.nj + 1) {
                    const int64_t c = m.coff + (int64_t)jp * m.pitch + ip;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        if constexpr (Fast<FT>::on) p.Woutf[v * ncp + c] = p.Winf[v * ncp + c];
                        else p.Wout[v * ncp + c] = p.Win[v * ncp + c];
                    }
                }
            }
        }
        if (p.last_stage && tid == 0) {
            p.norm_part[tix] = 0.0;
            p.norm_part[p.ntiles + tix] = 0.0;
            p.norm_part[2 * p.ntiles + tix] = 0.0;
            p.norm_part[3 * p.ntiles + tix] = MILLION;
        }
        return;
    }
    double l1 = 0.0, l2 = 0.0, mx = 0.0, dtn = MILLION;
    for (int q = tid; q < TI * TJ; q += UPD_NT) {
        const int ip = t.i0 + 2 + q % TI, jp = t.j0 + 2 + q / TI;
        if (ip > m.ni + 1 || jp > m.nj + 1) continue;
        const int64_t c = m.coff + (int64_t)jp * m.pitch + ip;
        const double A = p.geom[GEO_A * ncp + c];
        const double dtc = p.lts ? p.dt[c] : bitsd(*p.dtmin_in);
        double dU[4] = {p.Sd[c], p.Sd[ncp + c], p.Sd[2 * ncp + c], p.Sd[3 * ncp + c]};
        double a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = MILLION;
        const double lW = rdgeo<FT>(p, GEO_LE, GF_LE, c - 1), lE = rdgeo<FT>(p, GEO_LE, GF_LE, c);
        const double lS = rdgeo<FT>(p, GEO_LN, GF_LN, c - m.pitch), lN = rdgeo<FT>(p, GEO_LN, GF_LN, c);
        const double U0[4] = {p.U[c], p.U[ncp + c], p.U[2 * ncp + c], p.U[3 * ncp + c]};
        fused_update<FT>(p, m, t, c, ip, jp, A, dtc, lW, lE, lS, lN, U0, false, U0, dU, a1, a2, a3, a4);
        l1 += a1; l2 += a2; mx = dmax(mx, a3); dtn = dmin(dtn, a4);
    }
    if (!p.last_stage) return;
    for (int off = 32; off > 0; off >>= 1) {
        l1 += __shfl_xor(l1, off);
        l2 += __shfl_xor(l2, off);
        mx = dmax(mx, __shfl_xor(mx, off));
        dtn = dmin(dtn, __shfl_xor(dtn, off));
    }
    const int wv = tid >> 6;
    if ((tid & 63) == 0) { sRed[wv][0] = l1; sRed[wv][1] = l2; sRed[wv][2] = mx; sRed[wv][3] = dtn; }
    __syncthreads();
    if (tid == 0) {
        double a = sRed[0][0], b = sRed[0][1], cmx = sRed[0][2], dmn = sRed[0][3];
        for (int w = 1; w < UPD_NT / 64; ++w) {
            a += sRed[w][0]; b += sRed[w][1]; cmx = dmax(cmx, sRed[w][2]); dmn = dmin(dmn, sRed[w][3]);
        }
        p.norm_part[tix] = a;
        p.norm_part[p.ntiles + tix] = b;
        p.norm_part[2 * p.ntiles + tix] = cmx;
        p.norm_part[3 * p.ntiles + tix] = dmn;
    }
}

}  // namespace cffc
