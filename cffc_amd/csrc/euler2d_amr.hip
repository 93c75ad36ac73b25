// euler2d_amr.hip -- resolution changes of quadtree block meshes on gfx950.
//
// Three small list-driven kernels, one thread per operation.  The
// restriction and prolongation run at the start of every stage, before the
// BCs (and repeatedly while the ghosts of a new state are primed); the flux
// correction runs between the residual and the update of the deferred tiles
// (a separate launch: inside the update kernel its load latency would add
// to every deferred tile's):
//   restrict_kernel   fine-to-coarse leg of Send_All_Messages: a coarse ghost
//                     cell takes the area-weighted average of the 2 x 2 fine
//                     cells it covers (LoadSendBuffer_F2C, Euler2DQuad.h:
//                     1597-1624), then W = W(U) (UnloadReceiveBuffer_F2C
//                     :1658-1684)
//   prolong_kernel    coarse-to-fine leg: the limited least-squares
//                     reconstruction of one coarse cell (SubcellReconstruction,
//                     Euler2DQuad.h:1721-1950, LIMITER_VENKATAKRISHNAN, the
//                     8-point stencil) evaluated at the centroids of its four
//                     quarters, sent as U (LoadSendBuffer_C2F, Euler2DQuad.cc:
//                     266-360); the fine ghost takes W = W(U)
//   flux_fix_kernel   Send_Conservative_Flux_Corrections (AdaptiveBlock2D_
//                     MessagePassing.h:8328) + Apply_Boundary_Flux_Corrections_
//                     Multistage_Explicit (Euler2DQuadSingleBlock.cc:4992-5080)
// The operation lists are built once per mesh by the host (capi.cpp).  The
// restriction runs before the prolongation: the coarse cells next to a fine
// block read the coarse ghosts the restriction fills, as in the reference,
// where the fine-to-coarse messages are unloaded before the coarse-to-fine
// ones are loaded (AdaptiveBlock2D_MessagePassing.h:7883-8020).  All
// arithmetic is FP64 and, compiled with -ffp-contract=off, follows the
// reference's operation order; the mixed path reads and writes FP32 W.
#include <hip/hip_runtime.h>

#include "euler2d_device.hpp"
#include "euler2d_kernels.hpp"
#include "euler2d_state.hpp"

namespace cffc {

template <typename ST> __device__ __forceinline__ PS<double> amr_ldW(const ST *W, int64_t ncp, int64_t c) {
    PS<double> r = {(double)W[c], (double)W[ncp + c], (double)W[2 * ncp + c], (double)W[3 * ncp + c]};
    return r;
}
template <typename ST> __device__ __forceinline__ void amr_stW(ST *W, int64_t ncp, int64_t c, const PS<double> &r) {
    W[c] = (ST)r.d; W[ncp + c] = (ST)r.vx; W[2 * ncp + c] = (ST)r.vy; W[3 * ncp + c] = (ST)r.p;
}

template <typename ST>
__global__ __launch_bounds__(64) void restrict_kernel(const RestrictOp *ops, int nops, const double *U,
                                                       const double *geom, int64_t ncp, ST *W,
                                                       const int32_t *stop_flag, int copy_from_U) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nops || (stop_flag && *stop_flag)) return;
    const RestrictOp o = ops[q];
    if (o.kind == 1) {
        // same-level ghost of a lattice mesh: the neighbour cell's W, from its
        // U where the reference transfers U (the first stage: W = W(U))
        if (copy_from_U) {
            const CS<double> Us = {U[o.src], U[ncp + o.src], U[2 * ncp + o.src], U[3 * ncp + o.src]};
            amr_stW(W, ncp, o.dst, W_of_U(Us));
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) W[v * ncp + o.dst] = W[v * ncp + o.src];
        }
        return;
    }
    const int64_t c00 = o.src, c10 = o.src + 1, c01 = o.src + o.pitch, c11 = o.src + o.pitch + 1;
    const double *A = geom + GEO_A * ncp;
    const double A00 = A[c00], A10 = A[c10], A01 = A[c01], A11 = A[c11];
    // four correctly rounded quotients over one reciprocal of the area sum
    const Recip<double> rden = recip(A00 + A10 + A01 + A11);
    double r[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double *u = U + v * ncp;
        r[v] = rdiv(A00 * u[c00] + A10 * u[c10] + A01 * u[c01] + A11 * u[c11], rden);
    }
    const CS<double> Ur = {r[0], r[1], r[2], r[3]};
    amr_stW(W, ncp, o.dst, W_of_U(Ur));
}

template <typename ST>
__global__ __launch_bounds__(64) void prolong_kernel(const ProlongOp *ops, int nops, const BlockMeta *meta,
                                                      const double *geom, const double *phi, int frozen,
                                                      const int32_t *part_state, int64_t ncp, ST *W,
                                                      const int32_t *stop_flag) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nops || (stop_flag && *stop_flag)) return;
    // limiters frozen by the caller, or by a one-part steady run
    if (part_state && part_state[0] != PART_RUNNING) frozen = 1;
    const ProlongOp o = ops[q];
    const int64_t c = o.src;
    const int64_t pitch = meta[o.blk].pitch;
    const double *G = geom;
    const double xc0 = G[GEO_XC * ncp + c], yc0 = G[GEO_YC * ncp + c];
    const PS<double> W0 = amr_ldW(W, ncp, c);
    // the 8-point stencil in the reference's order
    const int64_t nb[8] = {c - pitch - 1, c - pitch, c - pitch + 1, c - 1, c + 1, c + pitch - 1, c + pitch,
                           c + pitch + 1};
    double Dx[4] = {0, 0, 0, 0}, Dy[4] = {0, 0, 0, 0};
    double DxDx = 0.0, DxDy = 0.0, DyDy = 0.0;
    double wmin[4] = {W0.d, W0.vx, W0.vy, W0.p}, wmax[4] = {W0.d, W0.vx, W0.vy, W0.p};
    const double w0[4] = {W0.d, W0.vx, W0.vy, W0.p};
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) {
        const int64_t cn = nb[n2];
        const double dXx = G[GEO_XC * ncp + cn] - xc0, dXy = G[GEO_YC * ncp + cn] - yc0;
        const PS<double> Wn = amr_ldW(W, ncp, cn);
        const double wn[4] = {Wn.d, Wn.vx, Wn.vy, Wn.p};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double DU = wn[v] - w0[v];
            Dx[v] += DU * dXx;
            Dy[v] += DU * dXy;
            wmin[v] = wmin[v] < wn[v] ? wmin[v] : wn[v];   // min(u0Min, v), Math.h:73-92
            wmax[v] = wmax[v] > wn[v] ? wmax[v] : wn[v];
        }
        DxDx += dXx * dXx;
        DxDy += dXx * dXy;
        DyDy += dXy * dXy;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) { Dx[v] /= 8.0; Dy[v] /= 8.0; }
    DxDx /= 8.0; DxDy /= 8.0; DyDy /= 8.0;
    const Recip<double> rdet = recip(DxDx * DyDy - DxDy * DxDy);
    double gx[4], gy[4], ph[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        gx[v] = rdiv(Dx[v] * DyDy - Dy[v] * DxDy, rdet);
        gy[v] = rdiv(Dy[v] * DxDx - Dx[v] * DxDy, rdet);
    }
    if (frozen) {
#pragma unroll
        for (int v = 0; v < 4; ++v) ph[v] = phi[v * ncp + c];
    } else {
        // Limiter_Venkatakrishnan over the E, W, N, S face midpoints
        const double xE = G[GEO_XFE * ncp + c], yE = G[GEO_YFE * ncp + c];
        const double xW = G[GEO_XFE * ncp + c - 1], yW = G[GEO_YFE * ncp + c - 1];
        const double xN = G[GEO_XFN * ncp + c], yN = G[GEO_YFN * ncp + c];
        const double xS = G[GEO_XFN * ncp + c - pitch], yS = G[GEO_YFN * ncp + c - pitch];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double u0 = w0[v];
            double p = 1.0;
            p = venkat_fold<double>(p, u0 + gx[v] * (xE - xc0) + gy[v] * (yE - yc0), u0, wmin[v], wmax[v]);
            p = venkat_fold<double>(p, u0 + gx[v] * (xW - xc0) + gy[v] * (yW - yc0), u0, wmin[v], wmax[v]);
            p = venkat_fold<double>(p, u0 + gx[v] * (xN - xc0) + gy[v] * (yN - yc0), u0, wmin[v], wmax[v]);
            p = venkat_fold<double>(p, u0 + gx[v] * (xS - xc0) + gy[v] * (yS - yc0), u0, wmin[v], wmax[v]);
            ph[v] = p;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double dXx = o.xq[k] - xc0, dXy = o.yq[k] - yc0;
        double r[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = w0[v] + (ph[v] * gx[v]) * dXx + (ph[v] * gy[v]) * dXy;
        const PS<double> Wq = {r[0], r[1], r[2], r[3]};
        amr_stW(W, ncp, o.dst[k], W_of_U(U_of_W(Wq)));
    }
}

__global__ __launch_bounds__(256) void flux_fix_kernel(const FluxFixOp *ops, int nops, double *S,
                                                       const double *eflux, const double *geom, const double *dt,
                                                       const unsigned long long *dtmin, int lts, double cfl,
                                                       int64_t ncp, const int32_t *stop_flag) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nops || (stop_flag && *stop_flag)) return;
    const FluxFixOp o = ops[q];
    const int64_t c = o.cell;
    const double dtc = lts ? dt[c] : __longlong_as_double((long long)*dtmin);
    const double cdt = cfl * dtc;
    const Recip<double> rA = recip(geom[GEO_A * ncp + c]);
    double d[4] = {S[c], S[ncp + c], S[2 * ncp + c], S[3 * ncp + c]};
    for (int k = 0; k < o.n; ++k) {
        const double *own = eflux + 4 * (int64_t)o.own[k];
        const double *fa = eflux + 4 * (int64_t)o.fine[k];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            // FluxX = -(fine pair) - own; dUdt -= (CFL*dt)*FluxX/A
            const double F = -(fa[v] + fa[4 + v]) - own[v];
            d[v] -= rdiv(cdt * F, rA);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) S[v * ncp + c] = d[v];
}

}  // namespace cffc

using namespace cffc;

extern "C" hipError_t cffc_launch_restrict(const RestrictOp *ops, int nops, const double *U, const double *geom,
                                           int64_t ncp, double *W, float *Wf, const int32_t *stop_flag,
                                           int copy_from_U, hipStream_t st) {
    if (nops <= 0) return hipSuccess;
    // one wavefront per workgroup: short lists spread over more CUs
    const int nt = 64;
    const int nb = (nops + nt - 1) / nt;
    if (Wf)
        hipLaunchKernelGGL((restrict_kernel<float>), dim3(nb), dim3(nt), 0, st, ops, nops, U, geom, ncp, Wf,
                           stop_flag, copy_from_U);
    else
        hipLaunchKernelGGL((restrict_kernel<double>), dim3(nb), dim3(nt), 0, st, ops, nops, U, geom, ncp, W,
                           stop_flag, copy_from_U);
    return hipGetLastError();
}

extern "C" hipError_t cffc_launch_prolong(const ProlongOp *ops, int nops, const BlockMeta *meta, const double *geom,
                                          const double *phi, int frozen, const int32_t *part_state, int64_t ncp,
                                          double *W, float *Wf, const int32_t *stop_flag, hipStream_t st) {
    if (nops <= 0) return hipSuccess;
    const int nb = (nops + 63) / 64;
    if (Wf)
        hipLaunchKernelGGL((prolong_kernel<float>), dim3(nb), dim3(64), 0, st, ops, nops, meta, geom, phi, frozen,
                           part_state, ncp, Wf, stop_flag);
    else
        hipLaunchKernelGGL((prolong_kernel<double>), dim3(nb), dim3(64), 0, st, ops, nops, meta, geom, phi, frozen,
                           part_state, ncp, W, stop_flag);
    return hipGetLastError();
}

extern "C" hipError_t cffc_launch_flux_fix(const FluxFixOp *ops, int nops, double *S, const double *eflux,
                                           const double *geom, const double *dt, const unsigned long long *dtmin,
                                           int lts, double cfl, int64_t ncp, const int32_t *stop_flag,
                                           hipStream_t st) {
    if (nops <= 0) return hipSuccess;
    hipLaunchKernelGGL(flux_fix_kernel, dim3((nops + 255) / 256), dim3(256), 0, st, ops, nops, S, eflux, geom, dt,
                       dtmin, lts, cfl, ncp, stop_flag);
    return hipGetLastError();
}
