// unet_kernels.hip -- BF16 MFMA kernels of the warm-start U-Net (gfx950).
//
// The network has no counterpart in the reference: it predicts an initial
// state for the steady bump solves.  Activations are NHWC; weights and
// activations are BF16, accumulation and gradients FP32.
//
// 3x3 "same" convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   forward      Y[p,co]  = act( sum_k X[p+off(k), ci(k)] * Wt[co][k] + b[co] )
//                M = pixels, N = Cout, K = 9*Cin (k = tap*Cin + ci)
//   backward-in  dX[p,ci] = sum_k' G[p-off(k'), co(k')] * Wd[ci][k']
//                M = pixels, N = Cin,  K = 9*Cout (k' = tap*Cout + co)
//   backward-w   dW[co][k] = sum_p G[p,co] * X[p+off(k), ci(k)]
//                LDS-staged tiles, split over pixel chunks, FP32 atomics
// with G = dY * relu'(Y) formed on the fly.  Operand layout of the 16x16x32
// MFMA (wave64): lane l holds A[l%16][8*(l/16)+0..7], B[8*(l/16)+0..7][l%16]
// and D[4*(l/16)+0..3][l%16]; unet_mfma_layout_kernel checks it on the
// device against a host product.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;

namespace cffc_unet {

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (bf16_t)(u >> 16);  // inf/nan
    u += 0x7FFFu + ((u >> 16) & 1u);                                  // round to nearest even
    return (bf16_t)(u >> 16);
}

union Frag {
    bf16x8 v;
    uint4 u;
    bf16_t h[8];
};

__device__ __forceinline__ Frag zero_frag() {
    Frag f;
    f.u = make_uint4(0, 0, 0, 0);
    return f;
}

// ---------------------------------------------------------------------------
// layout self-test: D = A(16x32) * B(32x16) with A[m][k] = m + k/64,
// B[k][n] = (k == n) + (k == n + 16); each lane writes its 4 D values.
__global__ void mfma_layout_kernel(float *out) {
    const int l = threadIdx.x;
    Frag a, b;
    for (int i = 0; i < 8; ++i) {
        const int m = l % 16, k = 8 * (l / 16) + i;
        a.h[i] = f2bf((float)m + (float)k / 64.0f);
        const int kb = 8 * (l / 16) + i, n = l % 16;
        b.h[i] = f2bf((kb == n || kb == n + 16) ? 1.0f : 0.0f);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc, 0, 0, 0);
    for (int i = 0; i < 4; ++i) out[l * 4 + i] = acc[i];
}

// ---------------------------------------------------------------------------
// forward convolution.  Workgroup: 4 waves x 16 pixels = 64 pixels, NT
// output channels (NB = NT/16 MFMA column blocks per wave).
template <int NB>
__global__ __launch_bounds__(256) void conv3x3_fwd_kernel(const bf16_t *__restrict__ xA, int CA,
                                                          const bf16_t *__restrict__ xB, int CB, int B, int H,
                                                          int W, const bf16_t *__restrict__ wt, int Kpad,
                                                          const float *__restrict__ bias, int Cout, int relu,
                                                          bf16_t *__restrict__ y, float *__restrict__ y32) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Cin = CA + CB, K = 9 * Cin;
    const int P = B * H * W;
    const int p = blockIdx.x * 64 + wave * 16 + (lane & 15);
    const int n0 = blockIdx.y * (NB * 16);
    int pb = 0, py = 0, px = 0;
    const bool pvalid = p < P;
    if (pvalid) { px = p % W; py = (p / W) % H; pb = p / (W * H); }
    f32x4 acc[NB];
    for (int nb = 0; nb < NB; ++nb) acc[nb] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int kl = 8 * (lane >> 4);
    // k = tap*Cin + ci, stepped by 32 along with k (no division per step)
    int tap = kl / Cin, ci = kl - tap * Cin;
    for (int k0 = 0; k0 < Kpad; k0 += 32) {
        const int k = k0 + kl;
        Frag a = zero_frag();
        if (pvalid && k < K) {
            const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const int64_t q = ((int64_t)pb * H + yy) * W + xx;
                a.u = (ci < CA) ? *(const uint4 *)(xA + q * CA + ci) : *(const uint4 *)(xB + q * CB + (ci - CA));
            }
        }
        for (int nb = 0; nb < NB; ++nb) {
            const int n = n0 + nb * 16 + (lane & 15);
            Frag b;
            b.u = *(const uint4 *)(wt + (int64_t)n * Kpad + k);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[nb], 0, 0, 0);
        }
        ci += 32;
        while (ci >= Cin) { ci -= Cin; ++tap; }
    }
    const int pr0 = blockIdx.x * 64 + wave * 16 + 4 * (lane >> 4);
    for (int nb = 0; nb < NB; ++nb) {
        const int n = n0 + nb * 16 + (lane & 15);
        const float bv = bias[n];
        for (int i = 0; i < 4; ++i) {
            const int pp = pr0 + i;
            if (pp >= P) continue;
            float v = acc[nb][i] + bv;
            if (relu) v = v > 0.f ? v : 0.f;
            y[(int64_t)pp * Cout + n] = f2bf(v);
            if (y32) y32[(int64_t)pp * Cout + n] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// backward wrt the input.  G = dY * relu'(Y) (relu: Y > 0) is formed while
// loading; dX is split into the A part (channels < CA) and the B part.
template <int NB>
__global__ __launch_bounds__(256) void conv3x3_bwd_data_kernel(const float *__restrict__ dy,
                                                               const bf16_t *__restrict__ yact, int Cout, int relu,
                                                               int B, int H, int W, const bf16_t *__restrict__ wd,
                                                               int Kpad, int CA, int CB, float *__restrict__ dxA,
                                                               int accA, float *__restrict__ dxB, int accB) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int K = 9 * Cout;
    const int P = B * H * W;
    const int p = blockIdx.x * 64 + wave * 16 + (lane & 15);
    const int n0 = blockIdx.y * (NB * 16);
    int pb = 0, py = 0, px = 0;
    const bool pvalid = p < P;
    if (pvalid) { px = p % W; py = (p / W) % H; pb = p / (W * H); }
    f32x4 acc[NB];
    for (int nb = 0; nb < NB; ++nb) acc[nb] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int kl = 8 * (lane >> 4);
    // k = tap*Cout + co, stepped by 32 along with k (no division per step)
    int tap = kl / Cout, co = kl - tap * Cout;
    for (int k0 = 0; k0 < Kpad; k0 += 32) {
        const int k = k0 + kl;
        Frag a = zero_frag();
        if (pvalid && k < K) {
            // dX[p] gathers G at p - off(tap)
            const int yy = py - (tap / 3 - 1), xx = px - (tap % 3 - 1);
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const int64_t q = (((int64_t)pb * H + yy) * W + xx) * Cout + co;
                const float4 g0 = *(const float4 *)(dy + q), g1 = *(const float4 *)(dy + q + 4);
                float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
                if (relu) {
                    Frag ya;
                    ya.u = *(const uint4 *)(yact + q);
                    for (int i = 0; i < 8; ++i)
                        if (!(ya.h[i] != 0 && !(ya.h[i] & 0x8000u))) g[i] = 0.f;  // Y > 0
                }
                for (int i = 0; i < 8; ++i) a.h[i] = f2bf(g[i]);
            }
        }
        for (int nb = 0; nb < NB; ++nb) {
            const int n = n0 + nb * 16 + (lane & 15);
            Frag b;
            b.u = *(const uint4 *)(wd + (int64_t)n * Kpad + k);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[nb], 0, 0, 0);
        }
        co += 32;
        while (co >= Cout) { co -= Cout; ++tap; }
    }
    const int pr0 = blockIdx.x * 64 + wave * 16 + 4 * (lane >> 4);
    for (int nb = 0; nb < NB; ++nb) {
        const int n = n0 + nb * 16 + (lane & 15);
        for (int i = 0; i < 4; ++i) {
            const int pp = pr0 + i;
            if (pp >= P) continue;
            const float v = acc[nb][i];
            if (n < CA) {
                float *d = dxA + (int64_t)pp * CA + n;
                *d = accA ? *d + v : v;
            } else {
                float *d = dxB + (int64_t)pp * CB + (n - CA);
                *d = accB ? *d + v : v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// backward wrt the weights: dW[co][k] += sum over a pixel chunk of
// G[p,co] * X[p+off(k), ci(k)].  Workgroup tile: 64 output channels x 64 k,
// pixels in sub-chunks of 32 staged in LDS (G as [co][p], X as [k][p]).
// Staging walks co (resp. k) fastest across the threads, so the NHWC loads of
// dY, Y and X are contiguous; each thread keeps one co and one k for the whole
// chunk.  The workgroups of the first k tile also sum G in FP32 over their
// pixels for the bias gradient db[co] (one atomic per co and thread).
__global__ __launch_bounds__(256) void conv3x3_bwd_weight_kernel(const float *__restrict__ dy,
                                                                 const bf16_t *__restrict__ yact, int Cout,
                                                                 int relu, const bf16_t *__restrict__ xA, int CA,
                                                                 const bf16_t *__restrict__ xB, int CB, int B,
                                                                 int H, int W, float *__restrict__ dw,
                                                                 float *__restrict__ db, int pix_per_wg) {
    __shared__ bf16_t sG[64][32 + 8];
    __shared__ bf16_t sX[64][32 + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, tid = threadIdx.x;
    const int Cin = CA + CB, K = 9 * Cin;
    const int P = B * H * W;
    const int co0 = blockIdx.y * 64, k0 = blockIdx.z * 64;
    const int pstart = blockIdx.x * pix_per_wg;
    const int pend = min(P, pstart + pix_per_wg);
    // this thread's staging row: channel co = co0 + r of G, tap/ci of k = k0 + r
    const int r = tid & 63, pp0 = tid >> 6;
    const int co = co0 + r, k = k0 + r;
    const bool co_ok = co < Cout, k_ok = k < K;
    const int tap = k_ok ? k / Cin : 0, ci = k_ok ? k - tap * Cin : 0;
    const int dyy = tap / 3 - 1, dxx = tap % 3 - 1;
    const bool useA = ci < CA;
    const bf16_t *xsrc = useA ? xA : xB;
    const int xc = useA ? CA : CB, xci = useA ? ci : ci - CA;
    const bool bias = db != nullptr && blockIdx.z == 0;
    float bsum = 0.f;
    // wave w owns output rows co0 + 16*w .. +16 and all 64 k (4 blocks)
    f32x4 acc[4];
    for (int nb = 0; nb < 4; ++nb) acc[nb] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int ps = pstart; ps < pend; ps += 32) {
        // stage G[64 co][32 p] and X[64 k][32 p]; the pixel's (x, y, image)
        // is divided out once per sub-chunk, then stepped by 4
        int p = ps + pp0;
        int px = p % W, py = (p / W) % H, pb = p / (W * H);
        for (int pp = pp0; pp < 32; pp += 4, p += 4) {
            float g = 0.f;
            bf16_t xv = 0;
            if (p < pend) {
                if (co_ok) {
                    g = dy[(int64_t)p * Cout + co];
                    if (relu) {
                        const bf16_t ya = yact[(int64_t)p * Cout + co];
                        if (!(ya != 0 && !(ya & 0x8000u))) g = 0.f;
                    }
                }
                if (k_ok) {
                    const int yy = py + dyy, xx = px + dxx;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                        xv = xsrc[(((int64_t)pb * H + yy) * W + xx) * xc + xci];
                }
            }
            bsum += g;
            sG[r][pp] = f2bf(g);
            sX[r][pp] = xv;
            px += 4;
            while (px >= W) {
                px -= W;
                if (++py == H) { py = 0; ++pb; }
            }
        }
        __syncthreads();
        // A = G (16 co x 32 p), B = X^T (32 p x 16 k)
        Frag a;
        const int m = 16 * wave + (lane & 15), kk = 8 * (lane >> 4);
        for (int i = 0; i < 8; ++i) a.h[i] = sG[m][kk + i];
        for (int nb = 0; nb < 4; ++nb) {
            Frag b;
            const int n = nb * 16 + (lane & 15);
            for (int i = 0; i < 8; ++i) b.h[i] = sX[n][kk + i];
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[nb], 0, 0, 0);
        }
        __syncthreads();
    }
    if (bias && co_ok) atomicAdd(&db[co], bsum);
    for (int nb = 0; nb < 4; ++nb) {
        const int kq = k0 + nb * 16 + (lane & 15);
        for (int i = 0; i < 4; ++i) {
            const int cq = co0 + 16 * wave + 4 * (lane >> 4) + i;
            if (cq < Cout && kq < K) atomicAdd(&dw[(int64_t)cq * K + kq], acc[nb][i]);
        }
    }
}

// 2x2 average pooling (NHWC, bf16) and its backward (fp32, +=)
__global__ void avgpool2_fwd_kernel(const bf16_t *x, int B, int H, int W, int C, bf16_t *y) {
    const int64_t n = (int64_t)B * (H / 2) * (W / 2) * C;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = q % C;
        const int64_t r = q / C;
        const int xo = r % (W / 2), yo = (r / (W / 2)) % (H / 2), b = r / ((W / 2) * (H / 2));
        const int64_t base = (((int64_t)b * H + 2 * yo) * W + 2 * xo) * C + c;
        const float s = bf2f(x[base]) + bf2f(x[base + C]) + bf2f(x[base + (int64_t)W * C]) + bf2f(x[base + (int64_t)W * C + C]);
        y[q] = f2bf(0.25f * s);
    }
}
__global__ void avgpool2_bwd_kernel(const float *dy, int B, int H, int W, int C, float *dx) {
    const int64_t n = (int64_t)B * H * W * C;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = q % C;
        const int64_t r = q / C;
        const int x = r % W, y = (r / W) % H, b = r / (W * H);
        dx[q] += 0.25f * dy[(((int64_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    }
}
// nearest 2x upsampling and its backward (sum of the 4 children)
__global__ void upsample2_fwd_kernel(const bf16_t *x, int B, int H, int W, int C, bf16_t *y) {
    const int64_t n = (int64_t)B * (2 * H) * (2 * W) * C;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = q % C;
        const int64_t r = q / C;
        const int xo = r % (2 * W), yo = (r / (2 * W)) % (2 * H), b = r / (4 * (int64_t)W * H);
        y[q] = x[(((int64_t)b * H + yo / 2) * W + xo / 2) * C + c];
    }
}
__global__ void upsample2_bwd_kernel(const float *dy, int B, int H, int W, int C, float *dx) {
    const int64_t n = (int64_t)B * H * W * C;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = q % C;
        const int64_t r = q / C;
        const int x = r % W, y = (r / W) % H, b = r / (W * H);
        const int64_t W2 = 2 * W;
        const int64_t base = (((int64_t)b * 2 * H + 2 * y) * W2 + 2 * x) * C + c;
        dx[q] = dy[base] + dy[base + C] + dy[base + W2 * C] + dy[base + W2 * C + C];
    }
}

// MSE over the first nout of Cout output channels; dy = 2*(y-t)/N; the
// loss partial sums go to loss_acc (atomicAdd).
__global__ void mse_kernel(const float *y32, int Cout, int nout, const float *t, int64_t P, float *dy,
                           float *loss_acc) {
    __shared__ float sm[256];
    const float inv = 1.0f / (float)(P * nout);
    float s = 0.f;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < P * Cout; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = q % Cout;
        const int64_t p = q / Cout;
        float g = 0.f;
        if (c < nout) {
            const float d = y32[q] - t[p * nout + c];
            s += d * d;
            g = 2.f * d * inv;
        }
        dy[q] = g;
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_acc, sm[0] * inv);
}

// Adam on FP32 master weights [Cout][K]; refreshes the BF16 forward copy
// wt[Cout][Kpad] and the backward-input copy wd[Cin][Kpad2] (k' = tap*Cout+co).
__global__ void adam_kernel(float *w, const float *g, float *m, float *v, int Cout, int Cin, int Kpad, int Kpad2,
                            bf16_t *wt, bf16_t *wd, float lr, float b1, float b2, float eps, float c1, float c2,
                            int update) {
    const int K = 9 * Cin;
    const int64_t n = (int64_t)Cout * K;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        float wv = w[q];
        if (update) {
            const float gv = g[q];
            const float mv = b1 * m[q] + (1.f - b1) * gv;
            const float vv = b2 * v[q] + (1.f - b2) * gv * gv;
            m[q] = mv;
            v[q] = vv;
            wv -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
            w[q] = wv;
        }
        const int co = q / K, k = q % K;
        const int tap = k / Cin, ci = k % Cin;
        const bf16_t hb = f2bf(wv);
        wt[(int64_t)co * Kpad + k] = hb;
        wd[(int64_t)ci * Kpad2 + tap * Cout + co] = hb;
    }
}
__global__ void adam_bias_kernel(float *b, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                                 float eps, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gv = g[q];
    const float mv = b1 * m[q] + (1.f - b1) * gv;
    const float vv = b2 * v[q] + (1.f - b2) * gv * gv;
    m[q] = mv;
    v[q] = vv;
    b[q] -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
}

// gather a batch of samples: dst[i] = src[idx[i]] (rows of row_elems elements)
template <typename T>
__global__ void gather_rows_kernel(const T *src, const int32_t *idx, int nrows, int64_t row_elems, T *dst) {
    const int64_t n = (int64_t)nrows * row_elems;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int r = q / row_elems;
        dst[q] = src[(int64_t)idx[r] * row_elems + q % row_elems];
    }
}

}  // namespace cffc_unet

using namespace cffc_unet;

static int grid_1d(int64_t n) { return (int)((n + 255) / 256 > 16384 ? 16384 : (n + 255) / 256); }

extern "C" {

hipError_t cffc_unet_launch_mfma_layout(float *out, hipStream_t st) {
    hipLaunchKernelGGL(mfma_layout_kernel, dim3(1), dim3(64), 0, st, out);
    return hipGetLastError();
}

hipError_t cffc_unet_launch_conv_fwd(const bf16_t *xA, int CA, const bf16_t *xB, int CB, int B, int H, int W,
                                     const bf16_t *wt, int Kpad, const float *bias, int Cout, int relu, bf16_t *y,
                                     float *y32, hipStream_t st) {
    const int P = B * H * W;
    const int NT = Cout >= 64 ? 64 : Cout;
    dim3 grid((P + 63) / 64, Cout / NT);
    if (NT == 64) hipLaunchKernelGGL(conv3x3_fwd_kernel<4>, grid, dim3(256), 0, st, xA, CA, xB, CB, B, H, W, wt, Kpad, bias, Cout, relu, y, y32);
    else if (NT == 32) hipLaunchKernelGGL(conv3x3_fwd_kernel<2>, grid, dim3(256), 0, st, xA, CA, xB, CB, B, H, W, wt, Kpad, bias, Cout, relu, y, y32);
    else if (NT == 16) hipLaunchKernelGGL(conv3x3_fwd_kernel<1>, grid, dim3(256), 0, st, xA, CA, xB, CB, B, H, W, wt, Kpad, bias, Cout, relu, y, y32);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t cffc_unet_launch_conv_bwd_data(const float *dy, const bf16_t *yact, int Cout, int relu, int B, int H,
                                          int W, const bf16_t *wd, int Kpad2, int CA, int CB, float *dxA, int accA,
                                          float *dxB, int accB, hipStream_t st) {
    const int P = B * H * W;
    const int Cin = CA + CB;
    const int NT = (Cin % 64 == 0) ? 64 : (Cin % 32 == 0) ? 32 : 16;
    if (Cin % 16) return hipErrorInvalidValue;
    dim3 grid((P + 63) / 64, Cin / NT);
    if (NT == 64) hipLaunchKernelGGL(conv3x3_bwd_data_kernel<4>, grid, dim3(256), 0, st, dy, yact, Cout, relu, B, H, W, wd, Kpad2, CA, CB, dxA, accA, dxB, accB);
    else if (NT == 32) hipLaunchKernelGGL(conv3x3_bwd_data_kernel<2>, grid, dim3(256), 0, st, dy, yact, Cout, relu, B, H, W, wd, Kpad2, CA, CB, dxA, accA, dxB, accB);
    else hipLaunchKernelGGL(conv3x3_bwd_data_kernel<1>, grid, dim3(256), 0, st, dy, yact, Cout, relu, B, H, W, wd, Kpad2, CA, CB, dxA, accA, dxB, accB);
    return hipGetLastError();
}

hipError_t cffc_unet_launch_conv_bwd_weight(const float *dy, const bf16_t *yact, int Cout, int relu, const bf16_t *xA,
                                            int CA, const bf16_t *xB, int CB, int B, int H, int W, float *dw,
                                            float *db, hipStream_t st) {
    const int P = B * H * W;
    const int K = 9 * (CA + CB);
    // pixel chunk per workgroup: about 1024 workgroups over the (co, k) tiles,
    // chunks of 128..4096 pixels (whole 32-pixel sub-chunks)
    const int64_t tiles = (int64_t)((Cout + 63) / 64) * ((K + 63) / 64);
    const int pix = (int)std::min<int64_t>(4096, std::max<int64_t>(128, ((int64_t)P * tiles / 1024 + 31) / 32 * 32));
    hipError_t e = hipMemsetAsync(dw, 0, (size_t)Cout * K * sizeof(float), st);
    if (e == hipSuccess && db) e = hipMemsetAsync(db, 0, (size_t)Cout * sizeof(float), st);
    if (e != hipSuccess) return e;
    dim3 grid((P + pix - 1) / pix, (Cout + 63) / 64, (K + 63) / 64);
    hipLaunchKernelGGL(conv3x3_bwd_weight_kernel, grid, dim3(256), 0, st, dy, yact, Cout, relu, xA, CA, xB, CB, B, H, W,
                       dw, db, pix);
    return hipGetLastError();
}

hipError_t cffc_unet_launch_avgpool_fwd(const bf16_t *x, int B, int H, int W, int C, bf16_t *y, hipStream_t st) {
    hipLaunchKernelGGL(avgpool2_fwd_kernel, dim3(grid_1d((int64_t)B * H * W * C / 4)), dim3(256), 0, st, x, B, H, W, C, y);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_avgpool_bwd(const float *dy, int B, int H, int W, int C, float *dx, hipStream_t st) {
    hipLaunchKernelGGL(avgpool2_bwd_kernel, dim3(grid_1d((int64_t)B * H * W * C)), dim3(256), 0, st, dy, B, H, W, C, dx);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_upsample_fwd(const bf16_t *x, int B, int H, int W, int C, bf16_t *y, hipStream_t st) {
    hipLaunchKernelGGL(upsample2_fwd_kernel, dim3(grid_1d((int64_t)B * H * W * C * 4)), dim3(256), 0, st, x, B, H, W, C, y);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_upsample_bwd(const float *dy, int B, int H, int W, int C, float *dx, hipStream_t st) {
    hipLaunchKernelGGL(upsample2_bwd_kernel, dim3(grid_1d((int64_t)B * H * W * C)), dim3(256), 0, st, dy, B, H, W, C, dx);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_mse(const float *y32, int Cout, int nout, const float *t, int64_t P, float *dy,
                                float *loss, hipStream_t st) {
    // grid-stride: 512 workgroups at most, so at most 512 atomics meet on the loss
    hipLaunchKernelGGL(mse_kernel, dim3(std::min(grid_1d(P * Cout), 512)), dim3(256), 0, st, y32, Cout, nout, t, P, dy,
                       loss);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_adam(float *w, const float *g, float *m, float *v, int Cout, int Cin, int Kpad, int Kpad2,
                                 bf16_t *wt, bf16_t *wd, float lr, float b1, float b2, float eps, float c1, float c2,
                                 int update, float *b, const float *gb, float *mb, float *vb, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3(grid_1d((int64_t)Cout * 9 * Cin)), dim3(256), 0, st, w, g, m, v, Cout, Cin,
                       Kpad, Kpad2, wt, wd, lr, b1, b2, eps, c1, c2, update);
    if (update)
        hipLaunchKernelGGL(adam_bias_kernel, dim3((Cout + 255) / 256), dim3(256), 0, st, b, gb, mb, vb, Cout, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_gather_bf16(const bf16_t *src, const int32_t *idx, int nrows, int64_t row_elems,
                                        bf16_t *dst, hipStream_t st) {
    hipLaunchKernelGGL(gather_rows_kernel<bf16_t>, dim3(grid_1d(nrows * row_elems)), dim3(256), 0, st, src, idx, nrows, row_elems, dst);
    return hipGetLastError();
}
hipError_t cffc_unet_launch_gather_f32(const float *src, const int32_t *idx, int nrows, int64_t row_elems, float *dst,
                                       hipStream_t st) {
    hipLaunchKernelGGL(gather_rows_kernel<float>, dim3(grid_1d(nrows * row_elems)), dim3(256), 0, st, src, idx, nrows, row_elems, dst);
    return hipGetLastError();
}

}  // extern "C"
