// unet.cpp -- host runtime of the warm-start U-Net (include/cffc_unet.h).
//
// Four resolution levels (H x W down to H/8 x W/8), channels 16-32-64-128,
// two 3x3 conv + ReLU per level, 2x2 average pooling down, nearest 2x
// upsampling up, skip connections by channel concatenation (a conv reads its
// input from two NHWC tensors, so no copy is made), and a final 3x3 conv to
// 16 channels of which the first CFFC_UNET_OUT are the prediction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/cffc_euler2d.h"
#include "../../include/cffc_unet.h"

typedef uint16_t bf16_t;

extern "C" {
hipError_t cffc_unet_launch_mfma_layout(float *out, hipStream_t st);
hipError_t cffc_unet_launch_conv_fwd(const bf16_t *xA, int CA, const bf16_t *xB, int CB, int B, int H, int W,
                                     const bf16_t *wt, int Kpad, const float *bias, int Cout, int relu, bf16_t *y,
                                     float *y32, hipStream_t st);
hipError_t cffc_unet_launch_conv_bwd_data(const float *dy, const bf16_t *yact, int Cout, int relu, int B, int H,
                                          int W, const bf16_t *wd, int Kpad2, int CA, int CB, float *dxA, int accA,
                                          float *dxB, int accB, hipStream_t st);
hipError_t cffc_unet_launch_conv_bwd_weight(const float *dy, const bf16_t *yact, int Cout, int relu, const bf16_t *xA,
                                            int CA, const bf16_t *xB, int CB, int B, int H, int W, float *dw,
                                            float *db, hipStream_t st);
hipError_t cffc_unet_launch_avgpool_fwd(const bf16_t *x, int B, int H, int W, int C, bf16_t *y, hipStream_t st);
hipError_t cffc_unet_launch_avgpool_bwd(const float *dy, int B, int H, int W, int C, float *dx, hipStream_t st);
hipError_t cffc_unet_launch_upsample_fwd(const bf16_t *x, int B, int H, int W, int C, bf16_t *y, hipStream_t st);
hipError_t cffc_unet_launch_upsample_bwd(const float *dy, int B, int H, int W, int C, float *dx, hipStream_t st);
hipError_t cffc_unet_launch_mse(const float *y32, int Cout, int nout, const float *t, int64_t P, float *dy,
                                float *loss, hipStream_t st);
hipError_t cffc_unet_launch_adam(float *w, const float *g, float *m, float *v, int Cout, int Cin, int Kpad, int Kpad2,
                                 bf16_t *wt, bf16_t *wd, float lr, float b1, float b2, float eps, float c1, float c2,
                                 int update, float *b, const float *gb, float *mb, float *vb, hipStream_t st);
hipError_t cffc_unet_launch_gather_bf16(const bf16_t *src, const int32_t *idx, int nrows, int64_t row_elems,
                                        bf16_t *dst, hipStream_t st);
hipError_t cffc_unet_launch_gather_f32(const float *src, const int32_t *idx, int nrows, int64_t row_elems, float *dst,
                                       hipStream_t st);
}

static thread_local std::string u_err;
static int ufail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    u_err = buf;
    fprintf(stderr, "cffc_unet: %s\n", buf);
    return code;
}
#define UCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return ufail(CFFC_E_HIP, "%s: %s", #x, hipGetErrorString(e_));      \
    } while (0)

static inline bf16_t host_f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (bf16_t)(u >> 16);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

struct Tensor {
    int C, level;
    bf16_t *act = nullptr;
    float *grad = nullptr;
};

struct Conv {
    int inA, inB, out;  // tensor ids (inB = -1: none)
    int CA, CB, Cout, level, relu;
    int K, Kpad, Kpad2;
    float *w = nullptr, *b = nullptr, *m = nullptr, *v = nullptr, *mb = nullptr, *vb = nullptr;
    float *gw = nullptr, *gb = nullptr;
    bf16_t *wt = nullptr, *wd = nullptr;
};

enum OpKind { OP_CONV, OP_POOL, OP_UP };
struct Op {
    OpKind kind;
    int a, b;  // conv: layer index; pool/up: src tensor, dst tensor
};

struct cffc_unet {
    int H = 0, W = 0, maxB = 0;
    hipStream_t stream = nullptr;
    std::vector<Tensor> t;
    std::vector<Conv> L;
    std::vector<Op> ops;
    int x0 = -1, yout = -1;
    float *y32 = nullptr, *target = nullptr, *loss = nullptr;
    int32_t *idx = nullptr;
    bf16_t *ds_in = nullptr;
    float *ds_tg = nullptr;
    int ds_n = 0;
    int step = 0;
    std::vector<void *> allocs;
};

static int ualloc(cffc_unet *n, void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return ufail(CFFC_E_NOMEM, "hipMalloc(%zu) failed", bytes);
    n->allocs.push_back(*p);
    if (hipMemset(*p, 0, bytes) != hipSuccess) return ufail(CFFC_E_HIP, "hipMemset failed");
    return 0;
}

static int npix(const cffc_unet *n, int level, int B) { return B * (n->H >> level) * (n->W >> level); }

static int add_tensor(cffc_unet *n, int C, int level) {
    Tensor x;
    x.C = C;
    x.level = level;
    n->t.push_back(x);
    return (int)n->t.size() - 1;
}

static int add_conv(cffc_unet *n, int inA, int inB, int Cout, int relu) {
    Conv c;
    c.inA = inA;
    c.inB = inB;
    c.CA = n->t[inA].C;
    c.CB = inB >= 0 ? n->t[inB].C : 0;
    c.Cout = Cout;
    c.level = n->t[inA].level;
    c.relu = relu;
    c.out = add_tensor(n, Cout, c.level);
    c.K = 9 * (c.CA + c.CB);
    c.Kpad = (c.K + 31) / 32 * 32;
    c.Kpad2 = (9 * Cout + 31) / 32 * 32;
    n->L.push_back(c);
    n->ops.push_back(Op{OP_CONV, (int)n->L.size() - 1, 0});
    return c.out;
}

static int add_pool(cffc_unet *n, int src) {
    const int dst = add_tensor(n, n->t[src].C, n->t[src].level + 1);
    n->ops.push_back(Op{OP_POOL, src, dst});
    return dst;
}

static int add_up(cffc_unet *n, int src) {
    const int dst = add_tensor(n, n->t[src].C, n->t[src].level - 1);
    n->ops.push_back(Op{OP_UP, src, dst});
    return dst;
}

static void build_graph(cffc_unet *n) {
    const int ch[4] = {16, 32, 64, 128};
    n->x0 = add_tensor(n, CFFC_UNET_IN, 0);
    int cur = n->x0;
    int skip[3];
    for (int l = 0; l < 4; ++l) {
        cur = add_conv(n, cur, -1, ch[l], 1);
        cur = add_conv(n, cur, -1, ch[l], 1);
        if (l < 3) {
            skip[l] = cur;
            cur = add_pool(n, cur);
        }
    }
    for (int l = 2; l >= 0; --l) {
        cur = add_up(n, cur);
        cur = add_conv(n, cur, skip[l], ch[l], 1);
        cur = add_conv(n, cur, -1, ch[l], 1);
    }
    n->yout = add_conv(n, cur, -1, 16, 0);
}

static int upload_weights_bf16(cffc_unet *n, Conv &c) {
    UCHK(cffc_unet_launch_adam(c.w, c.gw, c.m, c.v, c.Cout, c.CA + c.CB, c.Kpad, c.Kpad2, c.wt, c.wd, 0.f, 0.f, 0.f,
                               0.f, 1.f, 1.f, 0, c.b, c.gb, c.mb, c.vb, n->stream));
    return 0;
}

extern "C" void cffc_unet_destroy(cffc_unet *n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void *p : n->allocs) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

extern "C" int cffc_unet_create(int H, int W, int max_batch, uint64_t seed, cffc_unet **out) {
    if (!out || H <= 0 || W <= 0 || H % 8 || W % 8 || max_batch <= 0) return ufail(CFFC_E_ARG, "bad network shape");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ufail(CFFC_E_NODEVICE, "no HIP device");
    cffc_unet *n = new cffc_unet();
    n->H = H;
    n->W = W;
    n->maxB = max_batch;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return ufail(CFFC_E_HIP, "stream creation failed");
    }
    build_graph(n);
    int e = 0;
    for (auto &x : n->t) {
        const size_t np = (size_t)npix(n, x.level, max_batch) * x.C;
        if ((e = ualloc(n, (void **)&x.act, np * sizeof(bf16_t))) || (e = ualloc(n, (void **)&x.grad, np * sizeof(float)))) {
            cffc_unet_destroy(n);
            return e;
        }
    }
    std::mt19937_64 rng(seed);
    for (size_t li = 0; li < n->L.size(); ++li) {
        Conv &c = n->L[li];
        const size_t nw = (size_t)c.Cout * c.K;
        const int Cin = c.CA + c.CB;
        if ((e = ualloc(n, (void **)&c.w, nw * 4)) || (e = ualloc(n, (void **)&c.m, nw * 4)) ||
            (e = ualloc(n, (void **)&c.v, nw * 4)) || (e = ualloc(n, (void **)&c.gw, nw * 4)) ||
            (e = ualloc(n, (void **)&c.b, c.Cout * 4)) || (e = ualloc(n, (void **)&c.mb, c.Cout * 4)) ||
            (e = ualloc(n, (void **)&c.vb, c.Cout * 4)) || (e = ualloc(n, (void **)&c.gb, c.Cout * 4)) ||
            (e = ualloc(n, (void **)&c.wt, (size_t)c.Cout * c.Kpad * 2)) ||
            (e = ualloc(n, (void **)&c.wd, (size_t)Cin * c.Kpad2 * 2))) {
            cffc_unet_destroy(n);
            return e;
        }
        // He-normal initialisation; the output layer starts near zero so
        // the untrained prediction is the freestream
        std::vector<float> hw(nw);
        const double sd = std::sqrt(2.0 / (9.0 * Cin)) * (li + 1 == n->L.size() ? 0.05 : 1.0);
        std::normal_distribution<double> nd(0.0, sd);
        for (auto &x : hw) x = (float)nd(rng);
        UCHK(hipMemcpy(c.w, hw.data(), nw * 4, hipMemcpyHostToDevice));
        if ((e = upload_weights_bf16(n, c))) {
            cffc_unet_destroy(n);
            return e;
        }
    }
    const size_t np0 = (size_t)npix(n, 0, max_batch);
    if ((e = ualloc(n, (void **)&n->y32, np0 * 16 * 4)) || (e = ualloc(n, (void **)&n->target, np0 * CFFC_UNET_OUT * 4)) ||
        (e = ualloc(n, (void **)&n->loss, 4)) || (e = ualloc(n, (void **)&n->idx, max_batch * 4))) {
        cffc_unet_destroy(n);
        return e;
    }
    UCHK(hipStreamSynchronize(n->stream));
    *out = n;
    return 0;
}

extern "C" int cffc_unet_num_layers(const cffc_unet *n) { return n ? (int)n->L.size() : 0; }
extern "C" int cffc_unet_layer_shape(const cffc_unet *n, int l, int shape[4]) {
    if (!n || l < 0 || l >= (int)n->L.size()) return ufail(CFFC_E_ARG, "bad layer");
    shape[0] = n->L[l].CA;
    shape[1] = n->L[l].CB;
    shape[2] = n->L[l].Cout;
    shape[3] = n->L[l].level;
    return 0;
}
extern "C" int cffc_unet_get_weights(cffc_unet *n, int l, float *w, float *b) {
    if (!n || l < 0 || l >= (int)n->L.size()) return ufail(CFFC_E_ARG, "bad layer");
    Conv &c = n->L[l];
    UCHK(hipStreamSynchronize(n->stream));
    if (w) UCHK(hipMemcpy(w, c.w, (size_t)c.Cout * c.K * 4, hipMemcpyDeviceToHost));
    if (b) UCHK(hipMemcpy(b, c.b, (size_t)c.Cout * 4, hipMemcpyDeviceToHost));
    return 0;
}
extern "C" int cffc_unet_set_weights(cffc_unet *n, int l, const float *w, const float *b) {
    if (!n || l < 0 || l >= (int)n->L.size()) return ufail(CFFC_E_ARG, "bad layer");
    Conv &c = n->L[l];
    if (w) UCHK(hipMemcpy(c.w, w, (size_t)c.Cout * c.K * 4, hipMemcpyHostToDevice));
    if (b) UCHK(hipMemcpy(c.b, b, (size_t)c.Cout * 4, hipMemcpyHostToDevice));
    if (int e = upload_weights_bf16(n, c)) return e;
    UCHK(hipStreamSynchronize(n->stream));
    return 0;
}

extern "C" int cffc_unet_set_dataset(cffc_unet *n, int ns, const float *inputs, const float *targets) {
    if (!n || ns <= 0 || !inputs || !targets) return ufail(CFFC_E_ARG, "bad dataset");
    const size_t per_in = (size_t)n->H * n->W * CFFC_UNET_IN, per_tg = (size_t)n->H * n->W * CFFC_UNET_OUT;
    std::vector<bf16_t> hin(per_in * ns);
    for (size_t q = 0; q < hin.size(); ++q) hin[q] = host_f2bf(inputs[q]);
    if (n->ds_in) {
        (void)hipFree(n->ds_in);
        (void)hipFree(n->ds_tg);
        n->allocs.erase(std::remove(n->allocs.begin(), n->allocs.end(), (void *)n->ds_in), n->allocs.end());
        n->allocs.erase(std::remove(n->allocs.begin(), n->allocs.end(), (void *)n->ds_tg), n->allocs.end());
    }
    int e;
    if ((e = ualloc(n, (void **)&n->ds_in, hin.size() * 2)) || (e = ualloc(n, (void **)&n->ds_tg, per_tg * ns * 4))) return e;
    UCHK(hipMemcpy(n->ds_in, hin.data(), hin.size() * 2, hipMemcpyHostToDevice));
    UCHK(hipMemcpy(n->ds_tg, targets, per_tg * ns * 4, hipMemcpyHostToDevice));
    n->ds_n = ns;
    return 0;
}

static int forward(cffc_unet *n, int B) {
    for (const Op &op : n->ops) {
        if (op.kind == OP_CONV) {
            Conv &c = n->L[op.a];
            const int H = n->H >> c.level, W = n->W >> c.level;
            const bf16_t *xb = c.inB >= 0 ? n->t[c.inB].act : nullptr;
            UCHK(cffc_unet_launch_conv_fwd(n->t[c.inA].act, c.CA, xb, c.CB, B, H, W, c.wt, c.Kpad, c.b, c.Cout, c.relu,
                                           n->t[c.out].act, c.out == n->yout ? n->y32 : nullptr, n->stream));
        } else if (op.kind == OP_POOL) {
            const Tensor &s = n->t[op.a];
            UCHK(cffc_unet_launch_avgpool_fwd(s.act, B, n->H >> s.level, n->W >> s.level, s.C, n->t[op.b].act, n->stream));
        } else {
            const Tensor &s = n->t[op.a];
            UCHK(cffc_unet_launch_upsample_fwd(s.act, B, n->H >> s.level, n->W >> s.level, s.C, n->t[op.b].act, n->stream));
        }
    }
    return 0;
}

static int backward(cffc_unet *n, int B) {
    for (int q = (int)n->ops.size() - 1; q >= 0; --q) {
        const Op &op = n->ops[q];
        if (op.kind == OP_CONV) {
            Conv &c = n->L[op.a];
            const int H = n->H >> c.level, W = n->W >> c.level;
            const bf16_t *xb = c.inB >= 0 ? n->t[c.inB].act : nullptr;
            const Tensor &o = n->t[c.out];
            UCHK(cffc_unet_launch_conv_bwd_weight(o.grad, o.act, c.Cout, c.relu, n->t[c.inA].act, c.CA, xb, c.CB, B, H, W,
                                                  c.gw, c.gb, n->stream));
            if (c.inA != n->x0) {
                float *dxb = c.inB >= 0 ? n->t[c.inB].grad : nullptr;
                UCHK(cffc_unet_launch_conv_bwd_data(o.grad, o.act, c.Cout, c.relu, B, H, W, c.wd, c.Kpad2, c.CA, c.CB,
                                                    n->t[c.inA].grad, 0, dxb, 0, n->stream));
            }
        } else if (op.kind == OP_POOL) {
            // the pooled tensor's source also feeds a skip connection whose
            // gradient the decoder wrote first: accumulate
            const Tensor &s = n->t[op.a];
            UCHK(cffc_unet_launch_avgpool_bwd(n->t[op.b].grad, B, n->H >> s.level, n->W >> s.level, s.C, s.grad, n->stream));
        } else {
            const Tensor &s = n->t[op.a];
            UCHK(cffc_unet_launch_upsample_bwd(n->t[op.b].grad, B, n->H >> s.level, n->W >> s.level, s.C, s.grad, n->stream));
        }
    }
    return 0;
}

static int load_batch(cffc_unet *n, const int32_t *idx, int B) {
    if (!n->ds_n) return ufail(CFFC_E_ARG, "no dataset");
    for (int i = 0; i < B; ++i)
        if (idx[i] < 0 || idx[i] >= n->ds_n) return ufail(CFFC_E_ARG, "sample index out of range");
    UCHK(hipMemcpyAsync(n->idx, idx, B * 4, hipMemcpyHostToDevice, n->stream));
    const int64_t px = (int64_t)n->H * n->W;
    UCHK(cffc_unet_launch_gather_bf16(n->ds_in, n->idx, B, px * CFFC_UNET_IN, n->t[n->x0].act, n->stream));
    UCHK(cffc_unet_launch_gather_f32(n->ds_tg, n->idx, B, px * CFFC_UNET_OUT, n->target, n->stream));
    return 0;
}

static int loss_and_grad(cffc_unet *n, int B, float *loss) {
    UCHK(hipMemsetAsync(n->loss, 0, 4, n->stream));
    UCHK(cffc_unet_launch_mse(n->y32, 16, CFFC_UNET_OUT, n->target, (int64_t)npix(n, 0, B), n->t[n->yout].grad, n->loss,
                              n->stream));
    if (int e = backward(n, B)) return e;
    if (loss) {
        UCHK(hipMemcpyAsync(loss, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
    }
    return 0;
}

extern "C" int cffc_unet_train_step(cffc_unet *n, const int32_t *idx, int B, float lr, float *loss) {
    if (!n || B <= 0 || B > n->maxB) return ufail(CFFC_E_ARG, "bad batch");
    if (int e = load_batch(n, idx, B)) return e;
    if (int e = forward(n, B)) return e;
    float l = 0.f;
    if (int e = loss_and_grad(n, B, &l)) return e;
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
    for (Conv &c : n->L)
        UCHK(cffc_unet_launch_adam(c.w, c.gw, c.m, c.v, c.Cout, c.CA + c.CB, c.Kpad, c.Kpad2, c.wt, c.wd, lr, b1, b2,
                                   1e-8f, c1, c2, 1, c.b, c.gb, c.mb, c.vb, n->stream));
    UCHK(hipStreamSynchronize(n->stream));
    if (loss) *loss = l;
    return 0;
}

extern "C" int cffc_unet_loss_grad(cffc_unet *n, const int32_t *idx, int B, float *loss, int l, float *gw, float *gb) {
    if (!n || B <= 0 || B > n->maxB) return ufail(CFFC_E_ARG, "bad batch");
    if (int e = load_batch(n, idx, B)) return e;
    if (int e = forward(n, B)) return e;
    float lv = 0.f;
    if (int e = loss_and_grad(n, B, &lv)) return e;
    UCHK(hipStreamSynchronize(n->stream));
    if (loss) *loss = lv;
    if (l >= 0 && l < (int)n->L.size()) {
        Conv &c = n->L[l];
        if (gw) UCHK(hipMemcpy(gw, c.gw, (size_t)c.Cout * c.K * 4, hipMemcpyDeviceToHost));
        if (gb) UCHK(hipMemcpy(gb, c.gb, (size_t)c.Cout * 4, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int cffc_unet_predict(cffc_unet *n, const float *inputs, int B, float *outputs) {
    if (!n || B <= 0 || B > n->maxB || !inputs || !outputs) return ufail(CFFC_E_ARG, "bad batch");
    const size_t px = (size_t)n->H * n->W;
    std::vector<bf16_t> hin(px * CFFC_UNET_IN * B);
    for (size_t q = 0; q < hin.size(); ++q) hin[q] = host_f2bf(inputs[q]);
    UCHK(hipMemcpyAsync(n->t[n->x0].act, hin.data(), hin.size() * 2, hipMemcpyHostToDevice, n->stream));
    if (int e = forward(n, B)) return e;
    std::vector<float> y(px * 16 * B);
    UCHK(hipMemcpyAsync(y.data(), n->y32, y.size() * 4, hipMemcpyDeviceToHost, n->stream));
    UCHK(hipStreamSynchronize(n->stream));
    for (size_t p = 0; p < px * B; ++p)
        for (int c = 0; c < CFFC_UNET_OUT; ++c) outputs[p * CFFC_UNET_OUT + c] = y[p * 16 + c];
    return 0;
}

extern "C" int cffc_unet_mfma_layout_probe(float out[256]) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ufail(CFFC_E_NODEVICE, "no HIP device");
    float *d = nullptr;
    UCHK(hipMalloc((void **)&d, 256 * 4));
    hipError_t e = cffc_unet_launch_mfma_layout(d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, d, 256 * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return ufail(CFFC_E_HIP, "probe: %s", hipGetErrorString(e));
    return 0;
}

extern "C" void *cffc_unet_stream(cffc_unet *n) { return n ? (void *)n->stream : nullptr; }
