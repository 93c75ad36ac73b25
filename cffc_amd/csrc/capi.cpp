// capi.cpp -- host runtime behind include/cffc_euler2d.h.
//
// Owns the device state of one multi-block mesh, orders the gfx950 kernels on
// one HIP stream and maps each C entry point onto the reference call it
// replaces (the comments in cffc_euler2d.h cite them).  No torch, no CPU
// fallback: without a HIP device cffc_create fails with CFFC_E_NODEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cffc_euler2d.h"
#include "euler2d_device.hpp"
#include "euler2d_kernels.hpp"

using namespace cffc;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(CFFC_E_HIP, "%s: %s", #x, hipGetErrorString(e_));    \
    } while (0)

static int resolve_scratch(int dev);

static const unsigned long long MILLION_BITS = 0x412E848000000000ull;
static const unsigned long long ERR_NONE = ~0ull;

struct cffc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    cffc_scheme_desc s{};
    int nblk = 0;
    int64_t ncp = 0, ncell = 0;
    int max_int = 0, max_cells = 0;
    int max_edge = 0;   // largest padded block edge (nci or ncj): sizes the exchange workgroups
    std::vector<BlockMeta> meta_h;
    BlockMeta *meta = nullptr;
    int32_t *nbr = nullptr, *bc = nullptr;
    double *wo = nullptr, *geom = nullptr;
    float *geomf = nullptr;   // FP32 geometry planes (mixed path only)
    double *W[2] = {nullptr, nullptr};
    float *Wf[2] = {nullptr, nullptr};   // fused mixed path: W stored as FP32
    int wcur = 0;
    double *U = nullptr, *Uo = nullptr, *dUdt[3] = {nullptr, nullptr, nullptr};
    double *dt = nullptr, *phi = nullptr;
    // tiles
    int ti = 16, tj = 16, ntiles = 0;
    TileDesc *tiles = nullptr;
    // scalars: [0],[1] dt slots, [2] error key
    unsigned long long *scal = nullptr;
    double *norm_part = nullptr, *norm_small = nullptr;
    unsigned int *counter = nullptr;
    int32_t *step_ctr = nullptr, *stop_flag = nullptr;
    double *norm_out = nullptr, *dt_out = nullptr;
    int max_steps = 0;
    double *staging = nullptr;
    // parts: independent cases of a batched steady run (contiguous blocks)
    int npart = 0;
    std::vector<int32_t> part_of_block;
    int32_t *part_tile0 = nullptr, *part_state = nullptr, *part_steps = nullptr;
    // a multigrid level of a batch with per-case limiter freezing (owned by
    // the cffc_mg): the part of every block, each part's frozen flag
    const int32_t *mg_part_of_block = nullptr, *mg_part_frozen = nullptr;
    double *part_l2 = nullptr, *part_last = nullptr;
    int part_l2_steps = 0;
    // fused mixed path: FP32 RK accumulator; sf_live: it, not dUdt[0],
    // holds the last fused run's residual
    float *Sf = nullptr;
    bool sf_live = false;
    // stall freeze (serving policy): a running part's best L2 and its step
    int stall_window = 0;
    double *part_best = nullptr;
    int32_t *part_best_step = nullptr;
    int steady_done = 0;   // steps of the current steady run before this segment
    unsigned long long *step_clock = nullptr;  // [max_steps + 1]: start, then end of every step
    double *chunk_part = nullptr;               // [4][256] first-level step reduction (large meshes)
    // push exchange: the fused stage kernel fills same-level ghosts; the
    // per-stage exchange launch then only applies BCs on boundary blocks
    bool push_ok = false;
    int nbc = 0;
    int32_t *bc_blocks = nullptr;
    // resolution changes (quadtree meshes): level-change legs of the exchange
    // and the conservative flux correction, as operation lists
    bool amr = false;
    int max_level = 0;
    std::vector<uint8_t> fine_faces;   // per block: bit f = face f has two finer neighbours
    RestrictOp *rops = nullptr;
    ProlongOp *pops = nullptr;
    FluxFixOp *fops = nullptr;
    int nrops = 0, npops = 0, nfops = 0;
    double *eflux = nullptr;           // saved block-edge fluxes
    double *Ustage = nullptr;          // fused: U^(s) of cells a coarser neighbour restricts
    uint8_t *ustore = nullptr;         // per padded cell: that cell is such a source
    double *Sd = nullptr;              // fused: residual of deferred tiles
    uint8_t *tile_defer = nullptr;
    int32_t *defer_list = nullptr;
    int ndefer = 0;
    double rs_eps = 0.0; // Residual_Smoothing_Epsilon
    int rs_iters = 0;    // Residual_Smoothing_Gauss_Seidel_Iterations (0: no smoothing)
    int freeze = 0;      // limiters frozen (PHI_LOAD)
    int phi_valid = 0;   // phi array holds the last evaluated limiters
    int store_phi = 0;   // keep phi stored every step (freeze may follow)
    bool granular_ready = false;
    // optional per-kernel timing with event pairs (non-graph runs)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;   // 0 = exchange, 1 = stage (one entry per pair)
    size_t ev_used = 0;
    double t_ms[2] = {0, 0};
    long long t_n[2] = {0, 0};
    bool own_stream = true;   // false: a multigrid level on its finest level's stream
};

// ---------------------------------------------------------------------------
static void free_ctx(cffc_ctx *c) {
    if (!c) return;
    void *ptrs[] = {c->meta, c->nbr, c->bc, c->wo, c->geom, c->geomf, c->W[0], c->W[1], c->U, c->Uo, c->dUdt[0],
                    c->dUdt[1], c->dUdt[2], c->dt, c->phi, c->tiles, c->scal, c->norm_part, c->norm_small,
                    c->counter, c->step_ctr, c->stop_flag, c->norm_out, c->dt_out, c->staging,
                    c->part_tile0, c->part_state, c->part_steps, c->part_l2, c->part_last, c->step_clock,
                    c->bc_blocks, c->Wf[0], c->Wf[1], c->chunk_part, c->rops, c->pops, c->fops, c->eflux,
                    c->Ustage, c->ustore, c->Sd, c->tile_defer, c->defer_list, c->part_best,
                    c->part_best_step, c->Sf};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

template <typename T> static int dalloc(T **p, size_t n) {
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return fail(CFFC_E_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
    return 0;
}

// Tiles in block order (so the tiles of a part are contiguous), each tagged
// with its block's part; part_tile0 gets the first tile of every part.
static int build_tiles(cffc_ctx *c, int ti, int tj) {
    std::vector<TileDesc> t;
    std::vector<int32_t> tile0;
    for (int b = 0; b < c->nblk; ++b) {
        const BlockMeta &m = c->meta_h[b];
        const int part = c->npart ? c->part_of_block[b] : 0;
        if (c->npart && (int)tile0.size() == part) tile0.push_back((int32_t)t.size());
        for (int j0 = 0; j0 < m.nj; j0 += tj)
            for (int i0 = 0; i0 < m.ni; i0 += ti) t.push_back(TileDesc{b, i0, j0, part});
    }
    if (c->tiles) (void)hipFree(c->tiles);
    if (c->norm_part) (void)hipFree(c->norm_part);
    if (c->part_tile0) (void)hipFree(c->part_tile0);
    c->tiles = nullptr;
    c->norm_part = nullptr;
    c->part_tile0 = nullptr;
    c->ntiles = (int)t.size();
    if (int e = dalloc(&c->tiles, t.size())) return e;
    if (int e = dalloc(&c->norm_part, 4 * t.size())) return e;
    if (t.size() >= 16384 && !c->chunk_part)
        if (int e = dalloc(&c->chunk_part, 4 * 256)) return e;
    if (t.size() < 16384 && c->chunk_part) {
        (void)hipFree(c->chunk_part);
        c->chunk_part = nullptr;
    }
    HIPCHK(hipMemcpy(c->tiles, t.data(), t.size() * sizeof(TileDesc), hipMemcpyHostToDevice));
    if (c->amr) {
        // tiles touching a face with finer neighbours update after the flux
        // correction of that face
        std::vector<uint8_t> defer(t.size(), 0);
        std::vector<int32_t> list;
        for (size_t q = 0; q < t.size(); ++q) {
            const TileDesc &d = t[q];
            const BlockMeta &m = c->meta_h[d.b];
            const int ff = c->fine_faces[d.b];
            const bool hit = ((ff & 1) && d.j0 + tj >= m.nj) || ((ff & 2) && d.j0 == 0) ||
                             ((ff & 4) && d.i0 + ti >= m.ni) || ((ff & 8) && d.i0 == 0);
            if (hit) {
                defer[q] = 1;
                list.push_back((int32_t)q);
            }
        }
        if (c->tile_defer) (void)hipFree(c->tile_defer);
        if (c->defer_list) (void)hipFree(c->defer_list);
        c->tile_defer = nullptr;
        c->defer_list = nullptr;
        c->ndefer = (int)list.size();
        if (int e = dalloc(&c->tile_defer, t.size())) return e;
        HIPCHK(hipMemcpy(c->tile_defer, defer.data(), t.size(), hipMemcpyHostToDevice));
        if (!list.empty()) {
            if (int e = dalloc(&c->defer_list, list.size())) return e;
            HIPCHK(hipMemcpy(c->defer_list, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        }
    }
    if (c->npart) {
        tile0.push_back((int32_t)t.size());
        if (int e = dalloc(&c->part_tile0, tile0.size())) return e;
        HIPCHK(hipMemcpy(c->part_tile0, tile0.data(), tile0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    c->ti = ti;
    c->tj = tj;
    return 0;
}

static bool tile_supported(int ti, int tj, bool fp64) {
#define T(A, B) if (ti == A && tj == B) return true;
    CFFC_TILE_LIST(T)
    if (fp64) {
        CFFC_SMALL_TILE_LIST(T)
    }
#undef T
    return false;
}

static int check_scheme(const cffc_scheme_desc *s) {
    const int ti = s->i_time_integration;
    if (ti != CFFC_TIME_STEPPING_EXPLICIT_EULER && ti != CFFC_TIME_STEPPING_EXPLICIT_PREDICTOR_CORRECTOR &&
        ti != CFFC_TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING && ti != CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA)
        return fail(CFFC_E_ARG, "unsupported i_time_integration %d", ti);
    if (s->n_stage < 1 || s->n_stage > 5) return fail(CFFC_E_ARG, "unsupported N_Stage %d", s->n_stage);
    if (s->i_reconstruction != CFFC_RECONSTRUCTION_LEAST_SQUARES)
        return fail(CFFC_E_ARG, "only least-squares reconstruction is implemented");
    if (s->i_flux_function != CFFC_FLUX_FUNCTION_ROE) return fail(CFFC_E_ARG, "only the Roe flux is implemented");
    if (s->i_limiter != CFFC_LIMITER_VENKATAKRISHNAN && s->i_limiter != CFFC_LIMITER_ONE &&
        s->i_limiter != CFFC_LIMITER_ZERO)
        return fail(CFFC_E_ARG, "unsupported limiter %d", s->i_limiter);
    if (s->local_time_stepping != CFFC_GLOBAL_TIME_STEPPING && s->local_time_stepping != CFFC_SCALAR_LOCAL_TIME_STEPPING)
        return fail(CFFC_E_ARG, "unsupported Local_Time_Stepping %d", s->local_time_stepping);
    if (s->flux_precision != CFFC_FLUX_FP64 && s->flux_precision != CFFC_FLUX_FP32)
        return fail(CFFC_E_ARG, "unsupported flux precision %d", s->flux_precision);
    if (s->axisymmetric != 0 && s->axisymmetric != 1) return fail(CFFC_E_ARG, "Axisymmetric must be 0 or 1");
    return 0;
}

// Runge_Kutta (CFD.cc:901-939) and MultiStage_Optimally_Smoothing (CFD.cc:953-1030)
static double runge_kutta(int i, int n) {
    double beta = 1.0;
    switch (n) {
    case 2: if (i == 2) beta = 0.5 * beta; break;
    case 4: if (i == 1 || i == 2) beta = 0.5 * beta; if (i == 4) beta = beta / 6.0; break;
    case 5: if (i == 1) beta *= 0.059; if (i == 2) beta *= 0.14; if (i == 3) beta *= 0.273; if (i == 4) beta *= 0.5; break;
    default: break;
    }
    return beta;
}
static double optimal_smoothing(int i, int n, int lim) {
    // van Leer, Tai & Powell (1989) coefficient table (first-order / limited)
    static const double base0[6] = {0, 1.0, 1.0, 1.50, 2.00, 2.50};
    static const double base1[6] = {0, 1.0, 0.4693, 0.6936, 0.9214, 1.1508};
    static const double a0[6][5] = {{0}, {0}, {0.3333}, {0.1481, 0.4000}, {0.0833, 0.2069, 0.4265},
                                    {0.0533, 0.1263, 0.2375, 0.4414}};
    static const double a1[6][5] = {{0}, {0}, {0.4242}, {0.1918, 0.4929}, {0.1084, 0.2602, 0.5052},
                                    {0.0695, 0.1602, 0.2898, 0.5060}};
    if (n < 1 || n > 5) return 1.0;
    double beta = (lim == CFFC_LIMITER_ZERO) ? base0[n] : base1[n];
    if (i < n) beta = ((lim == CFFC_LIMITER_ZERO) ? a0[n][i - 1] : a1[n][i - 1]) * beta;
    return beta;
}
static double stage_omega(const cffc_scheme_desc &s, int i) {
    if (s.i_time_integration == CFFC_TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING)
        return optimal_smoothing(i, s.n_stage, s.i_limiter);
    return runge_kutta(i, s.n_stage);
}
static int k_residual(const cffc_scheme_desc &s, int i) {
    if (s.i_time_integration == CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA && s.n_stage == 4) return i == 4 ? 0 : i - 1;
    return 0;
}
// init of the residual storage at stage i (Euler2DQuad.cc:1306-1343)
static int granular_init(const cffc_scheme_desc &s, int i) {
    if (i == 1) return GINIT_ZERO;
    if (s.i_time_integration == CFFC_TIME_STEPPING_EXPLICIT_PREDICTOR_CORRECTOR) return GINIT_KEEP;
    if (s.i_time_integration == CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA) {
        if (s.n_stage == 2) return GINIT_KEEP;
        if (s.n_stage == 4 && i == 4) return GINIT_RK4;
    }
    return GINIT_ZERO;
}
// the same, folded into the single accumulator of the fused path
static int fused_mode(const cffc_scheme_desc &s, int i) {
    const int g = granular_init(s, i);
    if (g == GINIT_KEEP || g == GINIT_RK4) return MODE_FROM_S;
    if (s.i_time_integration == CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA && s.n_stage == 4 && (i == 2 || i == 3))
        return MODE_ZERO_ACC2;
    return MODE_ZERO;
}


// ---- resolution changes -------------------------------------------------------
// quadCentroid (Math.h:311-350) of SW, SE, NE, NW and the centroid of quarter
// q (0 SW, 1 SE, 2 NW, 3 NE) of a cell (Grid2D_Quad_Block_HO::centroidSW/SE/
// NW/NE, HO_Grid2DQuad.h:1486-1660), in the reference's operation order
#pragma clang fp contract(off)
static void host_quad_centroid(const double *x, const double *y, double *xc, double *yc) {
    double ai, atmp = 0, xtmp = 0, ytmp = 0;
    ai = x[3] * y[0] - x[0] * y[3]; atmp += ai; xtmp += (x[0] + x[3]) * ai; ytmp += (y[0] + y[3]) * ai;
    ai = x[0] * y[1] - x[1] * y[0]; atmp += ai; xtmp += (x[1] + x[0]) * ai; ytmp += (y[1] + y[0]) * ai;
    ai = x[1] * y[2] - x[2] * y[1]; atmp += ai; xtmp += (x[2] + x[1]) * ai; ytmp += (y[2] + y[1]) * ai;
    ai = x[2] * y[3] - x[3] * y[2]; atmp += ai; xtmp += (x[3] + x[2]) * ai; ytmp += (y[3] + y[2]) * ai;
    if (atmp != 0) { *xc = xtmp / (3 * atmp); *yc = ytmp / (3 * atmp); }
    else { *xc = 0; *yc = 0; }
}
static void host_quarter_centroid(const cffc_mesh_desc *md, const BlockMeta &m, int i, int j, int q, double *xc,
                                  double *yc) {
    const int nn = m.nci + 1;
    auto X = [&](int a, int b) { return md->xnode[m.noff + (int64_t)b * nn + a]; };
    auto Y = [&](int a, int b) { return md->ynode[m.noff + (int64_t)b * nn + a]; };
    const double x00 = X(i, j), y00 = Y(i, j), x10 = X(i + 1, j), y10 = Y(i + 1, j);
    const double x01 = X(i, j + 1), y01 = Y(i, j + 1), x11 = X(i + 1, j + 1), y11 = Y(i + 1, j + 1);
    const double cx = 0.25 * (x00 + x10 + x01 + x11), cy = 0.25 * (y00 + y10 + y01 + y11);
    const double sx = 0.5 * (x00 + x10), sy = 0.5 * (y00 + y10), ex = 0.5 * (x10 + x11), ey = 0.5 * (y10 + y11);
    const double nx = 0.5 * (x01 + x11), ny = 0.5 * (y01 + y11), wx = 0.5 * (x00 + x01), wy = 0.5 * (y00 + y01);
    double Xq[4], Yq[4];
    switch (q) {
    case 0: Xq[0] = x00; Yq[0] = y00; Xq[1] = sx; Yq[1] = sy; Xq[2] = cx; Yq[2] = cy; Xq[3] = wx; Yq[3] = wy; break;
    case 1: Xq[0] = sx; Yq[0] = sy; Xq[1] = x10; Yq[1] = y10; Xq[2] = ex; Yq[2] = ey; Xq[3] = cx; Yq[3] = cy; break;
    case 2: Xq[0] = wx; Yq[0] = wy; Xq[1] = cx; Yq[1] = cy; Xq[2] = nx; Yq[2] = ny; Xq[3] = x01; Yq[3] = y01; break;
    default: Xq[0] = cx; Yq[0] = cy; Xq[1] = ex; Yq[1] = ey; Xq[2] = x11; Yq[2] = y11; Xq[3] = nx; Yq[3] = ny; break;
    }
    host_quad_centroid(Xq, Yq, xc, yc);
}
#pragma clang fp contract(on)

static int floor_half(int k) { return k >= 0 ? k / 2 : -((1 - k) / 2); }

// Levels, second face neighbours and the same-level table of the push /
// pull exchange; for a mesh with resolution changes, the restriction,
// prolongation and flux-correction lists, the per-block AMR flags and the
// edge-flux offsets.  Mirrors the reference's message rules
// (AdaptiveBlock2D_MessagePassing.h:1185-2398, 2399-5679, 8328).
static int setup_levels(cffc_ctx *c, const cffc_mesh_desc *md, std::vector<int32_t> &same, std::vector<RestrictOp> &rops,
                        std::vector<ProlongOp> &pops, std::vector<FluxFixOp> &fops, std::vector<uint8_t> &ustore,
                        int64_t &nentries) {
    const int nb = md->nblk;
    std::vector<int32_t> lev(nb, 0), n2(4 * (size_t)nb, -1);
    if (md->level)
        for (int b = 0; b < nb; ++b) {
            lev[b] = md->level[b];
            if (lev[b] < 0 || lev[b] > 30) return fail(CFFC_E_ARG, "block %d: level %d out of range", b, lev[b]);
        }
    if (md->nbr2)
        for (int q = 0; q < 4 * nb; ++q) {
            n2[q] = md->nbr2[q];
            if (n2[q] < -1 || n2[q] >= nb) return fail(CFFC_E_TOPOLOGY, "nbr2 entry %d out of range", q);
        }
    same.assign(md->nbr, md->nbr + 8 * (size_t)nb);
    c->amr = false;
    c->max_level = 0;
    c->fine_faces.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        c->max_level = std::max(c->max_level, lev[b]);
        for (int s = 0; s < 8; ++s) {
            const int o = md->nbr[8 * b + s];
            if (o < 0) continue;
            const int d = lev[o] - lev[b];
            if (d == 0) {
                if (s < 4 && n2[4 * b + s] >= 0)
                    return fail(CFFC_E_TOPOLOGY, "block %d face %d: second neighbour on a same-level face", b, s);
                continue;
            }
            if (d != 1 && d != -1) return fail(CFFC_E_TOPOLOGY, "blocks %d and %d differ by %d levels", b, o, d);
            const BlockMeta &m = c->meta_h[b], &n = c->meta_h[o];
            if (m.ni != n.ni || m.nj != n.nj || (m.ni & 1) || (m.nj & 1) || m.ni < 4 || m.nj < 4)
                return fail(CFFC_E_TOPOLOGY, "blocks %d and %d: a resolution change needs equal even ni, nj >= 4", b, o);
            same[8 * b + s] = -1;
            c->amr = true;
        }
    }
    if (!c->amr) return 0;
    if (c->s.n_stage < 1) return fail(CFFC_E_ARG, "bad scheme");
    static const int oppf[4] = {1, 0, 3, 2};
    auto cell = [&](int b, int i, int j) { const BlockMeta &m = c->meta_h[b]; return m.coff + (int64_t)j * m.pitch + i; };
    // sector of fine block b on its coarser face neighbour nbr[f]
    auto sector = [&](int b, int f) -> int {
        const int C = md->nbr[8 * b + f];
        if (md->nbr[8 * C + oppf[f]] == b) return 0;
        if (n2[4 * C + oppf[f]] == b) return 1;
        return -1;
    };
    std::vector<int32_t> amr(nb, 0);
    for (int b = 0; b < nb; ++b)
        for (int f = 0; f < 4; ++f) {
            const int o = md->nbr[8 * b + f];
            if (o < 0 || lev[o] == lev[b]) continue;
            if (lev[o] == lev[b] + 1) {
                const int o2 = n2[4 * b + f];
                if (o2 < 0 || lev[o2] != lev[b] + 1 || md->nbr[8 * o + oppf[f]] != b || md->nbr[8 * o2 + oppf[f]] != b)
                    return fail(CFFC_E_TOPOLOGY, "block %d face %d: finer neighbours %d, %d do not both face it", b, f, o, o2);
                c->fine_faces[b] |= (uint8_t)(1 << f);
                amr[b] |= AMR_FLUX_N << f;
                amr[o] |= AMR_FLUX_N << oppf[f];
                amr[o2] |= AMR_FLUX_N << oppf[f];
            } else if (sector(b, f) < 0) {
                return fail(CFFC_E_TOPOLOGY, "block %d face %d: coarser neighbour %d does not list it", b, f, o);
            }
        }
    // edge-flux entries: N [ni], S [ni], E [nj], W [nj] per flagged block
    nentries = 0;
    for (int b = 0; b < nb; ++b) {
        BlockMeta &m = c->meta_h[b];
        m.amr = amr[b];
        m.eoff = 0;
        if (amr[b] & AMR_FLUX_ANY) {
            m.eoff = nentries;
            nentries += 2 * (int64_t)m.ni + 2 * (int64_t)m.nj;
        }
    }
    ustore.assign((size_t)c->ncp, 0);
    auto src4 = [&](int F, int i, int j) {
        const BlockMeta &m = c->meta_h[F];
        const int64_t s0 = cell(F, i, j);
        ustore[s0] = ustore[s0 + 1] = ustore[s0 + m.pitch] = ustore[s0 + m.pitch + 1] = 1;
        c->meta_h[F].amr |= AMR_USTORE;
        return s0;
    };
    for (int b = 0; b < nb; ++b) {
        const BlockMeta m = c->meta_h[b];
        const int ICl = 2, ICu = m.ni + 1, JCl = 2, JCu = m.nj + 1, hi = m.ni / 2, hj = m.nj / 2;
        // restriction into b's ghosts from finer faces and corners
        for (int f = 0; f < 4; ++f) {
            const int o = md->nbr[8 * b + f];
            if (o < 0 || lev[o] != lev[b] + 1) continue;
            for (int g = 1; g <= 2; ++g) {
                if (f < 2) {
                    for (int i = ICl; i <= ICu; ++i) {
                        const int k = i - ICl, F = k < hi ? o : n2[4 * b + f];
                        const int i_f = 2 + 2 * (k % hi);
                        const int j_f = f == 0 ? 2 + 2 * (g - 1) : c->meta_h[F].nj + 2 - 2 * g;
                        rops.push_back(RestrictOp{cell(b, i, f == 0 ? JCu + g : JCl - g), src4(F, i_f, j_f),
                                                  c->meta_h[F].pitch, 0});
                    }
                } else {
                    for (int j = JCl; j <= JCu; ++j) {
                        const int k = j - JCl, F = k < hj ? o : n2[4 * b + f];
                        const int j_f = 2 + 2 * (k % hj);
                        const int i_f = f == 2 ? 2 + 2 * (g - 1) : c->meta_h[F].ni + 2 - 2 * g;
                        rops.push_back(RestrictOp{cell(b, f == 2 ? ICu + g : ICl - g, j), src4(F, i_f, j_f),
                                                  c->meta_h[F].pitch, 0});
                    }
                }
            }
        }
        for (int s = 4; s < 8; ++s) {
            const int F = md->nbr[8 * b + s];
            if (F < 0 || lev[F] != lev[b] + 1) continue;
            const bool east = s == 4 || s == 6, north = s == 4 || s == 5;
            const BlockMeta &fm = c->meta_h[F];
            for (int gj = 1; gj <= 2; ++gj)
                for (int gi = 1; gi <= 2; ++gi) {
                    const int i_f = east ? 2 + 2 * (gi - 1) : fm.ni + 2 - 2 * gi;
                    const int j_f = north ? 2 + 2 * (gj - 1) : fm.nj + 2 - 2 * gj;
                    rops.push_back(RestrictOp{cell(b, east ? ICu + gi : ICl - gi, north ? JCu + gj : JCl - gj),
                                              src4(F, i_f, j_f), fm.pitch, 0});
                }
        }
        // prolongation into b's ghosts from coarser faces (with the corner over
        // the middle of the coarse face) and the remaining coarser corners
        bool covered[8] = {false, false, false, false, false, false, false, false};
        int cover_blk[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
        for (int f = 0; f < 4; ++f) {
            const int C = md->nbr[8 * b + f];
            if (C < 0 || lev[C] != lev[b] - 1) continue;
            const int sig = sector(b, f);
            const BlockMeta &cm = c->meta_h[C];
            const int half = f < 2 ? m.ni / 2 : m.nj / 2;
            const int k0 = sig == 0 ? 0 : -1, k1 = sig == 0 ? half : half - 1;
            for (int kf = k0; kf <= k1; ++kf) {
                ProlongOp op{};
                int ic, jc;
                if (f < 2) { ic = 2 + sig * half + kf; jc = f == 0 ? 2 : cm.nj + 1; }
                else { ic = f == 2 ? 2 : cm.ni + 1; jc = 2 + sig * half + kf; }
                op.src = cell(C, ic, jc);
                op.blk = C;
                for (int q = 0; q < 4; ++q) {
                    const int qx = q & 1, qy = q >> 1;
                    int i, j;
                    if (f < 2) { i = ICl + 2 * kf + qx; j = f == 0 ? JCu + 1 + qy : JCl - 2 + qy; }
                    else { i = f == 2 ? ICu + 1 + qx : ICl - 2 + qx; j = JCl + 2 * kf + qy; }
                    op.dst[q] = cell(b, i, j);
                    host_quarter_centroid(md, cm, ic, jc, q, &op.xq[q], &op.yq[q]);
                }
                pops.push_back(op);
            }
            static const int cov[4][2] = {{4, 5}, {6, 7}, {4, 6}, {5, 7}};
            covered[cov[f][sig]] = true;
            cover_blk[cov[f][sig]] = C;
        }
        for (int s = 4; s < 8; ++s) {
            const int D = md->nbr[8 * b + s];
            if (covered[s]) {
                if (D >= 0 && D != cover_blk[s])
                    return fail(CFFC_E_TOPOLOGY, "block %d corner %d: lists block %d inside coarser block %d", b, s, D,
                                cover_blk[s]);
                same[8 * b + s] = -1;
                continue;
            }
            if (D < 0 || lev[D] != lev[b] - 1) continue;
            const bool east = s == 4 || s == 6, north = s == 4 || s == 5;
            const BlockMeta &dm = c->meta_h[D];
            const int ic = east ? 2 : dm.ni + 1, jc = north ? 2 : dm.nj + 1;
            const int i0 = east ? ICu + 1 : ICl - 2, j0 = north ? JCu + 1 : JCl - 2;
            ProlongOp op{};
            op.src = cell(D, ic, jc);
            op.blk = D;
            for (int q = 0; q < 4; ++q) {
                op.dst[q] = cell(b, i0 + (q & 1), j0 + (q >> 1));
                host_quarter_centroid(md, dm, ic, jc, q, &op.xq[q], &op.yq[q]);
            }
            pops.push_back(op);
        }
        // flux corrections of b's faces with finer neighbours, N, S, E, W
        std::vector<std::pair<int64_t, int>> at;   // cell -> op index of this block
        for (int f = 0; f < 4; ++f) {
            if (!(c->fine_faces[b] & (1 << f))) continue;
            const int n = f < 2 ? m.ni : m.nj, h = n / 2;
            static const int face_off_n[4] = {0, 1, 2, 2};   // N, S, E, W offsets: 0, ni, 2ni, 2ni+nj
            for (int q = 0; q < n; ++q) {
                const int F = q < h ? md->nbr[8 * b + f] : n2[4 * b + f];
                const BlockMeta &fm = c->meta_h[F];
                auto off = [&](const BlockMeta &x, int face) -> int64_t {
                    return x.eoff + (int64_t)face_off_n[face] * x.ni + (face == 3 ? x.nj : 0);
                };
                const int32_t own = (int32_t)(off(m, f) + q);
                const int32_t fine = (int32_t)(off(fm, oppf[f]) + 2 * (q % h));
                int64_t cc;
                if (f == 0) cc = cell(b, 2 + q, JCu);
                else if (f == 1) cc = cell(b, 2 + q, JCl);
                else if (f == 2) cc = cell(b, ICu, 2 + q);
                else cc = cell(b, ICl, 2 + q);
                int idx = -1;
                for (auto &pr : at)
                    if (pr.first == cc) idx = pr.second;
                if (idx < 0) {
                    idx = (int)fops.size();
                    fops.push_back(FluxFixOp{cc, 0, b, {0, 0}, {0, 0}});
                    at.push_back({cc, idx});
                }
                FluxFixOp &op = fops[idx];
                op.own[op.n] = own;
                op.fine[op.n] = fine;
                op.n += 1;
            }
        }
    }
    if (nentries > INT32_MAX) return fail(CFFC_E_ARG, "too many edge-flux entries");
    return 0;
}

// Blocks placed on per-level cell lattices (cffc_mesh_desc.lattice): blocks
// of any size, several per face.  Every ghost cell is filled from the block
// that holds its place, by the rule the reference applies between quadtree
// blocks of equal size (AdaptiveBlock2D_MessagePassing.h:32-1183 same level,
// 1185-2398 fine to coarse, 2399-5679 coarse to fine): a copy of the
// same-level cell, the restriction of the 2 x 2 finer cells, or the quarter
// of the coarser cell from its limited reconstruction; the BCs set ghost
// cells no block holds.  Coarse faces next to finer cells get the flux
// correction (MessagePassing.h:8328).  Same-level copies are RestrictOp
// entries of kind 1 in the restriction launch.
static int64_t floor_div(int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

static int setup_lattice(cffc_ctx *c, const cffc_mesh_desc *md, std::vector<int32_t> &same,
                         std::vector<RestrictOp> &rops, std::vector<ProlongOp> &pops, std::vector<FluxFixOp> &fops,
                         std::vector<uint8_t> &ustore, int64_t &nentries) {
    const int nb = md->nblk;
    if (!md->level) return fail(CFFC_E_ARG, "a lattice mesh needs block levels");
    std::vector<int32_t> lev(nb), I0(nb), J0(nb);
    int64_t g = 0;
    auto gcd = [](int64_t a, int64_t b) { while (b) { int64_t t = a % b; a = b; b = t; } return a < 0 ? -a : a; };
    c->max_level = 0;
    for (int b = 0; b < nb; ++b) {
        lev[b] = md->level[b];
        I0[b] = md->lattice[2 * b];
        J0[b] = md->lattice[2 * b + 1];
        if (lev[b] < 0 || lev[b] > 24) return fail(CFFC_E_ARG, "block %d: level %d out of range", b, lev[b]);
        const BlockMeta &m = c->meta_h[b];
        if ((m.ni & 1) || (m.nj & 1) || (I0[b] & 1) || (J0[b] & 1) || m.ni < 4 || m.nj < 4)
            return fail(CFFC_E_TOPOLOGY, "block %d: lattice blocks need even ni, nj >= 4 and even origins", b);
        g = gcd(g, gcd(gcd(m.ni, m.nj), gcd(I0[b], J0[b])));
        c->max_level = std::max(c->max_level, lev[b]);
    }
    if (g <= 0) g = 2;
    // index: squares of g x g cells of each level -> block
    auto key = [](int64_t l, int64_t x, int64_t y) {
        return ((uint64_t)l << 58) ^ ((uint64_t)(x + (1ll << 28)) << 29) ^ (uint64_t)(y + (1ll << 28));
    };
    std::unordered_map<uint64_t, int32_t> index;
    for (int b = 0; b < nb; ++b) {
        const BlockMeta &m = c->meta_h[b];
        for (int64_t y = J0[b] / g; y < (J0[b] + m.nj) / g; ++y)
            for (int64_t x = I0[b] / g; x < (I0[b] + m.ni) / g; ++x)
                if (!index.emplace(key(lev[b], x, y), b).second)
                    return fail(CFFC_E_TOPOLOGY, "block %d overlaps another block of level %d", b, lev[b]);
    }
    auto locate = [&](int64_t l, int64_t I, int64_t J) -> int {
        if (l < 0) return -1;
        auto it = index.find(key(l, floor_div(I, g), floor_div(J, g)));
        return it == index.end() ? -1 : it->second;
    };
    auto cell = [&](int b, int64_t i, int64_t j) { const BlockMeta &m = c->meta_h[b]; return m.coff + j * m.pitch + i; };
    same.assign(8 * (size_t)nb, -1);
    c->amr = true;
    c->fine_faces.assign(nb, 0);
    std::vector<int32_t> amr(nb, 0);
    // flux-correction faces first (they set the edge-flux flags and offsets)
    static const int oppf[4] = {1, 0, 3, 2};
    struct FixFace { int b, f, q, F; int64_t qf; };
    std::vector<FixFace> fixes;
    for (int b = 0; b < nb; ++b) {
        const BlockMeta &m = c->meta_h[b];
        const int L = lev[b];
        for (int f = 0; f < 4; ++f) {
            const int n = f < 2 ? m.ni : m.nj;
            for (int q = 0; q < n; ++q) {
                // the level-L cell across face f of boundary cell q
                int64_t I, J;
                if (f == 0) { I = I0[b] + q; J = J0[b] + m.nj; }
                else if (f == 1) { I = I0[b] + q; J = J0[b] - 1; }
                else if (f == 2) { I = I0[b] + m.ni; J = J0[b] + q; }
                else { I = I0[b] - 1; J = J0[b] + q; }
                if (locate(L, I, J) >= 0) continue;
                // the two finer cells along the face
                int64_t Ia = 2 * I, Ja = 2 * J, Ib, Jb;
                if (f == 0) { Ib = Ia + 1; Jb = Ja; }
                else if (f == 1) { Ja += 1; Ib = Ia + 1; Jb = Ja; }
                else if (f == 2) { Ib = Ia; Jb = Ja + 1; }
                else { Ia += 1; Ib = Ia; Jb = Ja + 1; }
                const int F = locate(L + 1, Ia, Ja);
                if (F < 0) continue;
                if (locate(L + 1, Ib, Jb) != F) return fail(CFFC_E_TOPOLOGY, "block %d face %d: finer cells split", b, f);
                const int64_t qf = (f < 2) ? Ia - I0[F] : Ja - J0[F];
                fixes.push_back(FixFace{b, f, q, F, qf});
                c->fine_faces[b] |= (uint8_t)(1 << f);
                amr[b] |= AMR_FLUX_N << f;
                amr[F] |= AMR_FLUX_N << oppf[f];
            }
        }
    }
    nentries = 0;
    for (int b = 0; b < nb; ++b) {
        BlockMeta &m = c->meta_h[b];
        m.amr = amr[b];
        m.eoff = 0;
        if (amr[b] & AMR_FLUX_ANY) {
            m.eoff = nentries;
            nentries += 2 * (int64_t)m.ni + 2 * (int64_t)m.nj;
        }
    }
    ustore.assign((size_t)c->ncp, 0);
    std::map<std::pair<int64_t, int>, int> pidx;   // (coarse cell, fine block) -> prolongation op
    for (int b = 0; b < nb; ++b) {
        const BlockMeta m = c->meta_h[b];
        const int L = lev[b];
        for (int j = 0; j < m.ncj; ++j)
            for (int i = 0; i < m.nci; ++i) {
                if (i >= 2 && i <= m.ni + 1 && j >= 2 && j <= m.nj + 1) continue;
                const int64_t I = I0[b] + i - 2, J = J0[b] + j - 2;
                const int64_t dst = cell(b, i, j);
                const int s = locate(L, I, J);
                if (s >= 0) {
                    rops.push_back(RestrictOp{dst, cell(s, I - I0[s] + 2, J - J0[s] + 2), 0, 1});
                    continue;
                }
                const int F = locate(L + 1, 2 * I, 2 * J);
                if (F >= 0) {
                    const BlockMeta &fm = c->meta_h[F];
                    const int64_t s0 = cell(F, 2 * I - I0[F] + 2, 2 * J - J0[F] + 2);
                    ustore[s0] = ustore[s0 + 1] = ustore[s0 + fm.pitch] = ustore[s0 + fm.pitch + 1] = 1;
                    c->meta_h[F].amr |= AMR_USTORE;
                    rops.push_back(RestrictOp{dst, s0, fm.pitch, 0});
                    continue;
                }
                const int C = locate(L - 1, floor_div(I, 2), floor_div(J, 2));
                if (C >= 0) {
                    const int64_t Ic = floor_div(I, 2), Jc = floor_div(J, 2);
                    const int q = (int)(I - 2 * Ic) + 2 * (int)(J - 2 * Jc);
                    const int ic = (int)(Ic - I0[C] + 2), jc = (int)(Jc - J0[C] + 2);
                    const int64_t src = cell(C, ic, jc);
                    auto it = pidx.find({src, b});
                    if (it == pidx.end()) {
                        ProlongOp op{};
                        op.src = src;
                        op.blk = C;
                        for (int k = 0; k < 4; ++k) {
                            op.dst[k] = -1;
                            host_quarter_centroid(md, c->meta_h[C], ic, jc, k, &op.xq[k], &op.yq[k]);
                        }
                        it = pidx.emplace(std::make_pair(src, b), (int)pops.size()).first;
                        pops.push_back(op);
                    }
                    pops[it->second].dst[q] = dst;
                    continue;
                }
                if (locate(L - 2, floor_div(I, 4), floor_div(J, 4)) >= 0 || locate(L + 2, 4 * I, 4 * J) >= 0)
                    return fail(CFFC_E_TOPOLOGY, "block %d: a ghost cell two levels away", b);
                // no block holds this ghost cell: a physical boundary, whose BC must set it
                const bool gi = i < 2 || i > m.ni + 1, gj = j < 2 || j > m.nj + 1;
                const int32_t *bcb = md->bc + m.bcoff;
                const int bcv = gi && !gj ? bcb[(i < 2 ? 0 : m.ncj) + j]
                               : gj && !gi ? bcb[2 * m.ncj + (j < 2 ? 0 : m.nci) + i] : -1;
                if (bcv == CFFC_BC_NONE)
                    return fail(CFFC_E_TOPOLOGY, "block %d ghost (%d, %d): no block and no boundary condition", b, i, j);
            }
    }
    for (const ProlongOp &op : pops)
        for (int k = 0; k < 4; ++k)
            if (op.dst[k] < 0) return fail(CFFC_E_TOPOLOGY, "block %d: a coarse cell's quarters split", op.blk);
    // flux corrections in the order N, S, E, W per cell
    std::map<int64_t, int> at;
    static const int face_off_n[4] = {0, 1, 2, 2};   // N, S, E, W entries at 0, ni, 2ni, 2ni+nj
    auto off = [&](const BlockMeta &x, int face) -> int64_t {
        return x.eoff + (int64_t)face_off_n[face] * x.ni + (face == 3 ? x.nj : 0);
    };
    for (const FixFace &fx : fixes) {
        const BlockMeta &m = c->meta_h[fx.b], &fm = c->meta_h[fx.F];
        int64_t cc;
        if (fx.f == 0) cc = cell(fx.b, 2 + fx.q, m.nj + 1);
        else if (fx.f == 1) cc = cell(fx.b, 2 + fx.q, 2);
        else if (fx.f == 2) cc = cell(fx.b, m.ni + 1, 2 + fx.q);
        else cc = cell(fx.b, 2, 2 + fx.q);
        auto it = at.find(cc);
        if (it == at.end()) {
            it = at.emplace(cc, (int)fops.size()).first;
            fops.push_back(FluxFixOp{cc, 0, fx.b, {0, 0}, {0, 0}});
        }
        FluxFixOp &op = fops[it->second];
        if (op.n >= 2) return fail(CFFC_E_TOPOLOGY, "block %d: a cell with three corrected faces", fx.b);
        op.own[op.n] = (int32_t)(off(m, fx.f) + fx.q);
        op.fine[op.n] = (int32_t)(off(fm, oppf[fx.f]) + fx.qf);
        op.n += 1;
    }
    if (nentries > INT32_MAX) return fail(CFFC_E_ARG, "too many edge-flux entries");
    return 0;
}

// Block metadata (padded layout, offsets) and the same-level topology and BC
// checks; host only.
static int build_meta(cffc_ctx *c, const cffc_mesh_desc *md, int32_t &bcoff, int64_t &noff) {
    int64_t coff = 0, ioff = 0;
    noff = 0;
    bcoff = 0;
    c->meta_h.resize(md->nblk);
    for (int b = 0; b < md->nblk; ++b) {
        BlockMeta &m = c->meta_h[b];
        m.ni = md->ni[b];
        m.nj = md->nj[b];
        if (m.ni < 2 || m.nj < 2 || m.ni > 4096 || m.nj > 4096) {
            return fail(CFFC_E_ARG, "block %d: ni, nj must be in [2, 4096]", b);
        }
        m.nci = m.ni + 4;
        m.ncj = m.nj + 4;
        // rows pitch cells apart (a multiple of 16 doubles = 128 bytes) and
        // the block placed so that interior cell (2, j) starts a line
        m.pitch = (m.nci + 15) & ~15;
        m.spare = 0;
        m.coff = ((coff + 15) & ~(int64_t)15) + 14;
        m.ioff = ioff;
        m.noff = noff;
        m.bcoff = bcoff;
        m.has_bc = 0;
        m.amr = 0;
        m.skip = 0;
        m.eoff = 0;
        for (int q = 0; q < 2 * (m.ni + 4) + 2 * (m.nj + 4); ++q)
            if (md->bc[bcoff + q] != CFFC_BC_NONE) m.has_bc = 1;
        coff = m.coff + (int64_t)m.pitch * m.ncj;
        ioff += (int64_t)m.ni * m.nj;
        noff += (int64_t)(m.nci + 1) * (m.ncj + 1);
        bcoff += 2 * m.ncj + 2 * m.nci;
        c->max_int = std::max(c->max_int, m.ni * m.nj);
        c->max_cells = std::max(c->max_cells, m.nci * m.ncj);
        c->max_edge = std::max(c->max_edge, std::max(m.nci, m.ncj));
    }
    for (int b = 0; b < md->nblk; ++b) {
        const BlockMeta &m = c->meta_h[b];
        for (int s = 0; s < 8; ++s) {
            const int nb = md->nbr[8 * b + s];
            if (nb < -1 || nb >= md->nblk) {
                return fail(CFFC_E_TOPOLOGY, "block %d slot %d: neighbour %d out of range", b, s, nb);
            }
            if (nb < 0) continue;
            const BlockMeta &n = c->meta_h[nb];
            if ((s == CFFC_NBR_N || s == CFFC_NBR_S) && n.ni != m.ni) {
                return fail(CFFC_E_TOPOLOGY, "block %d: N/S neighbour %d has different ni", b, nb);
            }
            if ((s == CFFC_NBR_E || s == CFFC_NBR_W) && n.nj != m.nj) {
                return fail(CFFC_E_TOPOLOGY, "block %d: E/W neighbour %d has different nj", b, nb);
            }
        }
        const int32_t *bcb = md->bc + m.bcoff;
        for (int q = 0; q < 2 * m.ncj + 2 * m.nci; ++q) {
            const int v = bcb[q];
            if (v != CFFC_BC_NONE && v != CFFC_BC_FIXED && v != CFFC_BC_CONSTANT_EXTRAPOLATION &&
                v != CFFC_BC_REFLECTION && v != CFFC_BC_PERIODIC && v != CFFC_BC_CHARACTERISTIC) {
                return fail(CFFC_E_ARG, "block %d: unsupported boundary condition %d", b, v);
            }
        }
    }
    c->ncp = (coff + 15) & ~(int64_t)15;   // every SoA plane starts 128-byte aligned
    c->ncell = ioff;
    return 0;
}

// ---------------------------------------------------------------------------
extern "C" const char *cffc_last_error(void) { return g_err.c_str(); }

extern "C" int cffc_create(const cffc_mesh_desc *md, const cffc_scheme_desc *sd, cffc_ctx **out) {
    if (!md || !sd || !out) return fail(CFFC_E_ARG, "null argument");
    *out = nullptr;
    if (md->nblk <= 0 || md->nghost != 2) return fail(CFFC_E_ARG, "nblk must be > 0 and nghost == 2");
    if (int e = check_scheme(sd)) return e;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(CFFC_E_NODEVICE, "no HIP device");
    cffc_ctx *c = new cffc_ctx();
    c->s = *sd;
    c->nblk = md->nblk;
    HIPCHK(hipGetDevice(&c->device));
    if (int e = resolve_scratch(c->device)) {
        free_ctx(c);
        return e;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        free_ctx(c);
        return fail(CFFC_E_HIP, "hipStreamCreate failed");
    }
    // block metadata and topology checks
    int32_t bcoff = 0;
    int64_t noff = 0;
    if (int ea = build_meta(c, md, bcoff, noff)) {
        free_ctx(c);
        return ea;
    }
    // levels: the same-level table (push / pull exchange) and, across
    // resolution changes, the restriction / prolongation / correction lists
    std::vector<int32_t> same;
    std::vector<RestrictOp> rops;
    std::vector<ProlongOp> pops;
    std::vector<FluxFixOp> fops;
    std::vector<uint8_t> ustore;
    int64_t nentries = 0;
    if (int ea = md->lattice ? setup_lattice(c, md, same, rops, pops, fops, ustore, nentries)
                             : setup_levels(c, md, same, rops, pops, fops, ustore, nentries)) {
        free_ctx(c);
        return ea;
    }
    {
        // the push exchange needs every same-level link to be mutual
        static const int opp[8] = {1, 0, 3, 2, 7, 6, 5, 4};
        bool sym = true;
        std::vector<int32_t> bl;
        for (int b = 0; b < md->nblk; ++b) {
            for (int s = 0; s < 8; ++s) {
                const int nb = same[8 * b + s];
                if (nb >= 0 && same[8 * nb + opp[s]] != b) sym = false;
            }
            if (c->meta_h[b].has_bc) bl.push_back(b);
        }
        c->push_ok = sym;
        if (c->amr && !sym) {
            free_ctx(c);
            return fail(CFFC_E_TOPOLOGY, "a mesh with resolution changes needs mutual same-level neighbours");
        }
        c->nbc = (int)bl.size();
        if (!bl.empty()) {
            if (int ea = dalloc(&c->bc_blocks, bl.size())) {
                free_ctx(c);
                return ea;
            }
            HIPCHK(hipMemcpy(c->bc_blocks, bl.data(), bl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        }
    }
    const int64_t ncp = c->ncp;
    int e = 0;
    if ((e = dalloc(&c->meta, md->nblk)) || (e = dalloc(&c->nbr, 8 * (size_t)md->nblk)) ||
        (e = dalloc(&c->bc, (size_t)bcoff)) || (e = dalloc(&c->wo, 4 * (size_t)bcoff)) ||
        (e = dalloc(&c->geom, GEO_COUNT * (size_t)ncp)) || (e = dalloc(&c->W[0], 4 * (size_t)ncp)) ||
        (e = dalloc(&c->W[1], 4 * (size_t)ncp)) || (e = dalloc(&c->U, 4 * (size_t)ncp)) ||
        (e = dalloc(&c->dUdt[0], 4 * (size_t)ncp)) || (e = dalloc(&c->dt, (size_t)ncp)) ||
        (e = dalloc(&c->scal, 4)) || (e = dalloc(&c->counter, 4)) || (e = dalloc(&c->step_ctr, 4)) ||
        (e = dalloc(&c->stop_flag, 4)) || (e = dalloc(&c->norm_small, 3))) {
        free_ctx(c);
        return e;
    }
    HIPCHK(hipMemcpy(c->meta, c->meta_h.data(), md->nblk * sizeof(BlockMeta), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->nbr, same.data(), 8 * md->nblk * sizeof(int32_t), hipMemcpyHostToDevice));
    if (c->amr) {
        c->nrops = (int)rops.size();
        c->npops = (int)pops.size();
        c->nfops = (int)fops.size();
        if ((c->nrops && (e = dalloc(&c->rops, rops.size()))) || (c->npops && (e = dalloc(&c->pops, pops.size()))) ||
            (c->nfops && (e = dalloc(&c->fops, fops.size()))) || (e = dalloc(&c->eflux, 4 * (size_t)nentries + 4)) ||
            (e = dalloc(&c->ustore, (size_t)ncp)) || (e = dalloc(&c->Ustage, 4 * (size_t)ncp)) ||
            (e = dalloc(&c->Sd, 4 * (size_t)ncp))) {
            free_ctx(c);
            return e;
        }
        if (c->nrops) HIPCHK(hipMemcpy(c->rops, rops.data(), rops.size() * sizeof(RestrictOp), hipMemcpyHostToDevice));
        if (c->npops) HIPCHK(hipMemcpy(c->pops, pops.data(), pops.size() * sizeof(ProlongOp), hipMemcpyHostToDevice));
        if (c->nfops) HIPCHK(hipMemcpy(c->fops, fops.data(), fops.size() * sizeof(FluxFixOp), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->ustore, ustore.data(), (size_t)ncp, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(c->eflux, 0, (4 * (size_t)nentries + 4) * sizeof(double)));
        HIPCHK(hipMemset(c->Ustage, 0, 4 * (size_t)ncp * sizeof(double)));
        HIPCHK(hipMemset(c->Sd, 0, 4 * (size_t)ncp * sizeof(double)));
        // meta_h carries the AMR flags and edge-flux offsets set above
        HIPCHK(hipMemcpy(c->meta, c->meta_h.data(), md->nblk * sizeof(BlockMeta), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->bc, md->bc, bcoff * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->wo, md->wo, 4 * (size_t)bcoff * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->W[0], 0, 4 * ncp * sizeof(double)));
    HIPCHK(hipMemset(c->W[1], 0, 4 * ncp * sizeof(double)));
    HIPCHK(hipMemset(c->U, 0, 4 * ncp * sizeof(double)));
    HIPCHK(hipMemset(c->dUdt[0], 0, 4 * ncp * sizeof(double)));
    HIPCHK(hipMemset(c->dt, 0, ncp * sizeof(double)));
    HIPCHK(hipMemset(c->counter, 0, 4 * sizeof(unsigned)));
    // geometry from the nodes
    double *xn = nullptr, *yn = nullptr;
    unsigned int *ndeg = nullptr;
    if ((e = dalloc(&xn, (size_t)noff)) || (e = dalloc(&yn, (size_t)noff)) || (e = dalloc(&ndeg, 1))) {
        if (xn) (void)hipFree(xn);
        if (yn) (void)hipFree(yn);
        free_ctx(c);
        return e;
    }
    HIPCHK(hipMemcpy(xn, md->xnode, noff * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(yn, md->ynode, noff * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ndeg, 0, sizeof(unsigned)));
    HIPCHK(cffc_launch_geometry(c->meta, c->nblk, c->max_cells, xn, yn, c->geom, ncp, ndeg, c->stream));
    unsigned nd = 0;
    HIPCHK(hipMemcpyAsync(&nd, ndeg, sizeof nd, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(xn);
    (void)hipFree(yn);
    (void)hipFree(ndeg);
    if (nd) {
        free_ctx(c);
        return fail(CFFC_E_ARG, "mesh has %u degenerate faces (length <= 1e-9 x cell perimeter)", nd);
    }
    if (c->s.axisymmetric) {
        // the source divides by the centroid radius: every interior cell
        // must lie off the axis (y > 0)
        std::vector<double> yc(ncp);
        HIPCHK(hipMemcpy(yc.data(), c->geom + (size_t)GEO_YC * ncp, ncp * sizeof(double), hipMemcpyDeviceToHost));
        for (int b = 0; b < c->nblk; ++b) {
            const BlockMeta &m = c->meta_h[b];
            for (int j = 2; j < m.nj + 2; ++j)
                for (int i = 2; i < m.ni + 2; ++i)
                    if (!(yc[m.coff + (int64_t)j * m.pitch + i] > 0.0)) {
                        free_ctx(c);
                        return fail(CFFC_E_ARG, "axisymmetric: block %d cell (%d, %d) has centroid y <= 0", b, i - 2,
                                    j - 2);
                    }
        }
    }
    if (c->s.flux_precision == CFFC_FLUX_FP32) {
        if ((e = dalloc(&c->geomf, GF_COUNT * (size_t)ncp))) {
            free_ctx(c);
            return e;
        }
        HIPCHK(cffc_launch_geometry_f32(c->meta, c->nblk, c->max_cells, c->geom, c->geomf, ncp, c->stream));
        for (int k = 0; k < 2; ++k)
            if ((e = dalloc(&c->Wf[k], 4 * (size_t)ncp))) {
                free_ctx(c);
                return e;
            }
        if ((e = dalloc(&c->Sf, 4 * (size_t)ncp))) {
            free_ctx(c);
            return e;
        }
    }
    // default tile: for blocks at least 32 cells wide the autotuner's picks
    // on the 16.8 M-cell and uneven quadtree workloads, 32 x 16 (FP64) and
    // 32 x 8 (mixed), both ring-pass tiles; 16 x 8 for blocks 16 to 31
    // cells wide (192 threads, three wavefronts); 8 x 8 for blocks narrower
    // than 16 cells; in FP64, 4 x 4 (one wavefront per block) for blocks
    // narrower than 8 cells
    const bool fp64 = c->s.flux_precision == CFFC_FLUX_FP64;
    int ti = 16, tj = 8;
    int mini = 1 << 30;
    for (auto &m : c->meta_h) mini = std::min(mini, m.ni);
    if (mini >= 32) { ti = 32; tj = fp64 ? 16 : 8; }
    if (mini < 16) ti = tj = 8;
    if (mini < 8 && fp64) ti = tj = 4;
    if ((e = build_tiles(c, ti, tj))) {
        free_ctx(c);
        return e;
    }
    *out = c;
    return 0;
}

extern "C" int cffc_check_arith(const double *a, const double *b, int64_t n, double *out) {
    if (!a || !b || !out || n < 0) return fail(CFFC_E_ARG, "null argument or negative count");
    if (n == 0) return 0;
    double *da = nullptr, *dout = nullptr;
    if (hipMalloc((void **)&da, 2 * n * sizeof(double)) != hipSuccess)
        return fail(CFFC_E_NOMEM, "hipMalloc of %zu bytes failed", (size_t)(2 * n * sizeof(double)));
    if (hipMalloc((void **)&dout, 8 * n * sizeof(double)) != hipSuccess) {
        (void)hipFree(da);
        return fail(CFFC_E_NOMEM, "hipMalloc of %zu bytes failed", (size_t)(8 * n * sizeof(double)));
    }
    hipError_t e = hipMemcpy(da, a, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(da + n, b, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = cffc_launch_arith_check(da, da + n, n, dout, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(CFFC_E_HIP, "cffc_check_arith: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int cffc_check_mesh(const cffc_mesh_desc *md, int64_t counts[4]) {
    if (!md) return fail(CFFC_E_ARG, "null argument");
    if (md->nblk <= 0 || md->nghost != 2) return fail(CFFC_E_ARG, "nblk must be > 0 and nghost == 2");
    cffc_ctx *c = new cffc_ctx();
    c->nblk = md->nblk;
    c->s.n_stage = 1;
    int32_t bcoff = 0;
    int64_t noff = 0;
    std::vector<int32_t> same;
    std::vector<RestrictOp> rops;
    std::vector<ProlongOp> pops;
    std::vector<FluxFixOp> fops;
    std::vector<uint8_t> ustore;
    int64_t nentries = 0;
    int e = build_meta(c, md, bcoff, noff);
    if (!e) e = md->lattice ? setup_lattice(c, md, same, rops, pops, fops, ustore, nentries)
                            : setup_levels(c, md, same, rops, pops, fops, ustore, nentries);
    if (!e && c->amr) {
        static const int opp[8] = {1, 0, 3, 2, 7, 6, 5, 4};
        for (int b = 0; b < md->nblk && !e; ++b)
            for (int s = 0; s < 8; ++s) {
                const int nb = same[8 * b + s];
                if (nb >= 0 && same[8 * nb + opp[s]] != b) {
                    e = fail(CFFC_E_TOPOLOGY, "a mesh with resolution changes needs mutual same-level neighbours");
                    break;
                }
            }
    }
    if (!e && counts) {
        counts[0] = (int64_t)rops.size();
        counts[1] = (int64_t)pops.size();
        counts[2] = (int64_t)fops.size();
        counts[3] = c->amr ? 1 : 0;
    }
    delete c;
    return e;
}

// The operation lists cffc_create would build for a mesh, host only:
// kind 0 restrictions, 1 same-level copies (RestrictOp, 24 bytes each),
// 2 prolongations (ProlongOp), 3 flux corrections (FluxFixOp), as raw structs.
extern "C" int cffc_mesh_ops(const cffc_mesh_desc *md, int kind, void *out, int64_t cap, int64_t *n) {
    if (!md || !n || kind < 0 || kind > 3) return fail(CFFC_E_ARG, "bad argument");
    if (md->nblk <= 0 || md->nghost != 2) return fail(CFFC_E_ARG, "nblk must be > 0 and nghost == 2");
    cffc_ctx *c = new cffc_ctx();
    c->nblk = md->nblk;
    c->s.n_stage = 1;
    int32_t bcoff = 0;
    int64_t noff = 0, nentries = 0;
    std::vector<int32_t> same;
    std::vector<RestrictOp> rops, sel;
    std::vector<ProlongOp> pops;
    std::vector<FluxFixOp> fops;
    std::vector<uint8_t> ustore;
    int e = build_meta(c, md, bcoff, noff);
    if (!e) e = md->lattice ? setup_lattice(c, md, same, rops, pops, fops, ustore, nentries)
                            : setup_levels(c, md, same, rops, pops, fops, ustore, nentries);
    if (!e) {
        const void *src = nullptr;
        size_t sz = 0, cnt = 0;
        if (kind <= 1) {
            for (const RestrictOp &o : rops)
                if (o.kind == kind) sel.push_back(o);
            src = sel.data(); sz = sizeof(RestrictOp); cnt = sel.size();
        } else if (kind == 2) {
            src = pops.data(); sz = sizeof(ProlongOp); cnt = pops.size();
        } else {
            src = fops.data(); sz = sizeof(FluxFixOp); cnt = fops.size();
        }
        *n = (int64_t)cnt;
        if (out && cap > 0) memcpy(out, src, sz * std::min<size_t>(cnt, (size_t)cap));
    }
    delete c;
    return e;
}

extern "C" void cffc_destroy(cffc_ctx *c) {
    if (c && c->stream) (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
}

extern "C" int64_t cffc_num_cells(const cffc_ctx *c) { return c ? c->ncell : 0; }
extern "C" void *cffc_stream(cffc_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int cffc_set_tile(cffc_ctx *c, int ti, int tj) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    if (!tile_supported(ti, tj, c->s.flux_precision == CFFC_FLUX_FP64))
        return fail(CFFC_E_ARG, "tile %dx%d is not instantiated", ti, tj);
    HIPCHK(hipStreamSynchronize(c->stream));
    return build_tiles(c, ti, tj);
}
extern "C" int cffc_get_tile(const cffc_ctx *c, int *ti, int *tj) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    *ti = c->ti;
    *tj = c->tj;
    return 0;
}

static int ensure_staging(cffc_ctx *c) {
    if (!c->staging) return dalloc(&c->staging, 4 * (size_t)c->ncell);
    return 0;
}

static int set_state(cffc_ctx *c, const double *host, int64_t n, int kind) {
    if (!c || !host) return fail(CFFC_E_ARG, "null argument");
    if (n != c->ncell) return fail(CFFC_E_ARG, "state has %lld cells, mesh has %lld", (long long)n, (long long)c->ncell);
    if (int e = ensure_staging(c)) return e;
    HIPCHK(hipMemcpyAsync(c->staging, host, 4 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(cffc_launch_scatter(c->meta, c->nblk, c->max_int, c->staging, c->W[c->wcur], c->ncp, kind, c->U, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->phi_valid = 0;
    c->freeze = 0;  // a new state starts with limiters evaluated again
    return 0;
}
extern "C" int cffc_set_W(cffc_ctx *c, const double *W, int64_t n) { return set_state(c, W, n, 0); }
extern "C" int cffc_set_U(cffc_ctx *c, const double *U, int64_t n) { return set_state(c, U, n, 1); }

static int get_field(cffc_ctx *c, const double *src, double *host, int64_t n, int nvar) {
    if (!c || !host) return fail(CFFC_E_ARG, "null argument");
    if (n != c->ncell) return fail(CFFC_E_ARG, "buffer has %lld cells, mesh has %lld", (long long)n, (long long)c->ncell);
    if (int e = ensure_staging(c)) return e;
    HIPCHK(cffc_launch_gather(c->meta, c->nblk, c->max_int, src, c->staging, c->ncp, nvar, c->stream));
    HIPCHK(hipMemcpyAsync(host, c->staging, nvar * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
extern "C" int cffc_get_W(cffc_ctx *c, double *W, int64_t n) { return c ? get_field(c, c->W[c->wcur], W, n, 4) : CFFC_E_ARG; }
extern "C" int cffc_get_U(cffc_ctx *c, double *U, int64_t n) { return c ? get_field(c, c->U, U, n, 4) : CFFC_E_ARG; }
extern "C" int cffc_get_dt(cffc_ctx *c, double *dt, int64_t n) { return c ? get_field(c, c->dt, dt, n, 1) : CFFC_E_ARG; }
extern "C" int cffc_get_dUdt(cffc_ctx *c, int k, double *d, int64_t n) {
    if (!c || k < 0 || k > 2) return fail(CFFC_E_ARG, "bad residual index");
    if (!c->dUdt[k]) return fail(CFFC_E_ARG, "dUdt[%d] has not been computed", k);
    if (k == 0 && c->sf_live) {
        // the last fused mixed run kept its residual in FP32
        HIPCHK(cffc_launch_widen(c->Sf, c->dUdt[0], 4 * c->ncp, c->stream));
        c->sf_live = false;
    }
    return get_field(c, c->dUdt[k], d, n, 4);
}

extern "C" int cffc_synchronize(cffc_ctx *c) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

static ExchangeParams exchange_params(cffc_ctx *c, int from_U, unsigned long long *reset_slot) {
    ExchangeParams p;
    p.meta = c->meta;
    p.nbr = c->nbr;
    p.bc = c->bc;
    p.wo = c->wo;
    p.geom = c->geom;
    p.ncp = c->ncp;
    p.W = c->W[c->wcur];
    p.U = c->U;
    p.from_U = from_U;
    p.stop_flag = nullptr;
    p.reset_slot = reset_slot;
    p.blocks = nullptr;
    p.bcs_only = 0;
    p.no_bcs = 0;
    p.store_U = 0;
    p.Wf = nullptr;
    p.W2 = nullptr;
    p.Wf2 = nullptr;
    // the same-level phase has 4 (ni + nj) + 16 ghost cells per block, about
    // 8 x (padded edge) - 16, the BC phases 2 x (padded edge): small blocks
    // (the coarse multigrid levels) get one wavefront
    const int items = 8 * c->max_edge - 16;
    p.nthreads = items <= 64 ? 64 : items <= 128 ? 128 : 256;
    return p;
}

static int ensure_phi(cffc_ctx *c);

// The resolution-change legs of Send_All_Messages on the W buffer (W or Wf)
// that the next stage reads: restriction from U (the state U^(s) of the
// restricted fine cells), then prolongation, which reads the ghosts the
// restriction filled.
static int amr_legs(cffc_ctx *c, const double *U, double *W, float *Wf, const int32_t *part_state,
                    const int32_t *stop_flag) {
    if (!c->amr) return 0;
    // same-level copies (lattice meshes) take W(U) where U is the current
    // state: the first stage and the granular path
    HIPCHK(cffc_launch_restrict(c->rops, c->nrops, U, c->geom, c->ncp, Wf ? nullptr : W, Wf, stop_flag, U == c->U,
                                c->stream));
    HIPCHK(cffc_launch_prolong(c->pops, c->npops, c->meta, c->geom, c->phi, c->freeze, part_state, c->ncp,
                               Wf ? nullptr : W, Wf, stop_flag, c->stream));
    return 0;
}

// ---- granular entry points (one reference call each) -----------------------
static int ensure_granular(cffc_ctx *c) {
    if (c->granular_ready) return 0;
    const size_t n = 4 * (size_t)c->ncp;
    int e;
    if (!c->Uo && (e = dalloc(&c->Uo, n))) return e;
    for (int k = 1; k < 3; ++k)
        if (!c->dUdt[k]) {
            if ((e = dalloc(&c->dUdt[k], n))) return e;
            HIPCHK(hipMemsetAsync(c->dUdt[k], 0, n * sizeof(double), c->stream));
        }
    c->granular_ready = true;
    return 0;
}

// Residual_Smoothing scratch is dUdt[k+1]: the reference's 4-stage Runge-Kutta
// smooths dUdt[2] at stage 3 into a fourth residual vector it does not have
// (NUMBER_OF_RESIDUAL_VECTORS_EULER2D = 3), so that scheme is refused.
static bool smoothing_supported(const cffc_scheme_desc &s) {
    return !(s.i_time_integration == CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA && s.n_stage == 4);
}

extern "C" int cffc_set_residual_smoothing(cffc_ctx *c, double epsilon, int iterations) {
    if (!c || iterations < 0 || !(epsilon >= 0.0)) return fail(CFFC_E_ARG, "bad argument");
    if (iterations > 0 && !smoothing_supported(c->s))
        return fail(CFFC_E_ARG, "residual smoothing is not supported with 4-stage Runge-Kutta");
    c->rs_eps = epsilon;
    c->rs_iters = iterations;
    return 0;
}

extern "C" int cffc_residual_smoothing(cffc_ctx *c, int i_stage) {
    if (!c || i_stage < 1 || i_stage > c->s.n_stage) return fail(CFFC_E_ARG, "bad argument");
    if (int e = ensure_granular(c)) return e;
    if (c->rs_iters == 0) return 0;
    const int k = k_residual(c->s, i_stage);
    HIPCHK(cffc_launch_residual_smoothing(c->meta, c->nblk, c->dUdt[k], c->dUdt[k + 1], c->ncp, c->rs_eps,
                                          c->rs_iters, nullptr, nullptr, c->stream));
    return 0;
}

extern "C" int cffc_exchange_and_bcs(cffc_ctx *c) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    // the granular path keeps U current, so ghosts are always W(U)
    ExchangeParams p = exchange_params(c, 1, nullptr);
    if (c->amr) {
        // same-level, fine-to-coarse and coarse-to-fine legs, then the BCs
        if (int e = ensure_phi(c)) return e;
        p.no_bcs = 1;
        HIPCHK(cffc_launch_exchange(&p, c->nblk, c->stream));
        if (int e = amr_legs(c, c->U, c->W[c->wcur], nullptr, nullptr, nullptr)) return e;
        ExchangeParams q = exchange_params(c, 1, nullptr);
        q.bcs_only = 1;
        q.blocks = c->bc_blocks;
        if (c->nbc) HIPCHK(cffc_launch_exchange(&q, c->nbc, c->stream));
    } else {
        HIPCHK(cffc_launch_exchange(&p, c->nblk, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int cffc_apply_flux_corrections(cffc_ctx *c, int i_stage) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    if (i_stage < 1 || i_stage > c->s.n_stage) return fail(CFFC_E_ARG, "bad stage %d", i_stage);
    if (!c->amr || c->nfops == 0) return 0;
    if (int e = ensure_granular(c)) return e;
    const int k = k_residual(c->s, i_stage);
    // the reference reads the cell's dt array (global stepping filled it)
    HIPCHK(cffc_launch_flux_fix(c->fops, c->nfops, c->dUdt[k], c->eflux, c->geom, c->dt, c->scal, 1,
                                c->s.cfl_number, c->ncp, nullptr, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int cffc_cfl(cffc_ctx *c, double *dt_min, int set_global) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    HIPCHK(cffc_launch_set_u64(c->scal, MILLION_BITS, 0, 1, c->stream));
    HIPCHK(cffc_launch_cfl(c->meta, c->tiles, c->ntiles, c->ti, c->tj, c->W[c->wcur], c->geom, c->ncp, c->dt, c->scal,
                           c->stream));
    if (set_global) HIPCHK(cffc_launch_set_global_dt(c->meta, c->nblk, c->max_int, c->dt, c->scal, c->stream));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, c->scal, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double d;
    memcpy(&d, &bits, sizeof d);
    if (dt_min) *dt_min = d;
    return 0;
}

extern "C" int cffc_set_global_timestep(cffc_ctx *c, double dt) {
    if (!c || !(dt > 0.0)) return fail(CFFC_E_ARG, "bad argument");
    unsigned long long bits;
    memcpy(&bits, &dt, sizeof bits);
    HIPCHK(cffc_launch_set_u64(c->scal, bits, 0, 1, c->stream));
    HIPCHK(cffc_launch_set_global_dt(c->meta, c->nblk, c->max_int, c->dt, c->scal, c->stream));
    return 0;
}

static StageParams base_stage_params(cffc_ctx *c) {
    StageParams p;
    memset(&p, 0, sizeof p);
    p.meta = c->meta;
    p.tiles = c->tiles;
    p.ntiles = c->ntiles;
    p.ncp = c->ncp;
    p.geom = c->geom;
    p.geomf = c->geomf;
    p.bc = c->bc;
    p.wo = c->wo;
    p.U = c->U;
    p.dt = c->dt;
    p.phi = c->phi;
    p.cfl = c->s.cfl_number;
    p.limiter = c->s.i_limiter;
    p.lts = c->s.local_time_stepping == CFFC_SCALAR_LOCAL_TIME_STEPPING;
    p.err = c->scal + 2;
    p.norm_part = c->norm_part;
    p.stop_flag = nullptr;
    p.nbr = c->nbr;
    p.push = 0;
    p.phi_mode = c->freeze ? PHI_LOAD : PHI_COMPUTE;
    p.eflux = c->eflux;
    p.ustore = c->ustore;
    p.axisym = c->s.axisymmetric;
    return p;
}

static hipError_t launch_stage(cffc_ctx *c, const StageParams &p, int fused) {
    if (c->s.flux_precision == CFFC_FLUX_FP32) return cffc_launch_stage_mixed(&p, c->ti, c->tj, fused, c->stream);
    return cffc_launch_stage_fp64(&p, c->ti, c->tj, fused, c->stream);
}

static int ensure_phi(cffc_ctx *c) {
    if (!c->phi) {
        if (int e = dalloc(&c->phi, 4 * (size_t)c->ncp)) return e;
        HIPCHK(hipMemsetAsync(c->phi, 0, 4 * c->ncp * sizeof(double), c->stream));
    }
    return 0;
}

extern "C" int cffc_freeze_limiters(cffc_ctx *c, int freeze) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    if (freeze && !c->phi_valid) return fail(CFFC_E_ARG, "no evaluated limiters to freeze (run a stage first)");
    c->freeze = freeze ? 1 : 0;
    return 0;
}

extern "C" int cffc_dudt_stage(cffc_ctx *c, int i_stage) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    if (i_stage < 1 || i_stage > c->s.n_stage) return fail(CFFC_E_ARG, "bad stage %d", i_stage);
    c->sf_live = false;   // dUdt[k] is written in FP64 here
    if (int e = ensure_granular(c)) return e;
    if (int e = ensure_phi(c)) return e;
    if (i_stage == 1)
        HIPCHK(cffc_launch_copy_interior(c->meta, c->nblk, c->max_int, c->U, c->Uo, c->ncp, c->stream));
    StageParams p = base_stage_params(c);
    p.Win = c->W[c->wcur];
    p.Wout = c->W[c->wcur];
    const int k = k_residual(c->s, i_stage);
    p.S = c->dUdt[k];
    p.S1 = c->dUdt[1];
    p.S2 = c->dUdt[2];
    p.mode = granular_init(c->s, i_stage);
    p.dtmin_in = c->scal;
    p.phi_mode = c->freeze ? PHI_LOAD : PHI_COMPUTE_STORE;
    HIPCHK(launch_stage(c, p, 0));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->freeze) c->phi_valid = 1;
    return 0;
}

extern "C" int cffc_update_stage(cffc_ctx *c, int i_stage, int64_t *bad_cell) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    if (i_stage < 1 || i_stage > c->s.n_stage) return fail(CFFC_E_ARG, "bad stage %d", i_stage);
    if (int e = ensure_granular(c)) return e;
    const int k = k_residual(c->s, i_stage);
    HIPCHK(cffc_launch_set_u64(c->scal + 2, ERR_NONE, 0, 1, c->stream));
    HIPCHK(cffc_launch_update(c->meta, c->nblk, c->max_int, c->W[c->wcur], c->U, c->Uo, c->dUdt[k], c->dt, c->ncp,
                              stage_omega(c->s, i_stage),
                              c->s.local_time_stepping == CFFC_SCALAR_LOCAL_TIME_STEPPING, c->scal + 2, c->stream));
    unsigned long long key = 0;
    HIPCHK(hipMemcpyAsync(&key, c->scal + 2, sizeof key, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (key != ERR_NONE) {
        if (bad_cell) *bad_cell = (int64_t)key;
        return (int)(key & 0xFFFF);  // the reference returns the i-index
    }
    if (bad_cell) *bad_cell = -1;
    return 0;
}

extern "C" int cffc_residual_norms(cffc_ctx *c, int var, double out[3]) {
    if (!c || !out || var < 1 || var > 4) return fail(CFFC_E_ARG, "bad argument");
    HIPCHK(cffc_launch_norms(c->meta, c->tiles, c->ntiles, c->ti, c->tj, c->dUdt[0], c->ncp, var, c->norm_part,
                             c->norm_small, c->stream));
    HIPCHK(hipMemcpyAsync(out, c->norm_small, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- the fused time-marching loop ------------------------------------------
static int ensure_steps(cffc_ctx *c, int nsteps) {
    if (nsteps <= c->max_steps) return 0;
    if (c->norm_out) (void)hipFree(c->norm_out);
    if (c->dt_out) (void)hipFree(c->dt_out);
    if (c->step_clock) (void)hipFree(c->step_clock);
    c->norm_out = c->dt_out = nullptr;
    c->step_clock = nullptr;
    int cap = std::max(nsteps, 1024);
    if (int e = dalloc(&c->norm_out, 3 * (size_t)cap)) return e;
    if (int e = dalloc(&c->dt_out, (size_t)cap)) return e;
    if (int e = dalloc(&c->step_clock, (size_t)cap + 1)) return e;
    c->max_steps = cap;
    return 0;
}

static hipEvent_t next_event(cffc_ctx *c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}
static void mark(cffc_ctx *c, int kind, bool begin) {
    if (!c->timing) return;
    hipEvent_t e = next_event(c);
    if (!e) return;
    (void)hipEventRecord(e, c->stream);
    if (begin) c->ev_kind.push_back(kind);
}
static void collect_timing(cffc_ctx *c) {
    if (!c->timing) return;
    for (size_t q = 0; q + 1 < c->ev_used && q / 2 < c->ev_kind.size(); q += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pool[q], c->ev_pool[q + 1]) == hipSuccess) {
            c->t_ms[c->ev_kind[q / 2]] += ms;
            c->t_n[c->ev_kind[q / 2]] += 1;
        }
    }
    c->ev_used = 0;
    c->ev_kind.clear();
}

// enqueue one step (all stages) with step parity par
// One fused time step.  parts: run the batched steady form (per-part
// freeze/convergence with freeze_level and tol per part).
static int enqueue_step(cffc_ctx *c, int par, double tol, bool parts = false, double freeze_level = 0.0) {
    const int N = c->s.n_stage;
    for (int st = 1; st <= N; ++st) {
        ExchangeParams ex = exchange_params(c, st == 1, nullptr);
        ex.stop_flag = c->stop_flag;
        ex.Wf = c->Wf[c->wcur];
        mark(c, 0, true);
        // messages across resolution changes (the same-level ones were pushed
        // by the previous stage), then the BCs: Euler2DQuadSolvers.cc:807-826.
        // The restricted fine cells hold U^n at stage 1, U^(s-1) after.
        if (int e = amr_legs(c, st == 1 ? c->U : c->Ustage, c->W[c->wcur], c->Wf[c->wcur],
                             parts ? c->part_state : nullptr, c->stop_flag))
            return e;
        if (c->amr) {
            // the next stage's coarse-to-fine messages read BC ghosts of this
            // stage's BCs call, as in the reference's single W array
            if (c->Wf[0]) ex.Wf2 = c->Wf[1 - c->wcur];
            else ex.W2 = c->W[1 - c->wcur];
        }
        if (c->push_ok) {
            // same-level ghosts were pushed by the previous stage kernel (or
            // filled by begin_run); only the physical BCs remain
            ex.bcs_only = 1;
            ex.blocks = c->bc_blocks;
            if (c->nbc) HIPCHK(cffc_launch_exchange(&ex, c->nbc, c->stream));
        } else {
            HIPCHK(cffc_launch_exchange(&ex, c->nblk, c->stream));
        }
        mark(c, 0, false);
        StageParams p = base_stage_params(c);
        p.Win = c->W[c->wcur];
        p.Wout = c->W[1 - c->wcur];
        p.Winf = c->Wf[c->wcur];
        p.Woutf = c->Wf[1 - c->wcur];
        p.S = c->dUdt[0];
        // the FP32 accumulator, except where the residual smoothing shares
        // dUdt[0] between the stored residual and the update
        p.Sf = c->rs_iters > 0 ? nullptr : c->Sf;
        p.mode = fused_mode(c->s, st);
        p.omega = stage_omega(c->s, st);
        p.last_stage = (st == N);
        // S is dead after this stage when the next one starts from zero
        // (multistage optimal smoothing); the last stage always stores it
        p.skip_S = (st < N && fused_mode(c->s, st + 1) == MODE_ZERO) ? 1 : 0;
        p.dtmin_in = c->scal + par;
        p.stop_flag = c->stop_flag;
        p.push = c->push_ok ? 1 : 0;
        if (parts) {
            p.part_state = c->part_state;
            p.done_copy = N & 1;
        }
        if (c->freeze) p.phi_mode = PHI_LOAD;
        else if (c->store_phi && st == N) p.phi_mode = PHI_COMPUTE_STORE;
        else p.phi_mode = PHI_COMPUTE;
        mark(c, 1, true);
        p.Ustage = c->Ustage;
        const bool fp32 = c->s.flux_precision == CFFC_FLUX_FP32;
        if (c->rs_iters > 0) {
            // residual, flux corrections, implicit residual smoothing, update
            // (Euler2DQuadSolvers.cc:841-877)
            p.defer_update = 1;
            p.Sd = c->dUdt[0];
            HIPCHK(launch_stage(c, p, 1));
            if (c->nfops)
                HIPCHK(cffc_launch_flux_fix(c->fops, c->nfops, c->dUdt[0], c->eflux, c->geom, c->dt, c->scal + par,
                                            p.lts, c->s.cfl_number, c->ncp, c->stop_flag, c->stream));
            HIPCHK(cffc_launch_residual_smoothing(c->meta, c->nblk, c->dUdt[0], c->dUdt[1], c->ncp, c->rs_eps,
                                                  c->rs_iters, parts ? c->part_state : nullptr, c->stop_flag,
                                                  c->stream));
            p.defer_update = 0;
            if (fp32) HIPCHK(cffc_launch_update_tile_mixed(&p, c->ti, c->tj, c->stream));
            else HIPCHK(cffc_launch_update_tile_fp64(&p, c->ti, c->tj, c->stream));
        } else if (c->amr) {
            // tiles next to a finer face store their residual; the coarse face
            // fluxes are corrected (Euler2DQuadSolvers.cc:841-853), then those
            // tiles update
            p.tile_defer = c->tile_defer;
            p.Sd = c->Sd;
            HIPCHK(launch_stage(c, p, 1));
            if (c->nfops)
                HIPCHK(cffc_launch_flux_fix(c->fops, c->nfops, c->Sd, c->eflux, c->geom, c->dt, c->scal + par, p.lts,
                                            c->s.cfl_number, c->ncp, c->stop_flag, c->stream));
            if (c->ndefer) {
                p.tile_list = c->defer_list;
                p.nlist = c->ndefer;
                if (fp32) HIPCHK(cffc_launch_update_tile_mixed(&p, c->ti, c->tj, c->stream));
                else HIPCHK(cffc_launch_update_tile_fp64(&p, c->ti, c->tj, c->stream));
            }
        } else {
            HIPCHK(launch_stage(c, p, 1));
        }
        mark(c, 1, false);
        c->wcur = 1 - c->wcur;
    }
    FinishParams f;
    memset(&f, 0, sizeof f);
    f.norm_part = c->norm_part;
    f.ntiles = c->ntiles;
    f.dtmin_in = c->scal + par;
    f.dtmin_out = c->scal + (1 - par);
    f.norm_out = c->norm_out;
    f.dt_out = c->dt_out;
    f.step_ctr = c->step_ctr;
    f.stop_flag = c->stop_flag;
    f.tol = tol;
    if (parts) {
        f.npart = c->npart;
        f.part_tile0 = c->part_tile0;
        f.part_state = c->part_state;
        f.part_steps = c->part_steps;
        f.part_l2 = c->part_l2;
        f.part_last = c->part_last;
        f.freeze_level = freeze_level;
        f.step_offset = c->steady_done;
        f.stall_window = c->stall_window;
        f.part_best = c->part_best;
        f.part_best_step = c->part_best_step;
    }
    f.step_clock = c->step_clock ? c->step_clock + 1 : nullptr;
    // meshes of many tiles reduce their partials in two levels (256 chunks)
    const int nchunk = (!parts && c->chunk_part) ? 256 : 0;
    HIPCHK(cffc_launch_finish(&f, c->chunk_part, nchunk, c->stream));
    return 0;
}

static int begin_run(cffc_ctx *c, int nsteps) {
    if (int e = ensure_steps(c, nsteps)) return e;
    c->sf_live = c->Sf && c->rs_iters == 0;
    if (c->rs_iters > 0)
        if (int e = ensure_granular(c)) return e;   // dUdt[1] is the smoothing scratch
    if (c->store_phi || c->freeze)
        if (int e = ensure_phi(c)) return e;
    int32_t zero[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(c->step_ctr, zero, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->stop_flag, zero, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->counter, 0, sizeof(unsigned), c->stream));
    HIPCHK(cffc_launch_set_u64(c->scal + 2, ERR_NONE, 0, 1, c->stream));
    // CFL of the first step from the current state (slot 0)
    HIPCHK(cffc_launch_set_u64(c->scal, MILLION_BITS, MILLION_BITS, 2, c->stream));
    HIPCHK(cffc_launch_cfl(c->meta, c->tiles, c->ntiles, c->ti, c->tj, c->W[c->wcur], c->geom, c->ncp, c->dt, c->scal,
                           c->stream));
    if (c->Wf[0]) {
        // the fused mixed path keeps W in FP32: derive it from U
        HIPCHK(cffc_launch_w_from_u(c->meta, c->nblk, c->max_int, c->U, nullptr, c->Wf[c->wcur], c->ncp, c->stream));
    }
    if (c->push_ok || c->Wf[0]) {
        // full exchange (ghosts = W(U) of the neighbours, then BCs): later
        // stages get their same-level ghosts pushed by the stage kernel
        ExchangeParams ex = exchange_params(c, 1, nullptr);
        ex.Wf = c->Wf[c->wcur];
        HIPCHK(cffc_launch_exchange(&ex, c->nblk, c->stream));
    }
    if (c->amr) {
        // Ghosts of the new state across resolution changes.  The
        // coarse-to-fine leg reads coarse ghosts that the BCs fill after it,
        // and BC corner rules read ghosts of that leg one level coarser:
        // max_level + 2 rounds of messages + BCs make every ghost a function
        // of the interior state (the first stage then repeats the messages).
        if (int e = ensure_phi(c)) return e;
        for (int r = 0; r < c->max_level + 2; ++r) {
            if (int e = amr_legs(c, c->U, c->W[c->wcur], c->Wf[c->wcur], nullptr, nullptr)) return e;
            ExchangeParams q = exchange_params(c, 1, nullptr);
            q.Wf = c->Wf[c->wcur];
            q.bcs_only = 1;
            q.blocks = c->bc_blocks;
            if (c->nbc) HIPCHK(cffc_launch_exchange(&q, c->nbc, c->stream));
        }
    }
    HIPCHK(cffc_launch_stamp(c->step_clock, c->stream));
    return 0;
}

static int end_run(cffc_ctx *c, int wstart, int nsteps, double *norms, double *dt_out, int *steps_done) {
    int32_t done = 0;
    unsigned long long key = 0;
    HIPCHK(hipMemcpyAsync(&done, c->step_ctr, sizeof done, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&key, c->scal + 2, sizeof key, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (done > nsteps) done = nsteps;
    if (norms && done > 0) HIPCHK(hipMemcpy(norms, c->norm_out, 3 * (size_t)done * sizeof(double), hipMemcpyDeviceToHost));
    if (dt_out && done > 0) HIPCHK(hipMemcpy(dt_out, c->dt_out, (size_t)done * sizeof(double), hipMemcpyDeviceToHost));
    if (steps_done) *steps_done = done;
    // kernels after a stop are skipped: the newest W is the one after `done` steps
    c->wcur = (wstart + done * c->s.n_stage) & 1;
    if (c->Wf[0]) {
        // FP64 W of the state for get_W and the granular entry points
        HIPCHK(cffc_launch_w_from_u(c->meta, c->nblk, c->max_int, c->U, c->W[c->wcur], nullptr, c->ncp, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (c->store_phi) c->phi_valid = 1;
    if (key != ERR_NONE) {
        g_err = "non-positive density or energy";
        return (int)(key & 0xFFFF);
    }
    return 0;
}

extern "C" int cffc_run_steps(cffc_ctx *c, int nsteps, double tol, double *norms, double *dt_out, int *steps_done) {
    if (!c || nsteps < 0) return fail(CFFC_E_ARG, "bad argument");
    if (int e = begin_run(c, nsteps)) return e;
    const int wstart = c->wcur;
    int32_t *hflag = nullptr;
    HIPCHK(hipHostMalloc((void **)&hflag, 2 * sizeof(int32_t), hipHostMallocDefault));
    hflag[0] = 0;
    for (int n = 0; n < nsteps; ++n) {
        if (int e = enqueue_step(c, n & 1, tol)) {
            (void)hipHostFree(hflag);
            return e;
        }
        // poll for convergence / failure now and then without draining the queue
        if ((n & 63) == 63) {
            if (c->timing) {
                (void)hipStreamSynchronize(c->stream);
                collect_timing(c);
            }
            if (hflag[0]) break;
            (void)hipMemcpyAsync(hflag, c->stop_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
        }
    }
    (void)hipStreamSynchronize(c->stream);
    collect_timing(c);
    (void)hipHostFree(hflag);
    return end_run(c, wstart, nsteps, norms, dt_out, steps_done);
}

// ---- the fused loop without allocations -------------------------------------
// cffc_run_steps grows per-context device buffers for its per-step records
// (norms, time steps, step clock) and polls through pinned host memory.  The
// variant below keeps those records in `scratch`, a file-scope device array
// of fixed size, and does not poll, so that a caller that may not allocate
// (graph capture, a serving loop, a harness that forbids it) can advance an
// existing context.  Every device holds its own copy of the array for the
// life of the process; host code uses the address hipGetSymbolAddress gave
// for that device (g_scratch), never the symbol.  Calls that overlap on one
// device share the array and race.
static constexpr int NOALLOC_MAX_STEPS = 4096;
// norms [3][max], dt [max], then the step clock [max + 1] (64-bit words)
static constexpr size_t NOALLOC_WORDS = 3 * (size_t)NOALLOC_MAX_STEPS + NOALLOC_MAX_STEPS + NOALLOC_MAX_STEPS + 1;
static __device__ __align__(256) double scratch[NOALLOC_WORDS];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[NOALLOC_MAX_DEVICES];

// Once per device, with that device current (cffc_create calls it).
static int resolve_scratch(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(CFFC_E_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(CFFC_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
    return 0;
}

extern "C" int cffc_run_steps_noalloc(cffc_ctx *c, int nsteps, double tol, double *norms, double *dt_out,
                                      int *steps_done) {
    if (!c || nsteps < 0) return fail(CFFC_E_ARG, "bad argument");
    if (nsteps > NOALLOC_MAX_STEPS)
        return fail(CFFC_E_ARG, "%d steps exceed the %d-step record of cffc_run_steps_noalloc", nsteps,
                    NOALLOC_MAX_STEPS);
    double *sp = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) sp = g_scratch[c->device];
    }
    if (!sp) return fail(CFFC_E_HIP, "no scratch address for device %d", c->device);
    // buffers the run needs must already exist (they are allocated by the
    // set-up calls); kernel timing records events and is refused
    if (c->timing) return fail(CFFC_E_ARG, "kernel timing is on: cffc_run_steps_noalloc records no events");
    if (c->rs_iters > 0 && !c->granular_ready)
        return fail(CFFC_E_ARG, "residual smoothing buffers are not allocated");
    if ((c->store_phi || c->freeze || c->amr) && !c->phi)
        return fail(CFFC_E_ARG, "limiter buffer is not allocated");
    double *own_norm = c->norm_out, *own_dt = c->dt_out;
    unsigned long long *own_clock = c->step_clock;
    const int own_max = c->max_steps;
    c->norm_out = sp;
    c->dt_out = sp + 3 * (size_t)NOALLOC_MAX_STEPS;
    c->step_clock = reinterpret_cast<unsigned long long *>(sp + 4 * (size_t)NOALLOC_MAX_STEPS);
    c->max_steps = NOALLOC_MAX_STEPS;
    int rc = begin_run(c, nsteps);
    const int wstart = c->wcur;
    for (int n = 0; !rc && n < nsteps; ++n) rc = enqueue_step(c, n & 1, tol);
    if (!rc) {
        // converged or failed parts skip their kernels; no host polling
        (void)hipStreamSynchronize(c->stream);
        rc = end_run(c, wstart, nsteps, norms, dt_out, steps_done);
    } else {
        (void)hipStreamSynchronize(c->stream);
    }
    c->norm_out = own_norm;
    c->dt_out = own_dt;
    c->step_clock = own_clock;
    c->max_steps = own_max;
    return rc;
}

extern "C" int cffc_run_steps_graph(cffc_ctx *c, int nsteps, double *norms, double *dt_out) {
    if (!c || nsteps < 0) return fail(CFFC_E_ARG, "bad argument");
    if (int e = begin_run(c, nsteps)) return e;
    const int wstart = c->wcur;
    const int pairs = nsteps / 2;
    if (pairs > 0) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        const int w0 = c->wcur;
            const bool tsave = c->timing;
        c->timing = false;  // no event pairs inside a captured graph
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int e = enqueue_step(c, 0, 0.0);
        if (!e) e = enqueue_step(c, 1, 0.0);
        hipError_t ce = hipStreamEndCapture(c->stream, &g);
        c->timing = tsave;
        if (e) return e;
        if (ce != hipSuccess) return fail(CFFC_E_HIP, "graph capture failed: %s", hipGetErrorString(ce));
        HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        for (int q = 0; q < pairs; ++q) HIPCHK(hipGraphLaunch(ge, c->stream));
        c->wcur = w0;  // every replay ends where it began (2*N_Stage buffer flips)
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
    }
    if (nsteps & 1)
        if (int e = enqueue_step(c, 0, 0.0)) return e;
    (void)hipStreamSynchronize(c->stream);
    int done = 0;
    int rc = end_run(c, wstart, nsteps, norms, dt_out, &done);
    return rc;
}

// ---- batched steady runs ----------------------------------------------------
extern "C" int cffc_set_parts(cffc_ctx *c, int npart, const int32_t *first_block) {
    if (!c || npart < 1 || npart > c->nblk || !first_block) return fail(CFFC_E_ARG, "bad argument");
    if (c->amr && npart > 1) return fail(CFFC_E_ARG, "a mesh with resolution changes runs as one part");
    if (first_block[0] != 0 || first_block[npart] != c->nblk) return fail(CFFC_E_ARG, "parts must cover all blocks");
    std::vector<int32_t> pob(c->nblk);
    for (int q = 0; q < npart; ++q) {
        if (first_block[q + 1] <= first_block[q]) return fail(CFFC_E_ARG, "part %d is empty", q);
        for (int b = first_block[q]; b < first_block[q + 1]; ++b) pob[b] = q;
    }
    // parts must not exchange ghost cells with each other
    std::vector<int32_t> nbr_h(8 * (size_t)c->nblk);
    HIPCHK(hipMemcpy(nbr_h.data(), c->nbr, nbr_h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int b = 0; b < c->nblk; ++b)
        for (int k = 0; k < 8; ++k) {
            const int nb = nbr_h[8 * (size_t)b + k];
            if (nb >= 0 && pob[nb] != pob[b]) return fail(CFFC_E_TOPOLOGY, "block %d neighbours block %d of another part", b, nb);
        }
    for (int32_t **q : {&c->part_state, &c->part_steps})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (c->part_last) { (void)hipFree(c->part_last); c->part_last = nullptr; }
    if (c->part_best) { (void)hipFree(c->part_best); c->part_best = nullptr; }
    if (c->part_best_step) { (void)hipFree(c->part_best_step); c->part_best_step = nullptr; }
    c->npart = npart;
    c->part_of_block = pob;
    for (int b = 0; b < c->nblk; ++b) c->meta_h[b].spare = pob[b];
    HIPCHK(hipMemcpy(c->meta, c->meta_h.data(), c->nblk * sizeof(BlockMeta), hipMemcpyHostToDevice));
    if (int e = dalloc(&c->part_state, npart)) return e;
    if (int e = dalloc(&c->part_steps, npart)) return e;
    if (int e = dalloc(&c->part_last, npart)) return e;
    if (int e = dalloc(&c->part_best, npart)) return e;
    if (int e = dalloc(&c->part_best_step, npart)) return e;
    return build_tiles(c, c->ti, c->tj);
}

extern "C" int cffc_run_steady(cffc_ctx *c, int max_steps, double tol, double freeze_level, int flags,
                               int32_t *part_steps, double *part_l2, int *steps_done) {
    if (!c || max_steps < 1 || !(tol > 0.0)) return fail(CFFC_E_ARG, "bad argument");
    const bool history = flags & CFFC_STEADY_HISTORY, resume = flags & CFFC_STEADY_RESUME;
    if (history && resume) return fail(CFFC_E_ARG, "history is kept for runs that are not resumed");
    if (resume && c->npart == 0) return fail(CFFC_E_ARG, "nothing to resume");
    if (steps_done) *steps_done = 0;
    if (c->npart == 0) {
        const int32_t fb[2] = {0, c->nblk};
        if (int e = cffc_set_parts(c, 1, fb)) return e;
    }
    if (int e = ensure_steps(c, max_steps)) return e;
    if (int e = ensure_phi(c)) return e;
    if (c->part_l2) { (void)hipFree(c->part_l2); c->part_l2 = nullptr; }
    c->part_l2_steps = 0;
    if (history) {
        if (int e = dalloc(&c->part_l2, (size_t)max_steps * c->npart)) return e;
        c->part_l2_steps = max_steps;
    }
    if (!resume) {
        HIPCHK(hipMemsetAsync(c->part_state, 0, c->npart * sizeof(int32_t), c->stream));       // PART_RUNNING
        HIPCHK(hipMemsetAsync(c->part_steps, 0xFF, c->npart * sizeof(int32_t), c->stream));    // -1
        HIPCHK(hipMemsetAsync(c->part_last, 0, c->npart * sizeof(double), c->stream));
        // best L2 so far: 0x7F7F...7F = 1.4e306, above any norm
        HIPCHK(hipMemsetAsync(c->part_best, 0x7F, c->npart * sizeof(double), c->stream));
        HIPCHK(hipMemsetAsync(c->part_best_step, 0, c->npart * sizeof(int32_t), c->stream));
        c->steady_done = 0;
    } else {
        // a resumed run whose parts have all converged has nothing to do
        std::vector<int32_t> st(c->npart);
        HIPCHK(hipMemcpyAsync(st.data(), c->part_state, c->npart * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        bool all = true;
        for (int32_t v : st) all = all && v == PART_DONE;
        if (all) {
            if (part_steps) HIPCHK(hipMemcpy(part_steps, c->part_steps, c->npart * sizeof(int32_t), hipMemcpyDeviceToHost));
            if (part_l2) HIPCHK(hipMemcpy(part_l2, c->part_last, c->npart * sizeof(double), hipMemcpyDeviceToHost));
            return 0;
        }
    }
    c->freeze = 0;
    if (int e = begin_run(c, max_steps)) return e;
    const int wstart = c->wcur;
    const int N = c->s.n_stage;
    // two steps (dt slot parity 0, 1) captured once and replayed
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    const bool tsave = c->timing;
    c->timing = false;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int e = enqueue_step(c, 0, tol, true, freeze_level);
    if (!e) e = enqueue_step(c, 1, tol, true, freeze_level);
    hipError_t ce = hipStreamEndCapture(c->stream, &g);
    c->timing = tsave;
    c->wcur = wstart;  // a pair of steps flips the W buffers 2*N times
    if (e) { if (g) (void)hipGraphDestroy(g); return e; }
    if (ce != hipSuccess) return fail(CFFC_E_HIP, "graph capture failed: %s", hipGetErrorString(ce));
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(CFFC_E_HIP, "graph instantiate failed");
    }
    int32_t *hflag = nullptr;
    HIPCHK(hipHostMalloc((void **)&hflag, sizeof(int32_t), hipHostMallocDefault));
    hflag[0] = 0;
    const int pairs = max_steps / 2;
    int rc = 0;
    for (int q = 0; q < pairs; ++q) {
        if (hipGraphLaunch(ge, c->stream) != hipSuccess) { rc = fail(CFFC_E_HIP, "graph launch failed"); break; }
        // poll the stop flag every 16 pairs without draining the queue
        if ((q & 15) == 15) {
            if (hflag[0]) break;
            (void)hipMemcpyAsync(hflag, c->stop_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
        }
    }
    if (!rc && (max_steps & 1)) rc = enqueue_step(c, 0, tol, true, freeze_level);
    // the single step flipped wcur N times; end_run recomputes it from the count
    (void)hipStreamSynchronize(c->stream);
    (void)hipHostFree(hflag);
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    c->wcur = wstart;
    int done = 0;
    rc = end_run(c, wstart, max_steps, nullptr, nullptr, &done);
    c->steady_done += done;
    if (steps_done) *steps_done = done;
    if (part_steps) HIPCHK(hipMemcpy(part_steps, c->part_steps, c->npart * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (part_l2) HIPCHK(hipMemcpy(part_l2, c->part_last, c->npart * sizeof(double), hipMemcpyDeviceToHost));
    (void)N;
    return rc;
}

extern "C" int cffc_num_parts(const cffc_ctx *c) { return c ? c->npart : CFFC_E_ARG; }

extern "C" int cffc_set_stall_freeze(cffc_ctx *c, int window) {
    if (!c || window < 0) return fail(CFFC_E_ARG, "bad argument");
    c->stall_window = window;
    return 0;
}

extern "C" int cffc_step_times(cffc_ctx *c, int nsteps, double *seconds) {
    if (!c || !seconds || nsteps < 0) return fail(CFFC_E_ARG, "bad argument");
    if (nsteps > c->max_steps) return fail(CFFC_E_ARG, "only %d steps are recorded", c->max_steps);
    std::vector<unsigned long long> t((size_t)nsteps + 1);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(t.data(), c->step_clock, t.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    if (khz <= 0) return fail(CFFC_E_HIP, "device wall clock rate unknown");
    for (int n = 0; n < nsteps; ++n) seconds[n] = (double)(t[n + 1] - t[0]) / (1e3 * khz);
    return 0;
}

// ---- tile autotuner -------------------------------------------------------------
// Times nsteps fused steps with every instantiated tile shape on the
// context's own mesh and state, restores the state after every trial and
// keeps the fastest shape.
extern "C" int cffc_autotune_tile(cffc_ctx *c, int nsteps, int *tile_i, int *tile_j, double *ms_per_step,
                                  int *ncand) {
    if (!c || nsteps < 1) return fail(CFFC_E_ARG, "bad argument");
    static const int cand[][2] = {
#define T(A, B) {A, B},
        CFFC_TILE_LIST(T)
#undef T
    };
    const int nc = (int)(sizeof cand / sizeof cand[0]);
    const size_t n4 = 4 * (size_t)c->ncp;
    double *sW = nullptr, *sU = nullptr, *sS = nullptr, *sdt = nullptr, *sphi = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    const int w0 = c->wcur, ti0 = c->ti, tj0 = c->tj;
    const bool tsave = c->timing;
    int best = -1;
    double best_ms = 0.0;
    auto snap = [&](bool save) -> hipError_t {
        hipError_t e = hipSuccess;
        const hipMemcpyKind k = hipMemcpyDeviceToDevice;
        if (save) {
            if ((e = hipMemcpyAsync(sW, c->W[w0], n4 * sizeof(double), k, c->stream))) return e;
            if ((e = hipMemcpyAsync(sU, c->U, n4 * sizeof(double), k, c->stream))) return e;
            if ((e = hipMemcpyAsync(sS, c->dUdt[0], n4 * sizeof(double), k, c->stream))) return e;
            if ((e = hipMemcpyAsync(sdt, c->dt, c->ncp * sizeof(double), k, c->stream))) return e;
            if (sphi) e = hipMemcpyAsync(sphi, c->phi, n4 * sizeof(double), k, c->stream);
        } else {
            if ((e = hipMemcpyAsync(c->W[w0], sW, n4 * sizeof(double), k, c->stream))) return e;
            if ((e = hipMemcpyAsync(c->U, sU, n4 * sizeof(double), k, c->stream))) return e;
            if ((e = hipMemcpyAsync(c->dUdt[0], sS, n4 * sizeof(double), k, c->stream))) return e;
            if ((e = hipMemcpyAsync(c->dt, sdt, c->ncp * sizeof(double), k, c->stream))) return e;
            if (sphi) e = hipMemcpyAsync(c->phi, sphi, n4 * sizeof(double), k, c->stream);
            c->wcur = w0;
            c->sf_live = false;   // dUdt[0] is the snapshot again
        }
        return e;
    };
    if ((rc = dalloc(&sW, n4)) || (rc = dalloc(&sU, n4)) || (rc = dalloc(&sS, n4)) || (rc = dalloc(&sdt, c->ncp)) ||
        (c->phi && (rc = dalloc(&sphi, n4))))
        goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(CFFC_E_HIP, "hipEventCreate failed");
        goto done;
    }
    c->timing = false;
    if (snap(true) != hipSuccess) { rc = fail(CFFC_E_HIP, "state snapshot failed"); goto done; }
    for (int k = 0; k < nc; ++k) {
        float ms = 0.0f;
        if ((rc = cffc_set_tile(c, cand[k][0], cand[k][1]))) goto done;
        if (snap(false) != hipSuccess) { rc = fail(CFFC_E_HIP, "state restore failed"); goto done; }
        if ((rc = begin_run(c, nsteps + 1))) goto done;
        if ((rc = enqueue_step(c, 0, 0.0))) goto done;   // untimed first step
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(CFFC_E_HIP, "event record failed"); goto done; }
        for (int n = 1; n <= nsteps; ++n)
            if ((rc = enqueue_step(c, n & 1, 0.0))) goto done;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(CFFC_E_HIP, "trial timing failed");
            goto done;
        }
        const double per = ms / nsteps;
        if (ms_per_step) ms_per_step[k] = per;
        if (best < 0 || per < best_ms) { best = k; best_ms = per; }
    }
done:
    c->timing = tsave;
    if (best >= 0 && !rc) {
        rc = cffc_set_tile(c, cand[best][0], cand[best][1]);
        if (!rc && snap(false) != hipSuccess) rc = fail(CFFC_E_HIP, "state restore failed");
        if (!rc) {
            if (tile_i) *tile_i = cand[best][0];
            if (tile_j) *tile_j = cand[best][1];
        }
    } else if (sW) {
        (void)cffc_set_tile(c, ti0, tj0);
        (void)snap(false);
    }
    (void)hipStreamSynchronize(c->stream);
    if (ncand) *ncand = nc;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (double *q : {sW, sU, sS, sdt, sphi})
        if (q) (void)hipFree(q);
    return rc;
}

extern "C" int cffc_get_part_history(cffc_ctx *c, int nsteps, double *out) {
    if (!c || !out || nsteps < 0) return fail(CFFC_E_ARG, "bad argument");
    if (!c->part_l2 || nsteps > c->part_l2_steps) return fail(CFFC_E_ARG, "no history of %d steps kept", nsteps);
    if (nsteps) HIPCHK(hipMemcpy(out, c->part_l2, (size_t)nsteps * c->npart * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// keep the last stage's limiters stored so that a freeze can follow a run
extern "C" int cffc_set_store_limiters(cffc_ctx *c, int on) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    c->store_phi = on ? 1 : 0;
    return 0;
}

extern "C" int cffc_set_kernel_timing(cffc_ctx *c, int on) {
    if (!c) return fail(CFFC_E_ARG, "null context");
    c->timing = on != 0;
    c->t_ms[0] = c->t_ms[1] = 0;
    c->t_n[0] = c->t_n[1] = 0;
    c->ev_used = 0;
    c->ev_kind.clear();
    return 0;
}

extern "C" int cffc_kernel_times(cffc_ctx *c, double out[4]) {
    if (!c || !out) return fail(CFFC_E_ARG, "null argument");
    out[0] = c->t_ms[1];
    out[1] = (double)c->t_n[1];
    out[2] = c->t_ms[0];
    out[3] = (double)c->t_n[0];
    return 0;
}

// ---- FAS multigrid (FASMultigrid2D/FASMultigrid2D.h) -------------------------
// One context per level on one stream, level 0 the finest; level l+1 halves
// level l block by block.  The cycle is the reference's Coarse_Grid_Correction
// from a top level (0, or a coarser one during the full-multigrid start-up),
// every smoothing stage through the granular FP64 stage kernel (HLLE and the
// zero limiter below the top level when first_order_coarse is set), the
// forcing term folded into that kernel after the face fluxes, and the level
// transfers of euler2d_mg.hip.  Per-cell dt on every level (Smooth's CFL call
// does not apply the global minimum).
struct cffc_mg {
    std::vector<cffc_ctx *> lv;
    cffc_mg_desc d{};
    std::vector<double *> P, uo;     // per level (P unused on level 0)
    std::vector<int> max_edge;       // per level: max ni + nj over blocks
    std::vector<double> Ahost;       // scratch: fine-level areas for the Wo restriction
    double rs_eps = 0.0;
    int rs_iters = 0;
    bool frozen = false;
    // the smoother's scheme (top level) and the one below the top level
    cffc_scheme_desc fine_s{}, coarse_s{};
    // Execute replays one captured cycle (kernels only); recaptured when the
    // limiters freeze (the stage kernels' phi mode changes) or the top level
    // changes
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_frozen = -1, graph_top = -1;
    // per-case stopping (cffc_mg_set_parts): part q owns level-0 blocks
    // part_block0[q] .. part_block0[q+1]-1; per part its tolerance, freeze
    // level, last checked L2 norm and the regular cycles done when its norm
    // first met the tolerance (-1 before); cycles counts the regular cycles
    // since cffc_mg_set_parts
    int npart = 0;
    std::vector<int32_t> part_block0, part_done, part_frozen;
    std::vector<double> part_tol, part_freeze, part_l2;
    int32_t *d_tile0 = nullptr, *d_part_of_block = nullptr, *d_part_frozen = nullptr;
    double *d_part_l2 = nullptr;
    long long cycles = 0;
};

// the flux function and limiter of every level for a top level: the
// smoother's on the top level and above, HLLE and the zero limiter below it
// when first_order_coarse is set (dUdt_Multistage_Explicit_for_Multigrid
// :2502-2509, Level > Top_Level)
static void mg_set_top(cffc_mg *mg, int top) {
    for (int l = 0; l < (int)mg->lv.size(); ++l) {
        const cffc_scheme_desc &s = l > top ? mg->coarse_s : mg->fine_s;
        mg->lv[l]->s.i_flux_function = s.i_flux_function;
        mg->lv[l]->s.i_limiter = s.i_limiter;
    }
}

static void mg_free(cffc_mg *mg) {
    if (!mg) return;
    for (int32_t *q : {mg->d_tile0, mg->d_part_of_block, mg->d_part_frozen})
        if (q) (void)hipFree(q);
    if (mg->d_part_l2) (void)hipFree(mg->d_part_l2);
    if (mg->gexec) (void)hipGraphExecDestroy(mg->gexec);
    if (mg->graph) (void)hipGraphDestroy(mg->graph);
    for (size_t l = 0; l < mg->lv.size(); ++l) {
        if (l < mg->P.size() && mg->P[l]) (void)hipFree(mg->P[l]);
        if (l < mg->uo.size() && mg->uo[l]) (void)hipFree(mg->uo[l]);
    }
    // coarse levels share the finest level's stream: free them first
    for (size_t l = mg->lv.size(); l-- > 0;)
        if (mg->lv[l]) free_ctx(mg->lv[l]);
    delete mg;
}

extern "C" void cffc_mg_destroy(cffc_mg *mg) { mg_free(mg); }

extern "C" cffc_ctx *cffc_mg_level(cffc_mg *mg, int level) {
    if (!mg || level < 0 || level >= (int)mg->lv.size()) return nullptr;
    return mg->lv[level];
}

extern "C" int cffc_mg_create(const cffc_mesh_desc *levels, int nlevels, const cffc_scheme_desc *smoother,
                              const cffc_mg_desc *d, cffc_mg **out) {
    if (!levels || !smoother || !d || !out) return fail(CFFC_E_ARG, "null argument");
    *out = nullptr;
    if (nlevels < 2 || nlevels != d->levels) return fail(CFFC_E_ARG, "levels (%d) must be >= 2 and match the meshes (%d)",
                                                         d->levels, nlevels);
    if (d->i_cycle != CFFC_MULTIGRID_V_CYCLE && d->i_cycle != CFFC_MULTIGRID_W_CYCLE)
        return fail(CFFC_E_ARG, "unknown multigrid cycle %d", d->i_cycle);
    if (smoother->flux_precision != CFFC_FLUX_FP64) return fail(CFFC_E_ARG, "multigrid runs the FP64 path only");
    if (d->smooths_on_finest_level < 0 || d->pre_smooths < 0 || d->post_smooths < 0 ||
        d->smooths_on_coarsest_level < 0 || d->max_update_reductions < 0)
        return fail(CFFC_E_ARG, "negative smoothing or reduction count");
    for (int l = 0; l < nlevels; ++l) {
        if (levels[l].level || levels[l].nbr2) return fail(CFFC_E_ARG, "multigrid levels must be single-level meshes");
        if (levels[l].nblk != levels[0].nblk) return fail(CFFC_E_ARG, "level %d has a different block count", l);
        for (int b = 0; b < levels[0].nblk && l > 0; ++b) {
            if (levels[l - 1].ni[b] != 2 * levels[l].ni[b] || levels[l - 1].nj[b] != 2 * levels[l].nj[b])
                return fail(CFFC_E_ARG, "block %d of level %d does not halve level %d", b, l, l - 1);
            // Half_Mesh_Resolution (HO_Grid2DQuad.cc:12618-12621): a coarse
            // block keeps at least Nghost cells a side
            if (levels[l].ni[b] < levels[l].nghost || levels[l].nj[b] < levels[l].nghost)
                return fail(CFFC_E_ARG, "block %d of level %d has fewer than %d cells a side", b, l,
                            levels[l].nghost);
        }
    }
    cffc_mg *mg = new cffc_mg();
    mg->d = *d;
    mg->fine_s = *smoother;
    mg->coarse_s = *smoother;
    if (d->first_order_coarse) {
        mg->coarse_s.i_flux_function = CFFC_FLUX_FUNCTION_HLLE;
        mg->coarse_s.i_limiter = CFFC_LIMITER_ZERO;
    }
    mg->lv.assign(nlevels, nullptr);
    mg->P.assign(nlevels, nullptr);
    mg->uo.assign(nlevels, nullptr);
    mg->max_edge.assign(nlevels, 0);
    for (int l = 0; l < nlevels; ++l) {
        if (int e = cffc_create(&levels[l], smoother, &mg->lv[l])) {
            mg_free(mg);
            return e;
        }
        cffc_ctx *c = mg->lv[l];
        if (l > 0) {
            (void)hipStreamDestroy(c->stream);
            c->stream = mg->lv[0]->stream;
            c->own_stream = false;
        }
        int e;
        if ((e = ensure_granular(c)) || (e = ensure_phi(c))) {
            mg_free(mg);
            return e;
        }
        const size_t n = 4 * (size_t)c->ncp;
        if ((e = dalloc(&mg->uo[l], n)) || (l > 0 && (e = dalloc(&mg->P[l], n)))) {
            mg_free(mg);
            return e;
        }
        if (l > 0 && hipMemsetAsync(mg->P[l], 0, n * sizeof(double), c->stream) != hipSuccess) {
            mg_free(mg);
            return fail(CFFC_E_HIP, "hipMemsetAsync failed");
        }
        for (int b = 0; b < c->nblk; ++b) mg->max_edge[l] = std::max(mg->max_edge[l], c->meta_h[b].ni + c->meta_h[b].nj);
    }
    mg_set_top(mg, 0);
    if (hipStreamSynchronize(mg->lv[0]->stream) != hipSuccess) {
        mg_free(mg);
        return fail(CFFC_E_HIP, "multigrid set-up failed");
    }
    *out = mg;
    return 0;
}

extern "C" int cffc_mg_set_residual_smoothing(cffc_mg *mg, double epsilon, int iterations) {
    if (!mg || iterations < 0 || !(epsilon >= 0.0)) return fail(CFFC_E_ARG, "bad argument");
    mg->rs_eps = epsilon;
    mg->rs_iters = iterations;
    return 0;
}

// Exchange_Solution_Information + BCs of one level (U sent, W = W(U))
static int mg_exchange(cffc_ctx *c) {
    ExchangeParams p = exchange_params(c, 1, nullptr);
    p.store_U = 1;
    HIPCHK(cffc_launch_exchange(&p, c->nblk, c->stream));
    return 0;
}

static const double *area(const cffc_ctx *c) { return c->geom + (size_t)GEO_A * c->ncp; }

// one granular FP64 stage kernel: a smoothing stage (residual = 0) or
// dUdt_Residual_Evaluation into dUdt[0] (residual = 1); P: forcing term or null
static int mg_stage(cffc_ctx *c, int i_stage, int residual, const double *P) {
    StageParams p = base_stage_params(c);
    p.Win = c->W[c->wcur];
    p.Wout = c->W[c->wcur];
    const int k = residual ? 0 : k_residual(c->s, i_stage);
    p.S = c->dUdt[k];
    p.S1 = c->dUdt[1];
    p.S2 = c->dUdt[2];
    p.mode = residual ? GINIT_ZERO : granular_init(c->s, i_stage);
    p.dtmin_in = c->scal;
    p.lts = 1;
    p.phi_mode = c->freeze ? PHI_LOAD : PHI_COMPUTE_STORE;
    p.part_of_block = c->mg_part_of_block;
    p.part_frozen = c->mg_part_frozen;
    p.hlle = c->s.i_flux_function == CFFC_FLUX_FUNCTION_HLLE;
    p.residual = residual;
    p.P = P;
    HIPCHK(cffc_launch_stage_fp64(&p, c->ti, c->tj, 0, c->stream));
    return 0;
}

// Smooth :2865-3110
static int mg_smooth(cffc_mg *mg, int top, int l) {
    cffc_ctx *c = mg->lv[l];
    // local time steps only (the levels' stages read dt per cell): no global
    // minimum is formed
    HIPCHK(cffc_launch_cfl(c->meta, c->tiles, c->ntiles, c->ti, c->tj, c->W[c->wcur], c->geom, c->ncp, c->dt, nullptr,
                           c->stream));
    if (l > top) {
        cffc_ctx *f = mg->lv[l - 1];
        HIPCHK(cffc_launch_mg_cfl(f->meta, c->meta, c->nblk, c->max_int, f->dt, area(f), c->dt, area(c), c->stream));
    }
    for (int st = 1; st <= c->s.n_stage; ++st) {
        if (int e = mg_exchange(c)) return e;
        if (st == 1) HIPCHK(cffc_launch_copy_interior(c->meta, c->nblk, c->max_int, c->U, c->Uo, c->ncp, c->stream));
        if (int e = mg_stage(c, st, 0, l > top ? mg->P[l] : nullptr)) return e;
        if (mg->rs_iters > 0)
            HIPCHK(cffc_launch_residual_smoothing(c->meta, c->nblk, c->dUdt[0], c->dUdt[1], c->ncp, mg->rs_eps,
                                                  mg->rs_iters, nullptr, nullptr, c->stream));
        const int k = k_residual(c->s, st);
        HIPCHK(cffc_launch_update(c->meta, c->nblk, c->max_int, c->W[c->wcur], c->U, c->Uo, c->dUdt[k], c->dt, c->ncp,
                                  stage_omega(c->s, st),
                                  c->s.local_time_stepping == CFFC_SCALAR_LOCAL_TIME_STEPPING, c->scal + 2,
                                  c->stream));
    }
    if (int e = mg_exchange(c)) return e;
    if (int e = mg_stage(c, 1, 1, l > top ? mg->P[l] : nullptr)) return e;
    if (!c->freeze) c->phi_valid = 1;
    return 0;
}

static int mg_store_uo(cffc_mg *mg, int l) {
    cffc_ctx *c = mg->lv[l];
    HIPCHK(cffc_launch_mg_copy(c->meta, c->nblk, c->max_cells, mg->uo[l], c->U, c->ncp, c->stream));
    return 0;
}

// Coarse_Grid_Correction :3118-3370
static int mg_cgc(cffc_mg *mg, int top, int l) {
    const cffc_mg_desc &d = mg->d;
    const int L = (int)mg->lv.size();
    cffc_ctx *c = mg->lv[l];
    int e;
    if (l == L - 1) {
        if ((e = mg_store_uo(mg, l))) return e;
        for (int nu = 1; nu <= d.smooths_on_coarsest_level; ++nu)
            if ((e = mg_smooth(mg, top, l))) return e;
    } else {
        cffc_ctx *cc = mg->lv[l + 1];
        // the W cycle's two passes apply below the finest level, whatever
        // the top level (:3151-3158)
        const int m_cycle = (d.i_cycle == CFFC_MULTIGRID_V_CYCLE || l == 0) ? 1 : 2;
        for (int m = 1; m <= m_cycle; ++m) {
            if ((e = mg_store_uo(mg, l))) return e;
            const int n_pre = l == top ? d.smooths_on_finest_level : d.pre_smooths;
            for (int nu = 0; nu < n_pre; ++nu)
                if ((e = mg_smooth(mg, top, l))) return e;
            HIPCHK(cffc_launch_mg_restrict(c->meta, cc->meta, c->nblk, cc->max_int, c->U, area(c), c->ncp, cc->U,
                                           area(cc), cc->ncp, c->stream));
            HIPCHK(cffc_launch_mg_primitive(cc->meta, cc->nblk, cc->max_int, cc->U, cc->W[cc->wcur], cc->ncp,
                                            c->stream));
            if ((e = mg_exchange(cc))) return e;
            if (d.defect_correction) {
                HIPCHK(cffc_launch_mg_restrict(c->meta, cc->meta, c->nblk, cc->max_int, c->dUdt[0], area(c), c->ncp,
                                               mg->P[l + 1], area(cc), cc->ncp, c->stream));
                if ((e = mg_stage(cc, 1, 1, nullptr))) return e;
                HIPCHK(cffc_launch_mg_subtract(cc->meta, cc->nblk, cc->max_int, mg->P[l + 1], cc->dUdt[0], cc->ncp,
                                               c->stream));
            }
            if ((e = mg_cgc(mg, top, l + 1))) return e;
            HIPCHK(cffc_launch_mg_prolong(cc->meta, c->meta, c->nblk, cc->max_int, mg->max_edge[l + 1], cc->bc, cc->U,
                                          cc->geom, cc->ncp, c->U, c->geom, c->ncp, !d.prolong_using_injection,
                                          d.injection_at_dirichlet, d.prolong_using_injection,
                                          d.update_stability_switch, d.max_update_reductions, 0, c->stream));
            HIPCHK(cffc_launch_mg_primitive(c->meta, c->nblk, c->max_int, c->U, c->W[c->wcur], c->ncp, c->stream));
            if ((e = mg_exchange(c))) return e;
            const int n_post = l == top ? 0 : d.post_smooths;
            for (int nu = 0; nu < n_post; ++nu)
                if ((e = mg_smooth(mg, top, l))) return e;
        }
    }
    if (l != top) HIPCHK(cffc_launch_mg_changes(c->meta, c->nblk, c->max_cells, c->U, mg->uo[l], c->ncp, c->stream));
    return 0;
}

// Restrict_Boundary_Ref_States :1112-1186 on the host (the fine areas read
// back once): coarse WoN/S/E/W over the interior span
static int mg_restrict_wo(cffc_mg *mg, int lf) {
    cffc_ctx *f = mg->lv[lf], *c = mg->lv[lf + 1];
    std::vector<double> Af(f->ncp);
    HIPCHK(hipMemcpyAsync(Af.data(), area(f), f->ncp * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    const size_t nwf = 4 * (size_t)(f->meta_h.back().bcoff + 2 * f->meta_h.back().ncj + 2 * f->meta_h.back().nci);
    const size_t nwc = 4 * (size_t)(c->meta_h.back().bcoff + 2 * c->meta_h.back().ncj + 2 * c->meta_h.back().nci);
    std::vector<double> wf(nwf), wc(nwc);
    HIPCHK(hipMemcpyAsync(wf.data(), f->wo, nwf * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipMemcpyAsync(wc.data(), c->wo, nwc * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    for (int b = 0; b < c->nblk; ++b) {
        const BlockMeta &F = f->meta_h[b], &C = c->meta_h[b];
        auto A = [&](int i, int j) { return Af[F.coff + (int64_t)j * F.pitch + i]; };
        // entry (side, k) -> AoS offset; sides W, E (ncj entries), S, N (nci)
        auto fo = [&](int side, int k) {
            const int base = side == 0 ? 0 : side == 1 ? F.ncj : side == 2 ? 2 * F.ncj : 2 * F.ncj + F.nci;
            return 4 * (size_t)(F.bcoff + base + k);
        };
        auto co = [&](int side, int k) {
            const int base = side == 0 ? 0 : side == 1 ? C.ncj : side == 2 ? 2 * C.ncj : 2 * C.ncj + C.nci;
            return 4 * (size_t)(C.bcoff + base + k);
        };
        auto avg = [&](int side_f, int k0, double A0, double A1, int side_c, int kc) {
            const double *w0 = &wf[fo(side_f, k0)], *w1 = &wf[fo(side_f, k0 + 1)];
            double *o = &wc[co(side_c, kc)];
            for (int v = 0; v < 4; ++v) o[v] = (w0[v] * A0 + w1[v] * A1) / (A0 + A1);
        };
        const int JCl = 2, JCu = F.nj + 1, ICl = 2, ICu = F.ni + 1;
        for (int ic = 2; ic <= C.ni + 1; ++ic) {
            const int i0 = 2 * (ic - 2) + 2;
            avg(3, i0, A(i0, JCu), A(i0 + 1, JCu), 3, ic);
            avg(2, i0, A(i0, JCl), A(i0 + 1, JCl), 2, ic);
        }
        for (int jc = 2; jc <= C.nj + 1; ++jc) {
            const int j0 = 2 * (jc - 2) + 2;
            avg(0, j0, A(ICl, j0), A(ICl, j0 + 1), 0, jc);
            avg(1, j0, A(ICu, j0), A(ICu, j0 + 1), 1, jc);
        }
    }
    HIPCHK(hipMemcpy(c->wo, wc.data(), nwc * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// The start of Execute :3395-3420 for a state set on level 0: coarse levels
// restricted from the level above (the IC_RESTART path of Apply_ICs
// :719-733) with their boundary reference states, then messages and BCs.
// a new start of a batch with parts: no case converged, frozen or skipped
static int mg_restart_parts(cffc_mg *mg) {
    for (cffc_ctx *c : mg->lv) {
        bool any = false;
        for (BlockMeta &m : c->meta_h) {
            any = any || m.skip;
            m.skip = 0;
        }
        if (any) HIPCHK(hipMemcpy(c->meta, c->meta_h.data(), c->nblk * sizeof(BlockMeta), hipMemcpyHostToDevice));
    }
    if (!mg->npart) return 0;
    mg->part_done.assign(mg->npart, -1);
    mg->part_l2.assign(mg->npart, -1.0);
    mg->part_frozen.assign(mg->npart, 0);
    HIPCHK(hipMemcpy(mg->d_part_frozen, mg->part_frozen.data(), mg->npart * sizeof(int32_t), hipMemcpyHostToDevice));
    mg->cycles = 0;
    return 0;
}

extern "C" int cffc_mg_initialize(cffc_mg *mg) {
    if (!mg) return fail(CFFC_E_ARG, "null multigrid solver");
    if (int e = mg_restart_parts(mg)) return e;
    if (int e = mg_exchange(mg->lv[0])) return e;
    for (size_t l = 1; l < mg->lv.size(); ++l) {
        cffc_ctx *f = mg->lv[l - 1], *c = mg->lv[l];
        HIPCHK(cffc_launch_mg_restrict(f->meta, c->meta, c->nblk, c->max_int, f->U, area(f), f->ncp, c->U, area(c),
                                       c->ncp, c->stream));
        HIPCHK(cffc_launch_mg_primitive(c->meta, c->nblk, c->max_int, c->U, c->W[c->wcur], c->ncp, c->stream));
        if (int e = mg_restrict_wo(mg, (int)l - 1)) return e;
        if (int e = mg_exchange(c)) return e;
    }
    HIPCHK(hipStreamSynchronize(mg->lv[0]->stream));
    return 0;
}

static int mg_reset_errors(cffc_mg *mg) {
    for (cffc_ctx *c : mg->lv) HIPCHK(cffc_launch_set_u64(c->scal + 2, ERR_NONE, 0, 1, c->stream));
    return 0;
}
// the first failing update of the last cycle, finest level first: level in
// bits 56-63 of *bad_cell, then the update kernel's block/j/i key
static int mg_check_errors(cffc_mg *mg, int64_t *bad_cell) {
    std::vector<unsigned long long> keys(mg->lv.size());
    for (size_t l = 0; l < mg->lv.size(); ++l)
        HIPCHK(hipMemcpyAsync(&keys[l], mg->lv[l]->scal + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              mg->lv[0]->stream));
    HIPCHK(hipStreamSynchronize(mg->lv[0]->stream));
    for (size_t l = 0; l < keys.size(); ++l)
        if (keys[l] != ERR_NONE) {
            if (bad_cell) *bad_cell = (int64_t)(keys[l] | ((unsigned long long)l << 56));
            return (int)(keys[l] & 0xFFFF);
        }
    if (bad_cell) *bad_cell = -1;
    return 0;
}

// one FAS cycle from level 0 (Coarse_Grid_Correction(0, 0)), synchronous
extern "C" int cffc_mg_cycle(cffc_mg *mg, int64_t *bad_cell) {
    if (!mg) return fail(CFFC_E_ARG, "null multigrid solver");
    mg_set_top(mg, 0);
    if (int e = mg_reset_errors(mg)) return e;
    if (int e = mg_cgc(mg, 0, 0)) return e;
    return mg_check_errors(mg, bad_cell);
}

extern "C" int cffc_mg_set_graph(cffc_mg *mg, int on) {
    if (!mg) return fail(CFFC_E_ARG, "null multigrid solver");
    mg->use_graph = on != 0;
    return 0;
}

// enqueue one cycle from a top level: a replay of the captured cycle,
// captured first (and again after the limiters freeze or the top level
// changes), or the launches themselves
static int mg_enqueue_cycle(cffc_mg *mg, int top) {
    cffc_ctx *f = mg->lv[0];
    if (!mg->use_graph) {
        if (int e = mg_reset_errors(mg)) return e;
        return mg_cgc(mg, top, top);
    }
    if (!mg->gexec || mg->graph_frozen != (int)mg->frozen || mg->graph_top != top) {
        if (mg->gexec) (void)hipGraphExecDestroy(mg->gexec);
        if (mg->graph) (void)hipGraphDestroy(mg->graph);
        mg->gexec = nullptr;
        mg->graph = nullptr;
        HIPCHK(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
        int e = mg_reset_errors(mg);
        if (!e) e = mg_cgc(mg, top, top);
        hipError_t ce = hipStreamEndCapture(f->stream, &mg->graph);
        if (e) return e;
        if (ce != hipSuccess) return fail(CFFC_E_HIP, "multigrid cycle capture failed: %s", hipGetErrorString(ce));
        HIPCHK(hipGraphInstantiate(&mg->gexec, mg->graph, nullptr, nullptr, 0));
        mg->graph_frozen = (int)mg->frozen;
        mg->graph_top = top;
    }
    HIPCHK(hipGraphLaunch(mg->gexec, f->stream));
    return 0;
}

extern "C" int cffc_mg_get_P(cffc_mg *mg, int level, double *P, int64_t ncell) {
    if (!mg || level < 1 || level >= (int)mg->lv.size()) return fail(CFFC_E_ARG, "bad level");
    return get_field(mg->lv[level], mg->P[level], P, ncell, 4);
}

// Execute :3375-3745: with fmg_cycles > 0 the full-multigrid start-up first
// (top level Levels-2 up to 1, fmg_cycles cycles each, stopping early on the
// FMG tolerances, then Prolong_Solution_Blocks to the next level with
// messages and BCs), then the regular cycles from level 0.  Before each
// cycle: the residual norms of the top level (the dUdt[0] the last
// smoothing left) and every level's error key, read with one
// synchronisation; the relative norm against the level's second check;
// limiter freezing at freeze_level (<= 0: never; unfrozen again at a new
// top level, :3462-3476); and the exits: max_cycles cycles in all, or the
// absolute / relative tolerance (the FMG ones above level 0).  A spent
// budget still prolongs through the remaining start-up levels, each
// stopping at its first check.  norms: 3 per check.
// the blocks of part q on every level: skipped by the cycle's kernels (a
// converged case) or not; the changed metadata go to the device in stream
// order
static int mg_mark_part(cffc_mg *mg, int q, int skip) {
    const int b0 = mg->part_block0[q], nb = mg->part_block0[q + 1] - b0;
    for (cffc_ctx *c : mg->lv) {
        for (int b = b0; b < b0 + nb; ++b) c->meta_h[b].skip = skip;
        HIPCHK(hipMemcpyAsync(c->meta + b0, c->meta_h.data() + b0, nb * sizeof(BlockMeta), hipMemcpyHostToDevice,
                              c->stream));
    }
    return 0;
}

extern "C" int cffc_mg_set_parts(cffc_mg *mg, int npart, const int32_t *first_block, const double *tol,
                                 const double *freeze) {
    if (!mg) return fail(CFFC_E_ARG, "null argument");
    cffc_ctx *f = mg->lv[0];
    for (cffc_ctx *c : mg->lv) {
        c->mg_part_of_block = nullptr;
        c->mg_part_frozen = nullptr;
        bool any = false;
        for (BlockMeta &m : c->meta_h) {
            any = any || m.skip;
            m.skip = 0;
        }
        if (any) HIPCHK(hipMemcpy(c->meta, c->meta_h.data(), c->nblk * sizeof(BlockMeta), hipMemcpyHostToDevice));
    }
    if (npart == 0) {
        mg->npart = 0;
        return 0;
    }
    if (npart < 0 || npart > f->nblk || !first_block || !tol) return fail(CFFC_E_ARG, "bad argument");
    if (first_block[0] != 0 || first_block[npart] != f->nblk) return fail(CFFC_E_ARG, "parts must cover all blocks");
    std::vector<int32_t> pob(f->nblk);
    for (int q = 0; q < npart; ++q) {
        if (first_block[q + 1] <= first_block[q]) return fail(CFFC_E_ARG, "part %d is empty", q);
        if (!(tol[q] > 0.0) || (freeze && !(freeze[q] >= 0.0))) return fail(CFFC_E_ARG, "bad tolerance of part %d", q);
        for (int b = first_block[q]; b < first_block[q + 1]; ++b) pob[b] = q;
    }
    // parts must not exchange ghost cells with each other
    std::vector<int32_t> nbr_h(8 * (size_t)f->nblk);
    HIPCHK(hipMemcpy(nbr_h.data(), f->nbr, nbr_h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int b = 0; b < f->nblk; ++b)
        for (int k = 0; k < 8; ++k) {
            const int nb = nbr_h[8 * (size_t)b + k];
            if (nb >= 0 && pob[nb] != pob[b])
                return fail(CFFC_E_TOPOLOGY, "block %d neighbours block %d of another part", b, nb);
        }
    for (int32_t **q : {&mg->d_tile0, &mg->d_part_of_block, &mg->d_part_frozen})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (mg->d_part_l2) (void)hipFree(mg->d_part_l2);
    mg->d_part_l2 = nullptr;
    if (int e = dalloc(&mg->d_tile0, npart + 1)) return e;
    if (int e = dalloc(&mg->d_part_l2, npart)) return e;
    if (int e = dalloc(&mg->d_part_of_block, f->nblk)) return e;
    if (int e = dalloc(&mg->d_part_frozen, npart)) return e;
    HIPCHK(hipMemcpy(mg->d_part_of_block, pob.data(), pob.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    mg->part_frozen.assign(npart, 0);
    HIPCHK(hipMemcpy(mg->d_part_frozen, mg->part_frozen.data(), npart * sizeof(int32_t), hipMemcpyHostToDevice));
    // every level has the finest level's blocks in the same order
    for (cffc_ctx *c : mg->lv) {
        c->mg_part_of_block = mg->d_part_of_block;
        c->mg_part_frozen = mg->d_part_frozen;
    }
    mg->npart = npart;
    mg->part_block0.assign(first_block, first_block + npart + 1);
    mg->part_tol.assign(tol, tol + npart);
    mg->part_freeze.assign(npart, 0.0);
    if (freeze) mg->part_freeze.assign(freeze, freeze + npart);
    mg->part_l2.assign(npart, -1.0);
    mg->part_done.assign(npart, -1);
    mg->cycles = 0;
    return 0;
}

extern "C" int cffc_mg_part_status(cffc_mg *mg, double *l2, int32_t *done_cycles) {
    if (!mg) return fail(CFFC_E_ARG, "null argument");
    if (!mg->npart) return fail(CFFC_E_ARG, "no parts set");
    if (l2) memcpy(l2, mg->part_l2.data(), mg->npart * sizeof(double));
    if (done_cycles) memcpy(done_cycles, mg->part_done.data(), mg->npart * sizeof(int32_t));
    return mg->npart;
}

// first level-0 tile of every part (tiles are in block order), uploaded for
// the per-part norms
static int mg_part_tiles(cffc_mg *mg) {
    const cffc_ctx *f = mg->lv[0];
    std::vector<int32_t> t0(mg->npart + 1);
    int32_t nt = 0;
    int q = 0;
    for (int b = 0; b < f->nblk; ++b) {
        while (q < mg->npart && mg->part_block0[q] == b) t0[q++] = nt;
        const BlockMeta &m = f->meta_h[b];
        nt += ((m.ni + f->ti - 1) / f->ti) * ((m.nj + f->tj - 1) / f->tj);
    }
    t0[mg->npart] = nt;
    if (nt != f->ntiles) return fail(CFFC_E_ARG, "tile count %d, expected %d", nt, f->ntiles);
    HIPCHK(hipMemcpy(mg->d_tile0, t0.data(), t0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int cffc_mg_execute(cffc_mg *mg, int max_cycles, double freeze_level, double *norms, int *cycles_done,
                               int *checks) {
    if (!mg || max_cycles < 0) return fail(CFFC_E_ARG, "bad argument");
    const int L = (int)mg->lv.size();
    bool first_step = true;
    int n = 0, nchk = 0, rc = 0;
    std::vector<unsigned long long> keys(mg->lv.size());
    if (int e = mg_reset_errors(mg)) return e;
    if (mg->npart)
        if (int e = mg_part_tiles(mg)) return e;
    const int top0 = mg->d.fmg_cycles > 0 ? L - 2 : 0;
    for (int top = top0; top >= 0 && !rc; --top) {
        mg_set_top(mg, top);
        cffc_ctx *t = mg->lv[top];
        if (!first_step && freeze_level > 0 && mg->frozen && mg->fine_s.i_limiter != CFFC_LIMITER_ZERO) {
            for (int l = top; l < L; ++l) mg->lv[l]->freeze = 0;
            mg->frozen = false;
        }
        const long long ncyc = top != 0 ? mg->d.fmg_cycles : 0x7fffffffffffffffLL;
        double init_l2 = -1.0, ratio = -1.0;
        for (long long cycles = 1; cycles <= ncyc; ++cycles) {
            double nr[3];
            HIPCHK(cffc_launch_norms(t->meta, t->tiles, t->ntiles, t->ti, t->tj, t->dUdt[0], t->ncp, 1,
                                     t->norm_part, t->norm_small, t->stream));
            HIPCHK(hipMemcpyAsync(nr, t->norm_small, 3 * sizeof(double), hipMemcpyDeviceToHost, t->stream));
            const bool parts = mg->npart && top == 0;
            if (parts) {
                HIPCHK(cffc_launch_part_norms(t->norm_part, mg->d_tile0, mg->npart, mg->d_part_l2, t->stream));
                HIPCHK(hipMemcpyAsync(mg->part_l2.data(), mg->d_part_l2, mg->npart * sizeof(double),
                                      hipMemcpyDeviceToHost, t->stream));
            }
            for (size_t l = 0; l < mg->lv.size(); ++l)
                HIPCHK(hipMemcpyAsync(&keys[l], mg->lv[l]->scal + 2, sizeof(unsigned long long),
                                      hipMemcpyDeviceToHost, t->stream));
            HIPCHK(hipStreamSynchronize(t->stream));
            for (size_t l = 0; l < keys.size() && !rc; ++l)
                if (keys[l] != ERR_NONE) rc = (int)(keys[l] & 0xFFFF);
            if (rc) break;
            if (norms) memcpy(norms + 3 * nchk, nr, sizeof nr);
            nchk++;
            if (cycles == 2) init_l2 = nr[1];
            else if (cycles > 2 && fabs(init_l2) > 1e-9) ratio = nr[1] / init_l2;
            bool stop = n >= max_cycles;
            bool freeze_now = nr[1] <= freeze_level;
            if (parts) {
                // per-case stopping: every part's own norm has met its
                // tolerance at a check after the second cycle (the single
                // norm's rule); each part's limiters freeze once its own norm
                // is at or below its freeze level (the single norm's rule)
                bool all = true, newly = false;
                freeze_now = false;
                for (int q = 0; q < mg->npart; ++q) {
                    if (cycles > 2 && mg->part_done[q] < 0 && mg->part_l2[q] < mg->part_tol[q]) {
                        // converged: the cycle's kernels leave the case alone
                        // from here on, so it keeps the state it converged in
                        mg->part_done[q] = (int32_t)mg->cycles;
                        if (int e = mg_mark_part(mg, q, 1)) return e;
                    }
                    if (mg->part_done[q] < 0) all = false;
                    if (!first_step && cycles > 1 && freeze_level > 0 && !mg->part_frozen[q] &&
                        mg->fine_s.i_limiter != CFFC_LIMITER_ZERO && mg->part_l2[q] <= mg->part_freeze[q]) {
                        mg->part_frozen[q] = 1;
                        newly = true;
                    }
                }
                if (all) stop = true;
                if (newly && !stop)
                    HIPCHK(hipMemcpyAsync(mg->d_part_frozen, mg->part_frozen.data(), mg->npart * sizeof(int32_t),
                                          hipMemcpyHostToDevice, t->stream));
            } else if (cycles > 2) {
                const double at = top == 0 ? mg->d.absolute_tolerance : mg->d.fmg_absolute_tolerance;
                const double rt = top == 0 ? mg->d.relative_tolerance : mg->d.fmg_relative_tolerance;
                if (nr[1] < at || ratio < rt) stop = true;
            }
            if (stop) break;
            if (!first_step && cycles > 1 && freeze_level > 0 && !mg->frozen &&
                mg->fine_s.i_limiter != CFFC_LIMITER_ZERO && freeze_now) {
                for (int l = top; l < L; ++l) mg->lv[l]->freeze = 1;
                mg->frozen = true;
            }
            if (int e = mg_enqueue_cycle(mg, top)) return e;
            first_step = false;
            n++;
            if (top == 0) mg->cycles++;
        }
        if (top != 0 && !rc) {
            cffc_ctx *f = mg->lv[top - 1];
            HIPCHK(cffc_launch_mg_prolong(t->meta, f->meta, t->nblk, t->max_int, mg->max_edge[top], t->bc, t->U,
                                          t->geom, t->ncp, f->U, f->geom, f->ncp, 0, 0, mg->d.prolong_using_injection,
                                          0, 0, 1, t->stream));
            if (int e = mg_exchange(f)) return e;
        }
    }
    mg_set_top(mg, 0);
    if (cycles_done) *cycles_done = n;
    if (checks) *checks = nchk;
    if (int e = mg_exchange(mg->lv[0])) return e;
    HIPCHK(hipStreamSynchronize(mg->lv[0]->stream));
    return rc;
}

// Time-accurate explicit loop, Euler2DQuadSolvers.cc:526-898 (Time_Accurate
// = 1): the reference's call order through the granular entry points, one
// host synchronisation per step for the clipped global step.  *time carries
// the start time in (the restart's Time) and the reached time out; the
// reference's loop has no step cap, max_steps > 0 adds one.
extern "C" int cffc_run_time_accurate(cffc_ctx *c, double t_max, int max_steps, double freeze_level, double *time,
                                      int *steps_done) {
    if (!c || !time || !(t_max >= 0.0) || !(*time >= 0.0) || !(freeze_level >= 0.0))
        return fail(CFFC_E_ARG, "bad argument");
    if (c->s.local_time_stepping != CFFC_GLOBAL_TIME_STEPPING)
        return fail(CFFC_E_ARG, "time-accurate runs use global time stepping");
    const bool optsm = c->s.i_time_integration == CFFC_TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING;
    // MultiStage_Optimally_Smoothing(N, N, limiter) enters the clip for every
    // scheme (:668-679) but the time update only for optimal smoothing (:888-896)
    const double beta = optimal_smoothing(c->s.n_stage, c->s.n_stage, c->s.i_limiter);
    const double CFL = c->s.cfl_number;
    double Time = *time;
    int n = 0, rc = 0;
    bool first_step = true;
    while (Time < t_max) {
        if (max_steps > 0 && n >= max_steps) {
            *time = Time;
            if (steps_done) *steps_done = n;
            return fail(CFFC_E_STEP_LIMIT, "max_steps reached before t_max");
        }
        double dTime;
        if (int e = cffc_cfl(c, &dTime, 0)) return e;
        // :662-680, the reference's if / else-if as written
        if (!optsm && Time + CFL * dTime > t_max) dTime = (t_max - Time) / CFL;
        else if (Time + CFL * dTime * beta > t_max) dTime = (t_max - Time) / (CFL * beta);
        if (int e = cffc_set_global_timestep(c, dTime)) return e;
        // :796-803 Freeze_Limiters once the (lagged) L2 norm is low enough
        if (freeze_level > 0.0 && !first_step && !c->freeze) {
            double nr[3];
            if (int e = cffc_residual_norms(c, 1, nr)) return e;
            if (nr[1] <= freeze_level)
                if (int e = cffc_freeze_limiters(c, 1)) return e;
        }
        for (int st = 1; st <= c->s.n_stage && !rc; ++st) {
            int e;
            if ((e = cffc_exchange_and_bcs(c)) || (e = cffc_dudt_stage(c, st)) ||
                (e = cffc_apply_flux_corrections(c, st)) || (e = cffc_residual_smoothing(c, st)))
                return e;
            rc = cffc_update_stage(c, st, nullptr);
            if (rc < 0) return rc;
        }
        if (rc) break;
        first_step = false;
        ++n;
        Time = optsm ? Time + CFL * dTime * beta : Time + CFL * dTime;
    }
    *time = Time;
    if (steps_done) *steps_done = n;
    return rc;
}
