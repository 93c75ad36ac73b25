// euler2d_kernels.hpp -- launch interface between the host runtime (capi.cpp)
// and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "euler2d_device.hpp"
#include "euler2d_stage.hpp"

namespace cffc {

struct ExchangeParams {
    const BlockMeta *meta;
    const int32_t *nbr;        // 8 per block
    const int32_t *bc;
    const double *wo;
    const double *geom;
    int64_t ncp;
    double *W;                 // W buffer whose ghosts are filled (FP64 storage)
    float *Wf;                 // or, if not null, the FP32 W buffer of the fused mixed path
    double *U;                 // conserved state (stage 1: ghosts = W(U))
    int32_t from_U;
    int32_t store_U;           // also fill the ghost U: copied with the neighbour's U, U(W) of a BC state
                               // (the multigrid levels read ghost U; needs from_U, FP64 storage)
    int32_t *stop_flag;
    unsigned long long *reset_slot;  // set to MILLION (next-step dt slot), may be null
    const int32_t *blocks;     // blocks to process (workgroup k -> blocks[k]), or null: workgroup = block
    int32_t bcs_only;          // skip the same-level exchange (ghosts already pushed by the stage kernel)
    int32_t no_bcs;            // same-level exchange only (the resolution-change legs and BCs follow)
    double *W2;                // if not null, the BC ghost states are also written here (FP64 storage) ...
    float *Wf2;                // ... or here (FP32 storage): the next stage's W buffer, see exchange_bc_kernel
    int32_t nthreads;          // workgroup size (64, 128 or 256; 0: 256)
};



// end of a fused time step: reduce the last stage's per-tile partials,
// record the step's norms and dt, set the next step's global dt, and raise
// the stop flag on convergence.  With parts (batched steady runs) every part
// is reduced on its own and moves RUNNING -> FROZEN -> DONE.
struct FinishParams {
    const double *norm_part;                // [4][ntiles]
    int32_t ntiles;
    const unsigned long long *dtmin_in;     // dt of the step that just ended (bits)
    unsigned long long *dtmin_out;          // next step's dt (bits)
    double *norm_out;                       // [max_steps][3]
    double *dt_out;                         // [max_steps]
    int32_t *step_ctr;
    int32_t *stop_flag;
    double tol;                             // L2 threshold (<= 0: none)
    // parts (npart == 0: none)
    int32_t npart;
    const int32_t *part_tile0;              // [npart + 1] first tile of each part
    int32_t *part_state;                    // [npart] PartState
    int32_t *part_steps;                    // [npart] steps to convergence
    double *part_l2;                        // [max_steps][npart] or null
    double *part_last;                      // [npart] L2 of the part's last step
    double freeze_level;                    // part L2 at which limiters freeze (<= 0: never)
    int32_t step_offset;                    // steps of earlier segments of the same steady run
    unsigned long long *step_clock;         // [max_steps] device wall clock at the end of each step, or null
    const double *chunk_part;               // [4][nchunk] chunk partials of norm_part (nchunk > 0, no parts)
    int32_t nchunk;
    // stall freeze: a running part whose L2 has not gone 1% below its best
    // for stall_window steps freezes its limiters (0: never)
    int32_t stall_window;
    double *part_best;                      // [npart] smallest L2 so far
    int32_t *part_best_step;                // [npart] step of that L2
};

}  // namespace cffc

// tile shapes the stage kernel is instantiated for (TI x TJ interior cells)
#define CFFC_TILE_LIST(X) X(16, 16) X(32, 8) X(8, 32) X(16, 8) X(8, 8) X(32, 4) X(32, 16) X(16, 4) X(8, 16)
// FP64 only, not autotuned: the default tile of meshes whose blocks are
// narrower than 8 cells (the coarse multigrid levels), one wavefront each
#define CFFC_SMALL_TILE_LIST(X) X(4, 4)

extern "C" {
hipError_t cffc_launch_widen(const float *src, double *dst, int64_t n, hipStream_t st);
hipError_t cffc_launch_arith_check(const double *a, const double *b, int64_t n, double *out, hipStream_t st);
hipError_t cffc_launch_part_norms(const double *part, const int32_t *tile0, int npart, double *out,
                                  hipStream_t st);
hipError_t cffc_launch_geometry(const cffc::BlockMeta *meta, int nblk, int max_cells, const double *xn,
                                const double *yn, double *geom, int64_t ncp, unsigned int *ndeg, hipStream_t st);
hipError_t cffc_launch_geometry_f32(const cffc::BlockMeta *meta, int nblk, int max_cells, const double *geom,
                                    float *gf, int64_t ncp, hipStream_t st);
hipError_t cffc_launch_scatter(const cffc::BlockMeta *meta, int nblk, int max_int, const double *aos, double *W,
                               int64_t ncp, int kind, double *U, hipStream_t st);
hipError_t cffc_launch_gather(const cffc::BlockMeta *meta, int nblk, int max_int, const double *src, double *aos,
                              int64_t ncp, int nvar, hipStream_t st);
hipError_t cffc_launch_exchange(const cffc::ExchangeParams *p, int nblk, hipStream_t st);
hipError_t cffc_launch_cfl(const cffc::BlockMeta *meta, const cffc::TileDesc *tiles, int ntiles, int ti, int tj,
                           const double *W, const double *geom, int64_t ncp, double *dt, unsigned long long *dtmin,
                           hipStream_t st);
hipError_t cffc_launch_set_global_dt(const cffc::BlockMeta *meta, int nblk, int max_int, double *dt,
                                     const unsigned long long *dtmin, hipStream_t st);
hipError_t cffc_launch_update(const cffc::BlockMeta *meta, int nblk, int max_int, double *W, double *U,
                              const double *Uo, double *dU, double *dt, int64_t ncp, double omega, int lts,
                              unsigned long long *err, hipStream_t st);
hipError_t cffc_launch_copy_interior(const cffc::BlockMeta *meta, int nblk, int max_int, const double *src,
                                     double *dst, int64_t ncp, hipStream_t st);
hipError_t cffc_launch_norms(const cffc::BlockMeta *meta, const cffc::TileDesc *tiles, int ntiles, int ti, int tj,
                             const double *dU, int64_t ncp, int var, double *part, double *out, hipStream_t st);
hipError_t cffc_launch_set_u64(unsigned long long *dst, unsigned long long v0, unsigned long long v1, int n,
                               hipStream_t st);
hipError_t cffc_launch_finish(const cffc::FinishParams *p, double *chunk_part, int nchunk, hipStream_t st);
hipError_t cffc_launch_stamp(unsigned long long *dst, hipStream_t st);
hipError_t cffc_launch_residual_smoothing(const cffc::BlockMeta *meta, int nblk, double *S, double *T, int64_t ncp,
                                          double eps, int iters, const int32_t *part_state, const int32_t *stop_flag,
                                          hipStream_t st);
hipError_t cffc_launch_update_tile_fp64(const cffc::StageParams *p, int ti, int tj, hipStream_t st);
hipError_t cffc_launch_update_tile_mixed(const cffc::StageParams *p, int ti, int tj, hipStream_t st);
hipError_t cffc_launch_w_from_u(const cffc::BlockMeta *meta, int nblk, int max_int, const double *U, double *W,
                                float *Wf, int64_t ncp, hipStream_t st);
hipError_t cffc_launch_restrict(const cffc::RestrictOp *ops, int nops, const double *U, const double *geom,
                                int64_t ncp, double *W, float *Wf, const int32_t *stop_flag, int copy_from_U,
                                hipStream_t st);
hipError_t cffc_launch_prolong(const cffc::ProlongOp *ops, int nops, const cffc::BlockMeta *meta, const double *geom,
                               const double *phi, int frozen, const int32_t *part_state, int64_t ncp, double *W,
                               float *Wf, const int32_t *stop_flag, hipStream_t st);
hipError_t cffc_launch_flux_fix(const cffc::FluxFixOp *ops, int nops, double *S, const double *eflux,
                                const double *geom, const double *dt, const unsigned long long *dtmin, int lts,
                                double cfl, int64_t ncp, const int32_t *stop_flag, hipStream_t st);
hipError_t cffc_launch_stage_fp64(const cffc::StageParams *p, int ti, int tj, int fused, hipStream_t st);
// multigrid level transfers (euler2d_mg.hip)
hipError_t cffc_launch_mg_restrict(const cffc::BlockMeta *mf, const cffc::BlockMeta *mc, int nblk, int max_int_c,
                                   const double *src, const double *Af, int64_t ncpf, double *dst, const double *Ac,
                                   int64_t ncpc, hipStream_t st);
hipError_t cffc_launch_mg_primitive(const cffc::BlockMeta *meta, int nblk, int max_int, const double *U, double *W,
                                    int64_t ncp, hipStream_t st);
hipError_t cffc_launch_mg_subtract(const cffc::BlockMeta *meta, int nblk, int max_int, double *P, const double *R,
                                   int64_t ncp, hipStream_t st);
hipError_t cffc_launch_mg_cfl(const cffc::BlockMeta *mf, const cffc::BlockMeta *mc, int nblk, int max_int_c,
                              const double *dtf, const double *Af, double *dtc, const double *Ac, hipStream_t st);
hipError_t cffc_launch_mg_changes(const cffc::BlockMeta *meta, int nblk, int max_cells, double *U, const double *uo,
                                  int64_t ncp, hipStream_t st);
hipError_t cffc_launch_mg_copy(const cffc::BlockMeta *meta, int nblk, int max_cells, double *dst, const double *src,
                               int64_t ncp, hipStream_t st);
hipError_t cffc_launch_mg_prolong(const cffc::BlockMeta *mc, const cffc::BlockMeta *mf, int nblk, int max_int_c,
                                  int max_edge_c, const int32_t *bc, double *Uc, const double *gc, int64_t ncpc,
                                  double *Uf, const double *gf, int64_t ncpf, int neumann, int inj_dirichlet,
                                  int inj_always, int stability, int max_red, int assign, hipStream_t st);
hipError_t cffc_launch_stage_mixed(const cffc::StageParams *p, int ti, int tj, int fused, hipStream_t st);
}
