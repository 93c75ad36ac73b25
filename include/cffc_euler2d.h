/*
 * cffc_euler2d.h -- C ABI of the MI355X (gfx950) implementation of CFFC's
 * explicit finite-volume residual/update path for the 2D Euler equations on
 * quadrilateral multi-block meshes.
 *
 * The reference calls this path through C++ free functions that loop over the
 * blocks of a processor (src_2D/Euler2D/Euler2DQuadMultiBlock.cc) and through
 * the message-passing template Send_All_Messages
 * (src_2D/AMR/AdaptiveBlock2D_MessagePassing.h:7883).  Each entry point below
 * replaces one of those calls; its comment cites the reference interface.
 * All arguments are plain pointers and sizes: no C++ types cross this
 * boundary and no exception escapes it.
 *
 * Threading: one host thread per context.  All work is ordered on the
 * context's HIP stream; entry points that return host data synchronise it.
 *
 * Return codes: 0 = success; > 0 = a failure the reference reports through
 * its int return (e.g. the i-index of the first cell whose density, total
 * energy or internal energy became non-positive, as returned by
 * Update_Solution_Multistage_Explicit, Euler2DQuadSingleBlock.cc:5375-5385);
 * < 0 = a CFFC_E_* runtime error below.
 */
#ifndef CFFC_EULER2D_H
#define CFFC_EULER2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime error codes (negative returns) ---------------------------- */
#define CFFC_E_ARG        (-1)  /* invalid argument / unsupported option      */
#define CFFC_E_HIP        (-2)  /* a HIP runtime call failed                  */
#define CFFC_E_NOMEM      (-3)  /* device allocation failed                   */
#define CFFC_E_TOPOLOGY   (-4)  /* inconsistent neighbour table (the reference
                                   reports these as 2200-2205 from
                                   Load_Send_Message_Buffers_NoResChange)     */
#define CFFC_E_NODEVICE   (-5)  /* no HIP device visible                      */
#define CFFC_E_STEP_LIMIT (-6)  /* a step cap ended a run before its end time  */

/* ---- reference constants (src_2D/CFD/CFD.h) ---------------------------- */
/* boundary-condition codes, CFD.h:273-351 */
#define CFFC_BC_NONE                    0
#define CFFC_BC_CONSTANT_EXTRAPOLATION  4001
#define CFFC_BC_PERIODIC                6000
#define CFFC_BC_REFLECTION              7001
#define CFFC_BC_CHARACTERISTIC          8001
#define CFFC_BC_FIXED                   11000
/* time integration, CFD.h:651-659 */
#define CFFC_TIME_STEPPING_EXPLICIT_EULER                 0
#define CFFC_TIME_STEPPING_EXPLICIT_PREDICTOR_CORRECTOR   3
#define CFFC_TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING   7
#define CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA           8
/* limiters, CFD.h:749-761 */
#define CFFC_LIMITER_ONE             (-1)
#define CFFC_LIMITER_ZERO              0
#define CFFC_LIMITER_VENKATAKRISHNAN   9
/* flux functions, CFD.h:769-771 */
#define CFFC_FLUX_FUNCTION_ROE         2
#define CFFC_FLUX_FUNCTION_HLLE        4
/* multigrid cycles, CFD.h:689-690 */
#define CFFC_MULTIGRID_V_CYCLE         1
#define CFFC_MULTIGRID_W_CYCLE         2
/* reconstruction, CFD.h */
#define CFFC_RECONSTRUCTION_LEAST_SQUARES 2
/* time-step type, CFD.h:801-802 */
#define CFFC_GLOBAL_TIME_STEPPING        0
#define CFFC_SCALAR_LOCAL_TIME_STEPPING  1

/* arithmetic of the reconstruction / limiter / Roe flux (the state, the
   update, the time-step minimum and the residual norms are always FP64) */
#define CFFC_FLUX_FP64 0
#define CFFC_FLUX_FP32 1

/* neighbour slots in cffc_mesh_desc.nbr */
enum { CFFC_NBR_N = 0, CFFC_NBR_S, CFFC_NBR_E, CFFC_NBR_W,
       CFFC_NBR_NE, CFFC_NBR_NW, CFFC_NBR_SE, CFFC_NBR_SW, CFFC_NBR_COUNT };

/*
 * Multi-block mesh: the data the reference keeps in Grid2D_Quad_Block_HO
 * (nodes, BCtypeN/S/E/W; HO_Grid2DQuad.h:264-330), in the solution block's
 * boundary reference states WoN/S/E/W (Euler2DQuad.h:140-205) and in the
 * same-level neighbour fields nN/nS/.../infoX[0].blknum of
 * AdaptiveBlock2D_List (AdaptiveBlock.h:820).
 *
 * Block b has ni[b] x nj[b] interior cells and nghost (= 2) ghost layers, so
 * nci = ni+2*nghost, ncj = nj+2*nghost cells including ghosts and
 * (nci+1) x (ncj+1) nodes.  Arrays of all blocks are concatenated in block
 * order; within a block the i index runs fastest.
 *   xnode/ynode : (nci+1)*(ncj+1) node coordinates per block, ghost nodes
 *                 included (node (i,j) is the SW corner of cell (i,j)).
 *   nbr         : 8 ints per block (CFFC_NBR_*), the same-level neighbour
 *                 block index or -1.  Blocks must share orientation.
 *   bc          : per block BCtypeW[ncj], BCtypeE[ncj], BCtypeS[nci],
 *                 BCtypeN[nci] (ghost entries included).
 *   wo          : the reference state (rho,u,v,p) for each bc entry, same
 *                 order, 4 doubles per entry (used by FIXED/CHARACTERISTIC).
 *   level       : quadtree refinement level per block (info.level), or NULL
 *                 for a single-level mesh.
 *   nbr2        : 4 ints per block (faces N, S, E, W), or NULL: the second
 *                 neighbour of a face that has two finer neighbours (nX == 2,
 *                 infoX[1]); nbr then holds the first (infoX[0]: the west one
 *                 of an N/S face, the south one of an E/W face); else -1.
 * Across a resolution change (neighbours one level apart, on faces or
 * corners) the blocks have equal, even ni and nj; the fine side's ghosts are
 * prolonged from the coarse block, the coarse side's restricted from the fine
 * blocks, and the coarse face fluxes corrected to the fine ones
 * (AdaptiveBlock2D_MessagePassing.h:1185-2398, 2399-5679, 6284-7882, 8328).
 * The coarse-to-fine messages read the coarse block's BC ghosts as the
 * previous BCs call left them, as in the reference.
 *   lattice     : 2 ints per block, or NULL (no counterpart in the reference,
 *                 whose blocks all have the same cell counts): the block's
 *                 first interior cell (I0, J0) on the cell lattice of its
 *                 level, where cell (I, J) of level l+1 lies in cell
 *                 (I>>1, J>>1) of level l (level required).  With it,
 *                 blocks of any size and position are allowed, and every
 *                 ghost cell is filled from the block that holds its place:
 *                 copied from the same level, restricted from the finer
 *                 level, prolonged from the coarser level (the reference's
 *                 message rules restated by position), or set by the BCs
 *                 where no block holds it; nbr and nbr2 are then unused
 *                 (all -1).  Levels of touching blocks differ by at most 1.
 */
typedef struct cffc_mesh_desc {
    int32_t        nblk;
    int32_t        nghost;
    const int32_t *ni;
    const int32_t *nj;
    const double  *xnode;
    const double  *ynode;
    const int32_t *nbr;
    const int32_t *bc;
    const double  *wo;
    const int32_t *level;
    const int32_t *nbr2;
    const int32_t *lattice;
} cffc_mesh_desc;

/* The scheme fields of Euler2D_Input_Parameters read on this path
   (Euler2DInput.h:57-118). */
typedef struct cffc_scheme_desc {
    int32_t i_time_integration;   /* CFFC_TIME_STEPPING_*                   */
    int32_t n_stage;              /* N_Stage                                */
    double  cfl_number;           /* CFL_Number                             */
    int32_t i_reconstruction;     /* CFFC_RECONSTRUCTION_LEAST_SQUARES      */
    int32_t i_limiter;            /* CFFC_LIMITER_*                         */
    int32_t i_flux_function;      /* CFFC_FLUX_FUNCTION_ROE / _HLLE         */
    int32_t local_time_stepping;  /* CFFC_GLOBAL/SCALAR_LOCAL_TIME_STEPPING */
    int32_t flux_precision;       /* CFFC_FLUX_FP64 / CFFC_FLUX_FP32        */
    int32_t axisymmetric;         /* Axisymmetric: 0 planar, 1 y is the radius
                                     (source S(W, X), Euler2DState.h:3179; every
                                     cell centroid must have y > 0)          */
} cffc_scheme_desc;

/* The FAS multigrid parameters of Multigrid_Input_Parameters
   (FASMultigrid2DInput.h:93-114; the reference's defaults in brackets).
   Coarse-mesh boundary conditions are always applied
   (Apply_Coarse_Mesh_Boundary_Conditions = ON, the default). */
typedef struct cffc_mg_desc {
    int32_t levels;                      /* Levels, >= 2                     [2] */
    int32_t i_cycle;                     /* CFFC_MULTIGRID_V/W_CYCLE         [V] */
    int32_t smooths_on_finest_level;     /* Number_of_Smooths_on_Finest_Level[1] */
    int32_t pre_smooths;                 /* Number_of_Pre_Smooths            [1] */
    int32_t post_smooths;                /* Number_of_Post_Smooths           [0] */
    int32_t smooths_on_coarsest_level;   /* Number_of_Smooths_on_Coarsest_Level [3] */
    int32_t defect_correction;           /* Defect_Correction                [1] */
    int32_t prolong_using_injection;     /* Prolong_Using_Injection          [0] */
    int32_t first_order_coarse;          /* First_Order_Coarse_Mesh_Reconstruction: HLLE
                                            flux, zero limiter below the top [1] */
    int32_t injection_at_dirichlet;      /* Injection_at_Dirichlet_Boundary_Conditions [1] */
    int32_t update_stability_switch;     /* Update_Stability_Switch          [1] */
    int32_t max_update_reductions;       /* Maximum_Number_of_Update_Reductions [5] */
    double  absolute_tolerance;          /* Absolute_Convergence_Tolerance  [-1] */
    double  relative_tolerance;          /* Relative_Convergence_Tolerance  [-1] */
    int32_t fmg_cycles;                  /* Number_of_Cycles_per_Stage_for_Full_Multigrid [0] */
    int32_t pad;
    double  fmg_absolute_tolerance;      /* FMG_Absolute_Convergence_Tolerance [-1] */
    double  fmg_relative_tolerance;      /* FMG_Relative_Convergence_Tolerance [-1] */
} cffc_mg_desc;

typedef struct cffc_ctx cffc_ctx;

/* Allocate the device state for a mesh and scheme (the reference's
   Allocate + Create_Initial_Solution_Blocks, Euler2DQuadMultiBlock.cc:22-33,
   AMR.h:39).  Geometry is derived from the nodes on the device side. */
int  cffc_create(const cffc_mesh_desc *mesh, const cffc_scheme_desc *scheme,
                 cffc_ctx **out);
void cffc_destroy(cffc_ctx *ctx);

/* Host-only check of a mesh descriptor: the metadata, topology, BC and
   resolution-change checks of cffc_create, without a device.  On success
   counts[0..3] receive the numbers of restriction, prolongation and
   flux-correction operations and 1 if the mesh has resolution changes. */
int  cffc_check_mesh(const cffc_mesh_desc *mesh, int64_t counts[4]);

/* Diagnostic, no reference counterpart: the device FP64 arithmetic of the
   stage kernels beside the compiler's, for n operand pairs (a[k], b[k]).
   out[8k..8k+7] = the kernels' correctly rounded quotient a/b (shared-
   reciprocal form), the compiler's a / b, the kernels' sqrt(|a|), the
   compiler's sqrt(|a|), the kernels' single-instruction min(a, b), fmin,
   the kernels' max(a, b), fmax.  Runs on the current device. */
int  cffc_check_arith(const double *a, const double *b, int64_t n, double *out);

/* Diagnostic, no reference counterpart: the operation lists cffc_create
   builds for a mesh, host only.  kind 0: restrictions, 1: same-level copies
   (lattice meshes), 2: prolongations, 3: flux corrections, as the library's
   packed records (24, 112 and 32 bytes; layout in
   cffc_amd/csrc/euler2d_device.hpp).  *n receives the count; at most cap
   records are copied to out (may be NULL). */
int  cffc_mesh_ops(const cffc_mesh_desc *mesh, int kind, void *out, int64_t cap, int64_t *n);

/* Number of interior cells over all blocks (the unit of the state arrays). */
int64_t cffc_num_cells(const cffc_ctx *ctx);

/* Host <-> device copies of the interior state, AoS, 4 doubles per cell,
   blocks in order, i fastest.  set_W sets W then U = U(W) (the reference ICs,
   Euler2DQuadSingleBlock.cc:1689); set_U sets U then W = W(U) (the restart
   path, operator>> Euler2DQuad.h:1340-1399). */
int  cffc_set_W(cffc_ctx *ctx, const double *W, int64_t ncell);
int  cffc_set_U(cffc_ctx *ctx, const double *U, int64_t ncell);
int  cffc_get_W(cffc_ctx *ctx, double *W, int64_t ncell);
int  cffc_get_U(cffc_ctx *ctx, double *U, int64_t ncell);
/* Residual storage dUdt[k] of the interior cells (AoS, 4 doubles/cell).
   k follows the reference's k_residual (Euler2DQuad.cc:1249-1279). */
int  cffc_get_dUdt(cffc_ctx *ctx, int k, double *dUdt, int64_t ncell);
/* Per-cell time step dt of the interior cells. */
int  cffc_get_dt(cffc_ctx *ctx, double *dt, int64_t ncell);

/* == Send_All_Messages(blocks, list, NUM_VAR_EULER2D, OFF)
      (AdaptiveBlock2D_MessagePassing.h:7883-8020) followed by
      BCs(blocks, list, IP) (Euler2DQuadMultiBlock.cc:937): fills both ghost
      layers of every block from its same-level neighbours (faces and
      corners), then applies the physical boundary conditions. */
int  cffc_exchange_and_bcs(cffc_ctx *ctx);

/* == Send_Conservative_Flux_Corrections(blocks, list, NUM_VAR_EULER2D)
      (AdaptiveBlock2D_MessagePassing.h:8328) followed by
      Apply_Boundary_Flux_Corrections_Multistage_Explicit(blocks, list, IP,
      i_stage) (Euler2DQuadMultiBlock.cc:1307 -> SingleBlock.cc:4992-5080):
      on every face with two finer neighbours, the coarse face flux saved by
      cffc_dudt_stage becomes -(sum of the two fine fluxes) - own, and
      dUdt[k] -= (CFL*dt)*Flux/A on the face's cells.  No-op without
      resolution changes. */
int  cffc_apply_flux_corrections(cffc_ctx *ctx, int i_stage);

/* == CFL(blocks, list, IP) (Euler2DQuadMultiBlock.cc:963) followed by
      CFFC_Minimum_MPI (MPI.h:184) and, when set_global != 0,
      Set_Global_TimeStep (Euler2DQuadMultiBlock.cc:994). */
int  cffc_cfl(cffc_ctx *ctx, double *dt_min, int set_global);
/* Set_Global_TimeStep (Euler2DQuadMultiBlock.cc:994): every cell's dt and
   the global step become dt (the time-accurate loop's clipped step) */
int  cffc_set_global_timestep(cffc_ctx *ctx, double dt);
/* The time-accurate explicit loop (Euler2DQuadSolvers.cc:526-898 with
   Time_Accurate = 1, global time stepping), from *time (0, or the Time of a
   restart) to t_max: per step CFL + global minimum, the step clipped by the
   reference's if / else-if (:662-680; the second branch, with
   MultiStage_Optimally_Smoothing(N, N, limiter), applies to every scheme),
   Set_Global_TimeStep, Freeze_Limiters when freeze_level > 0 and the lagged
   L2 norm of the density residual is <= freeze_level (not on the first
   step, :796-803), then per stage exchange + BCs, dUdt, flux corrections,
   residual smoothing (when set) and the update; Time += CFL*dTime (times
   MultiStage_Optimally_Smoothing(N, N, limiter) for that scheme, :884-896).
   Stops when Time >= t_max, like the reference, which has no step cap;
   max_steps > 0 adds one and returns CFFC_E_STEP_LIMIT when it ends the run
   first.  *time (required) receives the time reached, *steps_done (may be
   NULL) the steps taken; a positive return is the update's failing-cell
   code. */
int  cffc_run_time_accurate(cffc_ctx *ctx, double t_max, int max_steps, double freeze_level, double *time,
                            int *steps_done);

/* == dUdt_Multistage_Explicit(blocks, global, list, IP, i_stage)
      (Euler2DQuadMultiBlock.cc:1346 -> Euler2DQuad.cc:1235). */
int  cffc_dudt_stage(cffc_ctx *ctx, int i_stage);

/* == Update_Solution_Multistage_Explicit(blocks, list, IP, i_stage)
      (Euler2DQuadMultiBlock.cc:1383 -> Euler2DQuadSingleBlock.cc:5300).
      Returns the reference's code (> 0: i-index of the first failing cell);
      bad_cell, if not NULL, receives block*2^32 + j*2^16 + i of that cell. */
int  cffc_update_stage(cffc_ctx *ctx, int i_stage, int64_t *bad_cell);

/* == L1_Norm_Residual, L2_Norm_Residual, Max_Norm_Residual
      (Euler2DQuadMultiBlock.cc:1017-1104) of dUdt[..][0][var], var in 1..4.
      out = {L1, L2, max}. */
int  cffc_residual_norms(cffc_ctx *ctx, int var, double out[3]);

/* == Freeze_Limiters / Evaluate_Limiters (Euler2DQuadMultiBlock.cc:1190-1218).
      Freezing reuses the limiter values of the last evaluated stage. */
int  cffc_freeze_limiters(cffc_ctx *ctx, int freeze);
/* Keep the last stage's limiter values stored during cffc_run_steps so that
   a later cffc_freeze_limiters (the reference freezes once the residual falls
   below Freeze_Limiter_Residual_Level, Euler2DQuadSolvers.cc:796-803) has
   them; costs one extra 32-byte write per cell on the last stage. */
int  cffc_set_store_limiters(cffc_ctx *ctx, int on);

/* The explicit time-marching loop of Euler2DQuadSolver
   (Euler2DQuadSolvers.cc:549-898) for a fixed number of steps, with the
   stages fused (exchange/BC kernel + one fused residual/update kernel per
   stage; the next step's time step and the residual norms are produced by
   the last stage).  norms (may be NULL) receives {L1,L2,max} of the density
   residual per step (nsteps*3 doubles, the reference's lagged values: entry
   n is the norm printed at the top of step n+1).  dt_out (may be NULL)
   receives the global dt of each step.  Stops early (returning the number of
   steps taken in *steps_done) when the L2 norm drops to tol (tol <= 0: never). */
int  cffc_run_steps(cffc_ctx *ctx, int nsteps, double tol, double *norms,
                    double *dt_out, int *steps_done);

/* Same as cffc_run_steps but without host synchronisation per step: the
   steps are captured once into a hipGraph and replayed.  norms/dt_out are
   copied back at the end.  tol is ignored. */
int  cffc_run_steps_graph(cffc_ctx *ctx, int nsteps, double *norms,
                          double *dt_out);

/* No counterpart in the reference: cffc_run_steps for callers that cannot
   allocate (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same steps, results and arguments as
   cffc_run_steps, but it calls no hipMalloc / hipFree / hipHostMalloc: the
   per-step records live in a fixed device array of the library (one per
   device, kept until the process ends), and the host does not poll (steps
   after convergence or a failure are skipped by the kernels themselves).
   At most 4096 steps per call (CFFC_E_ARG beyond).  Refused (CFFC_E_ARG)
   while kernel timing is on, or when the context still lacks a buffer the
   run needs (limiters, smoothing), which the set-up calls allocate.  Calls
   that overlap on one device share that array and race. */
int  cffc_run_steps_noalloc(cffc_ctx *ctx, int nsteps, double tol, double *norms,
                            double *dt_out, int *steps_done);

/* Implicit residual smoothing (Residual_Smoothing, Euler2DQuadSingleBlock.cc:
   4521-4564, applied between dUdt_Multistage_Explicit and the update as in
   Euler2DQuadSolvers.cc:861-867).  Inputs Residual_Smoothing_Epsilon and
   Residual_Smoothing_Gauss_Seidel_Iterations; iterations = 0 switches it
   off (the default).  Gauss-Seidel sweeps in the reference's order, bit for
   bit.  Refused (CFFC_E_ARG) for 4-stage Runge-Kutta, whose stage 3 would
   need a fourth residual vector. */
int  cffc_set_residual_smoothing(cffc_ctx *ctx, double epsilon, int iterations);
/* == Residual_Smoothing(blocks, list, IP, i_stage) (MultiBlock.cc:1227),
   the granular form; cffc_run_steps / cffc_run_steady apply it by themselves. */
int  cffc_residual_smoothing(cffc_ctx *ctx, int i_stage);

/* Batched steady runs.  The reference runs one case per solver invocation
   (Euler2DQuadSolver, Euler2DQuadSolvers.cc:26, explicit loop :549-898, with
   Freeze_Limiters at Freeze_Limiter_Residual_Level, :796-803).  Here several
   independent cases ("parts") may share one context: part q owns blocks
   first_block[q] .. first_block[q+1]-1 (npart+1 entries, first 0, last
   nblk), and no block may have a neighbour in another part
   (CFFC_E_TOPOLOGY otherwise). */
int  cffc_set_parts(cffc_ctx *ctx, int npart, const int32_t *first_block);
int  cffc_num_parts(const cffc_ctx *ctx);

/* Stall freeze for batched steady runs: no counterpart in the reference,
   whose Freeze_Limiters waits for Freeze_Limiter_Residual_Level alone.  With
   window > 0, cffc_run_steady also freezes a running part's limiters once
   its L2 norm has not gone 1% below its best for window consecutive steps
   (the limiter cycling that keeps a case above the freeze level); window = 0
   (the default) keeps the reference's rule only.  Counted from the start of
   the run, across resumed segments. */
int  cffc_set_stall_freeze(cffc_ctx *ctx, int window);

/* Run every part to its own convergence in one stream of launches: each
   step is the fused step of cffc_run_steps; after each step every part's
   density-residual L2 norm (the reference's, over the part's cells) is
   checked: at <= freeze_level (> 0) the part's limiters freeze (its
   Freeze_Limiters), at <= tol the part stops updating and part_steps[q]
   receives the number of steps it took (-1: not converged within
   max_steps).  part_l2 (may be NULL) receives each part's last norm.
   Returns when every part has converged or after max_steps steps
   (*steps_done).  A part run alone and in a batch takes the same steps and
   ends in the same state.  With CFFC_STEADY_HISTORY the per-step norms of every
   part are kept for cffc_get_part_history ([nsteps][npart]).  Without
   cffc_set_parts the whole mesh is one part.  A non-positive density in any
   part stops the batch with the reference's error code.  Parts are
   independent only with local time stepping (a global dt couples them).
   flags: CFFC_STEADY_HISTORY keeps the per-step norms as above;
   CFFC_STEADY_RESUME continues the previous run (part states, frozen
   limiters, step counts) for another max_steps steps, so that a long solve
   can be driven in segments with the same result as one call (part_steps
   counts all segments; *steps_done this segment's steps). */
#define CFFC_STEADY_HISTORY 1
#define CFFC_STEADY_RESUME  2
int  cffc_run_steady(cffc_ctx *ctx, int max_steps, double tol, double freeze_level, int flags,
                     int32_t *part_steps, double *part_l2, int *steps_done);
int  cffc_get_part_history(cffc_ctx *ctx, int nsteps, double *out);

/* Device wall-clock time from the start of the last cffc_run_steps /
   cffc_run_steps_graph / cffc_run_steady to the end of each of its first
   nsteps steps (seconds; for per-case latencies of a batched run). */
int  cffc_step_times(cffc_ctx *ctx, int nsteps, double *seconds);

/* Wait for all work on the context's stream. */
int  cffc_synchronize(cffc_ctx *ctx);

/* HIP stream of the context (a hipStream_t), for callers that time kernels
   with their own events on the stream the kernels run on. */
void *cffc_stream(cffc_ctx *ctx);

/* Tile configuration of the fused stage kernel, chosen by the autotuner or
   set here: tile_i x tile_j interior cells per workgroup. */
int  cffc_set_tile(cffc_ctx *ctx, int tile_i, int tile_j);
int  cffc_get_tile(const cffc_ctx *ctx, int *tile_i, int *tile_j);

/* Tile autotuner (no reference counterpart: the reference's blocks are
   swept cell by cell).  Times nsteps fused time steps with every
   instantiated tile shape on this context's mesh and current state,
   restores the state after each trial, and keeps the fastest shape
   (returned in tile_i/tile_j).  ms_per_step (may be NULL, room for 16)
   receives the time per step of every candidate; *ncand their number. */
int  cffc_autotune_tile(cffc_ctx *ctx, int nsteps, int *tile_i, int *tile_j, double *ms_per_step, int *ncand);

/* Kernel timing for measurement: with timing on, cffc_run_steps brackets
   every exchange/BC kernel and every fused stage kernel with a pair of HIP
   events on the context stream; cffc_kernel_times returns {total stage-kernel
   ms, stage launches, total exchange-kernel ms, exchange launches} accumulated
   since timing was switched on.  (Not applied inside cffc_run_steps_graph.) */
int  cffc_set_kernel_timing(cffc_ctx *ctx, int on);
int  cffc_kernel_times(cffc_ctx *ctx, double out[4]);

/* ---- FAS multigrid: FAS_Multigrid2D_Solver (FASMultigrid2D/FASMultigrid2D.h) ----
   levels[0..nlevels-1]: the finest mesh and its coarsenings, level l+1
   halving every block of level l (Half_Mesh_Resolution, built by the host
   as the reference's allocate does, FASMultigrid2D.h:418-658; single-level
   meshes, FP64).  smoother: the scheme of Multigrid_IP.i_Smoothing (the
   reference overrides i_Time_Integration with it while smoothing); coarse
   levels switch to HLLE and the zero limiter when first_order_coarse is
   set.  Each level is a cffc_ctx on one stream; cffc_mg_level returns it
   (cffc_set_W / cffc_get_W on level 0). */
typedef struct cffc_mg cffc_mg;
/* allocate, FASMultigrid2D.h:418 */
int  cffc_mg_create(const cffc_mesh_desc *levels, int nlevels, const cffc_scheme_desc *smoother,
                    const cffc_mg_desc *mg, cffc_mg **out);
void cffc_mg_destroy(cffc_mg *mg);
cffc_ctx *cffc_mg_level(cffc_mg *mg, int level);
/* Residual_Smoothing inside Smooth (:2987-2993) */
int  cffc_mg_set_residual_smoothing(cffc_mg *mg, double epsilon, int iterations);
/* after cffc_set_W on level 0: the coarse levels restricted from it with
   their boundary reference states (Apply_ICs with IC_RESTART :719-733,
   Restrict_Boundary_Ref_States :1112), messages and BCs on every level */
int  cffc_mg_initialize(cffc_mg *mg);
/* one cycle, Coarse_Grid_Correction(0, 0) :3118; nonzero = the first
   unphysical update's i index, *bad_cell its level << 56 | block << 32 |
   j << 16 | i (may be NULL) */
int  cffc_mg_cycle(cffc_mg *mg, int64_t *bad_cell);
/* Execute :3375-3745: with fmg_cycles > 0 the full-multigrid start-up
   (top level levels-2 up to 1, at most fmg_cycles cycles each or to the
   fmg tolerances, then the solution prolonged to the next level,
   Prolong_Solution_Blocks :1200), then the regular cycles from level 0.
   norms: L1, L2, max of the top level's density residual before each
   cycle, 3 per check, room for max_cycles + levels checks; limiter
   freezing at freeze_level (<= 0: never; unfrozen at each new top level);
   exits on max_cycles cycles in all or the tolerances of cffc_mg_desc */
int  cffc_mg_execute(cffc_mg *mg, int max_cycles, double freeze_level, double *norms, int *cycles_done,
                     int *checks);
/* Per-case stopping for a batch of independent cases in one mesh (no
   counterpart in the reference, whose Execute stops on one norm over all
   blocks): part q owns level-0 blocks first_block[q] .. first_block[q+1]-1
   (npart+1 entries, first 0, last the level's block count; no block may
   neighbour a block of another part, CFFC_E_TOPOLOGY otherwise).  With
   parts set, cffc_mg_execute's regular cycles (top level 0) stop once every
   part's L2 norm of the density residual over its own cells has been below
   tol[q] at a check after the second cycle (the rule of the single norm),
   and with freeze_level > 0 the limiters freeze once every part's norm is
   at or below freeze[q] (freeze may be NULL: 0).  From the check at which
   a part has met its tolerance the cycles leave its blocks alone on every
   level, so it keeps the state it converged in: every part ends in the
   state it reaches alone.  cffc_mg_initialize starts the parts afresh
   (none converged, frozen or skipped; the tolerances are kept).  npart = 0
   clears the parts. */
int  cffc_mg_set_parts(cffc_mg *mg, int npart, const int32_t *first_block, const double *tol, const double *freeze);
/* per part: the last checked L2 norm (l2, may be NULL) and the regular
   cycles done since cffc_mg_set_parts when its norm first met its
   tolerance (done_cycles, -1: not yet; may be NULL); returns npart */
int  cffc_mg_part_status(cffc_mg *mg, double *l2, int32_t *done_cycles);
/* cffc_mg_execute replays one captured cycle as a HIP graph (on by
   default; recaptured once when the limiters freeze); off: the launches
   are issued one by one, with the same arithmetic.  No reference
   counterpart: launch overhead. */
int  cffc_mg_set_graph(cffc_mg *mg, int on);
/* the forcing term P of a coarse level, interior cells in block order */
int  cffc_mg_get_P(cffc_mg *mg, int level, double *P, int64_t ncell);

/* Last error message of this thread (static storage). */
const char *cffc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CFFC_EULER2D_H */
