// euler2d_device.hpp -- device data layout shared by the gfx950 kernels and
// the host runtime.
//
// HBM layout (one context):
//   * every block is stored padded with its two ghost layers,
//     nci = ni+4 by ncj = nj+4 cells, i fastest, rows pitch cells apart
//     (pitch = nci rounded up to 16) and the block offset chosen so that the
//     first interior cell of every row starts a 128-byte line; blocks are
//     concatenated, so a padded cell has one global index
//     c = coff + j*pitch + i;
//   * every field is SoA: variable v of field F lives at F + v*ncp, where ncp
//     is the total number of padded cells (coalesced 8-byte lanes along i);
//   * W (primitive, FP64) is double-buffered: a stage reads W^(s) with filled
//     ghosts and writes W^(s+1) for the interior; U holds the conserved state
//     U^n; dUdt[0] is the residual accumulator; geometry is precomputed once
//     per context from the nodes (GEO_* below).
#pragma once

#include <stdint.h>

namespace cffc {

enum {
    GEO_XC = 0, GEO_YC, GEO_A,                  // Cell[i][j].Xc, .A
    GEO_NEX, GEO_NEY, GEO_LE, GEO_XFE, GEO_YFE, // nfaceE, lfaceE, xfaceE
    GEO_NNX, GEO_NNY, GEO_LN, GEO_XFN, GEO_YFN, // nfaceN, lfaceN, xfaceN
    GEO_COUNT
};

// FP32 geometry planes of the mixed path: coordinates relative to the
// centroid of the block's first interior cell, normals, face lengths
enum {
    GF_XC = 0, GF_YC, GF_XFE, GF_YFE, GF_XFN, GF_YFN,
    GF_NEX, GF_NEY, GF_NNX, GF_NNY, GF_LE, GF_LN,
    GF_COUNT
};

struct BlockMeta {
    int32_t ni, nj, nci, ncj;
    int32_t pitch;   // row pitch of the padded block (cells, a multiple of 16)
    int32_t spare;
    int64_t coff;    // padded cell (0, 0) of the block; interior cell (2, j) is 128-byte aligned
    int64_t ioff;    // first interior cell of the block in the host state arrays
    int64_t noff;    // first node of the block in the node arrays
    int32_t bcoff;   // first entry of the block in the bc / wo arrays
    int32_t has_bc;  // nonzero if any BCtypeN/S/E/W entry is not BC_NONE
    int32_t amr;     // AMR_* flags (resolution changes)
    int32_t skip;    // nonzero: a converged case of a multigrid batch, left alone by the cycle's kernels
    int64_t eoff;    // first entry of the block's saved face fluxes (N [ni], S [ni], E [nj], W [nj])
};

// BlockMeta::amr: faces whose boundary fluxes the stage saves (the coarse side
// of a resolution change for its correction, the fine side for the sums it
// sends), and blocks some of whose cells feed a coarser neighbour's ghosts
enum { AMR_FLUX_N = 1, AMR_FLUX_S = 2, AMR_FLUX_E = 4, AMR_FLUX_W = 8, AMR_FLUX_ANY = 15, AMR_USTORE = 16 };

struct TileDesc {
    int32_t b, i0, j0, part;  // block, first interior cell of the tile, part (case) of the block
};

// Resolution-change operations (euler2d_amr.hip), built by the host per mesh.
// kind 0, restriction: coarse ghost dst <- fine cells src, src+1, src+pitch,
// src+pitch+1; kind 1 (lattice meshes): same-level ghost dst <- cell src
struct RestrictOp {
    int64_t dst, src;
    int32_t pitch, kind;
};
// prolongation: coarse cell src of block blk -> its quarters SW, SE, NW, NE,
// centred at (xq, yq), into fine ghost cells dst
struct ProlongOp {
    int64_t src;
    int64_t dst[4];
    double xq[4], yq[4];
    int32_t blk, pad;
};
// flux correction of coarse cell `cell` on n (1 or 2) faces, in the order N, S,
// E, W: own = the cell's saved face flux entry, fine = the first of the two
// fine entries it covers (entries of 4 doubles in the edge-flux buffer)
struct FluxFixOp {
    int64_t cell;
    int32_t n, blk;
    int32_t own[2], fine[2];
};

// per-part state of a batched steady run (one part = one independent case)
enum PartState { PART_RUNNING = 0, PART_FROZEN = 1, PART_DONE = 2 };

// stage modes of the residual accumulator (dUdt[0] in the fused path)
enum StageMode {
    MODE_ZERO = 0,       // dU starts at zero, store S = dU
    MODE_ZERO_ACC2 = 1,  // dU starts at zero, store S = S + 2*dU (RK4 stages 2, 3)
    MODE_FROM_S = 2,     // dU starts at S, store S = dU (RK4 stage 4, RK2 stage 2)
};

enum PhiMode { PHI_COMPUTE = 0, PHI_COMPUTE_STORE = 1, PHI_LOAD = 2 };

constexpr double MILLION = 1000000.00;

}  // namespace cffc
