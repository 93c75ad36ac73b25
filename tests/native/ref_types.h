// ref_types.h -- the few members of CFFC's host types that the GPU binding
// (euler2d_hip_shim.cpp) reads, under the reference's names, so that the
// binding compiles and runs here as it would inside CFFC.  These are this
// project's minimal declarations, not the reference's classes:
//   Euler2D_pState / Euler2D_cState   src_2D/Euler2D/Euler2DState.h:131-214, 655-742
//   Grid2D_Quad_Block_HO (Node, BCtype*)   src_2D/Grid/HO_Grid2DQuad.h:264-330
//   Euler2D_Quad_Block (W, U, Wo*, Grid, index bounds)   src_2D/Euler2D/Euler2DQuad.h:140-205
//   AdaptiveBlock2D_List / AdaptiveBlock2D (used, info, nN.., infoN..)   src_2D/AMR/AdaptiveBlock.h:580-830
//   Euler2D_Input_Parameters (scheme fields)   src_2D/Euler2D/Euler2DInput.h:57-118
// Arrays follow the reference: [i][j] row pointers over i, contiguous in j.
#pragma once

struct Vector2D { double x, y; };
struct Euler2D_pState { double d; Vector2D v; double p; };
struct Euler2D_cState { double d; Vector2D dv; double E; };
struct Node2D_HO { Vector2D X; };

struct Grid2D_Quad_Block_HO {
    int NNi, INl, INu, NNj, JNl, JNu;
    int NCi, ICl, ICu, NCj, JCl, JCu, Nghost;
    Node2D_HO **Node;                       // [NNi][NNj]
    int *BCtypeN, *BCtypeS, *BCtypeE, *BCtypeW;   // [NCi] (N, S), [NCj] (E, W)
};

struct Euler2D_Quad_Block {
    int NCi, ICl, ICu, NCj, JCl, JCu, Nghost;
    Euler2D_pState **W;                     // [NCi][NCj]
    Euler2D_cState **U;
    Euler2D_pState *WoN, *WoS, *WoE, *WoW;  // [NCi] (N, S), [NCj] (E, W)
    Grid2D_Quad_Block_HO Grid;
};

#define ADAPTIVEBLOCK2D_USED 1
#define ADAPTIVEBLOCK2D_NUMBER_NEIGHBOURS_MAX 2
struct AdaptiveBlock2D_Info { int cpu, blknum, level; };
struct AdaptiveBlock2D {
    int used;
    AdaptiveBlock2D_Info info;
    int nN, nS, nE, nW;
    AdaptiveBlock2D_Info infoN[ADAPTIVEBLOCK2D_NUMBER_NEIGHBOURS_MAX], infoS[ADAPTIVEBLOCK2D_NUMBER_NEIGHBOURS_MAX],
        infoE[ADAPTIVEBLOCK2D_NUMBER_NEIGHBOURS_MAX], infoW[ADAPTIVEBLOCK2D_NUMBER_NEIGHBOURS_MAX];
    int nNW, nNE, nSE, nSW;
    AdaptiveBlock2D_Info infoNW[1], infoNE[1], infoSE[1], infoSW[1];
};
struct AdaptiveBlock2D_List {
    int ThisCPU, Nblk;
    AdaptiveBlock2D *Block;
};
struct AdaptiveBlockResourceList { int Nused; };

struct Euler2D_Input_Parameters {
    int i_Time_Integration, N_Stage;
    double CFL_Number;
    int i_Reconstruction, i_Limiter, i_Flux_Function, Local_Time_Stepping, Axisymmetric;
};
