// euler2d_hip_shim.cpp -- the binding a CFFC maintainer adds next to
// Euler2DQuadMultiBlock.cc: the reference's multi-block functions, with the
// reference's signatures, routed to libcffc_hip.so through the C ABI
// (include/cffc_euler2d.h).  Compiled and run here against ref_types.h, which
// declares the members of CFFC's types that it reads.
//
//   Send_All_Messages          AdaptiveBlock2D_MessagePassing.h:7883
//   BCs                        Euler2DQuadMultiBlock.cc:937
//   CFL                        Euler2DQuadMultiBlock.cc:963
//   Set_Global_TimeStep        Euler2DQuadMultiBlock.cc:994
//   L1/L2/Max_Norm_Residual    Euler2DQuadMultiBlock.cc:1017-1104
//   Apply_Boundary_Flux_Corrections_Multistage_Explicit   :1307
//   dUdt_Multistage_Explicit   Euler2DQuadMultiBlock.cc:1346
//   Update_Solution_Multistage_Explicit                    :1383
//
// The device copy of a block array is attached once (after the ICs, where the
// reference would start its loop) and detached at the end, which writes the
// state back into the blocks' U and W.  Send_All_Messages also applies the
// BCs (the library fills ghosts by messages then BCs in one call, the
// reference's order); BCs itself then has nothing left to do.
#include <cstdio>
#include <map>
#include <vector>

#include "../../include/cffc_euler2d.h"
#include "ref_types.h"

namespace {
struct Attached {
    cffc_ctx *ctx = nullptr;
    int nblk = 0;
    int64_t ncell = 0;
};
std::map<const Euler2D_Quad_Block *, Attached> g_attached;

cffc_ctx *ctx_of(const Euler2D_Quad_Block *Soln_ptr) {
    auto it = g_attached.find(Soln_ptr);
    return it == g_attached.end() ? nullptr : it->second.ctx;
}
}  // namespace

// Build the mesh descriptor from the blocks and their list, create the
// device context and load the interior W (Soln_ptr[b].W after the ICs).
int Euler2D_HIP_Attach(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &List, Euler2D_Input_Parameters &IP) {
    const int nb = List.Nblk;
    std::vector<int32_t> ni(nb), nj(nb), nbr(8 * (size_t)nb, -1), nbr2(4 * (size_t)nb, -1), level(nb), bc;
    std::vector<double> xn, yn, wo, W;
    bool multilevel = false;
    for (int b = 0; b < nb; ++b) {
        if (List.Block[b].used != ADAPTIVEBLOCK2D_USED) return CFFC_E_ARG;
        const Euler2D_Quad_Block &B = Soln_ptr[b];
        if (B.Nghost != 2) return CFFC_E_ARG;
        ni[b] = B.ICu - B.ICl + 1;
        nj[b] = B.JCu - B.JCl + 1;
        for (int j = 0; j <= B.NCj; ++j)
            for (int i = 0; i <= B.NCi; ++i) {
                xn.push_back(B.Grid.Node[i][j].X.x);
                yn.push_back(B.Grid.Node[i][j].X.y);
            }
        // BCtypeW[ncj], BCtypeE[ncj], BCtypeS[nci], BCtypeN[nci] with the Wo states
        const int *bct[4] = {B.Grid.BCtypeW, B.Grid.BCtypeE, B.Grid.BCtypeS, B.Grid.BCtypeN};
        const Euler2D_pState *wos[4] = {B.WoW, B.WoE, B.WoS, B.WoN};
        for (int s = 0; s < 4; ++s)
            for (int k = 0; k < (s < 2 ? B.NCj : B.NCi); ++k) {
                bc.push_back(bct[s][k]);
                const Euler2D_pState &w = wos[s][k];
                wo.insert(wo.end(), {w.d, w.v.x, w.v.y, w.p});
            }
        for (int j = B.JCl; j <= B.JCu; ++j)
            for (int i = B.ICl; i <= B.ICu; ++i) {
                const Euler2D_pState &w = B.W[i][j];
                W.insert(W.end(), {w.d, w.v.x, w.v.y, w.p});
            }
        const AdaptiveBlock2D &A = List.Block[b];
        level[b] = A.info.level;
        multilevel = multilevel || A.info.level != List.Block[0].info.level;
        const int cnt[8] = {A.nN, A.nS, A.nE, A.nW, A.nNE, A.nNW, A.nSE, A.nSW};
        const AdaptiveBlock2D_Info *inf[8] = {A.infoN, A.infoS, A.infoE, A.infoW, A.infoNE, A.infoNW, A.infoSE, A.infoSW};
        for (int s = 0; s < 8; ++s) {
            if (cnt[s] >= 1) nbr[8 * b + s] = inf[s][0].blknum;
            if (s < 4 && cnt[s] == 2) nbr2[4 * b + s] = inf[s][1].blknum;
        }
    }
    cffc_mesh_desc m = {nb, 2, ni.data(), nj.data(), xn.data(), yn.data(), nbr.data(), bc.data(), wo.data(),
                        multilevel ? level.data() : nullptr, multilevel ? nbr2.data() : nullptr, nullptr};
    cffc_scheme_desc s = {IP.i_Time_Integration, IP.N_Stage, IP.CFL_Number, IP.i_Reconstruction, IP.i_Limiter,
                          IP.i_Flux_Function, IP.Local_Time_Stepping, CFFC_FLUX_FP64, IP.Axisymmetric};
    Attached a;
    if (int e = cffc_create(&m, &s, &a.ctx)) return e;
    a.nblk = nb;
    a.ncell = (int64_t)W.size() / 4;
    if (int e = cffc_set_W(a.ctx, W.data(), a.ncell)) {
        cffc_destroy(a.ctx);
        return e;
    }
    g_attached[Soln_ptr] = a;
    return 0;
}

// The device state back into the blocks' interior U and W; the context ends.
int Euler2D_HIP_Detach(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &List) {
    auto it = g_attached.find(Soln_ptr);
    if (it == g_attached.end()) return CFFC_E_ARG;
    Attached a = it->second;
    std::vector<double> U(4 * a.ncell), W(4 * a.ncell);
    int e = cffc_get_U(a.ctx, U.data(), a.ncell);
    if (!e) e = cffc_get_W(a.ctx, W.data(), a.ncell);
    size_t q = 0;
    for (int b = 0; b < List.Nblk && !e; ++b) {
        Euler2D_Quad_Block &B = Soln_ptr[b];
        for (int j = B.JCl; j <= B.JCu; ++j)
            for (int i = B.ICl; i <= B.ICu; ++i, q += 4) {
                B.U[i][j] = Euler2D_cState{U[q], {U[q + 1], U[q + 2]}, U[q + 3]};
                B.W[i][j] = Euler2D_pState{W[q], {W[q + 1], W[q + 2]}, W[q + 3]};
            }
    }
    cffc_destroy(a.ctx);
    g_attached.erase(it);
    return e;
}

template <class Quad_Soln_Block>
int Send_All_Messages(Quad_Soln_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List,
                      const int Number_of_Solution_Variables, const int Send_Mesh_Geometry_Only) {
    (void)Soln_Block_List;
    (void)Number_of_Solution_Variables;
    if (Send_Mesh_Geometry_Only) return 0;   // the geometry is resident on the device
    cffc_ctx *c = ctx_of(Soln_ptr);
    return c ? cffc_exchange_and_bcs(c) : 2200;
}
template int Send_All_Messages<Euler2D_Quad_Block>(Euler2D_Quad_Block *, AdaptiveBlock2D_List &, const int,
                                                   const int);

void BCs(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List, Euler2D_Input_Parameters &Input_Parameters) {
    (void)Soln_ptr;
    (void)Soln_Block_List;
    (void)Input_Parameters;   // applied by Send_All_Messages, after the messages
}

double CFL(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List, Euler2D_Input_Parameters &Input_Parameters) {
    (void)Soln_Block_List;
    (void)Input_Parameters;
    double dt = 0.0;
    if (cffc_ctx *c = ctx_of(Soln_ptr)) cffc_cfl(c, &dt, 0);
    return dt;
}

void Set_Global_TimeStep(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List, const double &Dt_min) {
    (void)Soln_Block_List;
    if (cffc_ctx *c = ctx_of(Soln_ptr)) cffc_set_global_timestep(c, Dt_min);
}

static double norm(Euler2D_Quad_Block *Soln_ptr, int k) {
    double n[3] = {0.0, 0.0, 0.0};
    if (cffc_ctx *c = ctx_of(Soln_ptr)) cffc_residual_norms(c, 1, n);
    return n[k];
}
double L1_Norm_Residual(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List) {
    (void)Soln_Block_List;
    return norm(Soln_ptr, 0);
}
double L2_Norm_Residual(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List) {
    (void)Soln_Block_List;
    return norm(Soln_ptr, 1);
}
double Max_Norm_Residual(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List) {
    (void)Soln_Block_List;
    return norm(Soln_ptr, 2);
}

int dUdt_Multistage_Explicit(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlockResourceList &Global_Soln_Block_List,
                             AdaptiveBlock2D_List &Soln_Block_List, Euler2D_Input_Parameters &Input_Parameters,
                             const int I_Stage) {
    (void)Global_Soln_Block_List;
    (void)Soln_Block_List;
    (void)Input_Parameters;
    cffc_ctx *c = ctx_of(Soln_ptr);
    return c ? cffc_dudt_stage(c, I_Stage) : 1;
}

void Apply_Boundary_Flux_Corrections_Multistage_Explicit(Euler2D_Quad_Block *Soln_ptr,
                                                         AdaptiveBlock2D_List &Soln_Block_List,
                                                         Euler2D_Input_Parameters &Input_Parameters,
                                                         const int I_Stage) {
    (void)Soln_Block_List;
    (void)Input_Parameters;
    if (cffc_ctx *c = ctx_of(Soln_ptr)) cffc_apply_flux_corrections(c, I_Stage);
}

int Update_Solution_Multistage_Explicit(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List,
                                        Euler2D_Input_Parameters &Input_Parameters, const int I_Stage) {
    (void)Soln_Block_List;
    (void)Input_Parameters;
    cffc_ctx *c = ctx_of(Soln_ptr);
    int64_t bad = -1;
    return c ? cffc_update_stage(c, I_Stage, &bad) : 1;
}
