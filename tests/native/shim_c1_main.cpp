// shim_c1_main.cpp -- drives the reference's explicit loop
// (Euler2DQuadSolvers.cc:549-898, Time_Accurate = 0, global time stepping)
// through the binding euler2d_hip_shim.cpp, with the mesh and initial state
// held in CFFC's block types (ref_types.h).
//
//   shim_c1 <case.bin> <nsteps> <out.bin>
//
// case.bin (little endian, written by tests/test_gpu_shim.py):
//   int32 nblk, int32 n_stage, float64 cfl,
//   int32 ni[nblk], nj[nblk], nbr[8 nblk], level[nblk],
//   int32 bc[nbc], float64 xnode[nnode], ynode[nnode], wo[4 nbc], W[4 ncell]
// with the arrays laid out as cffc_mesh_desc's (blocks concatenated, i fastest).
// out.bin: float64 W[4 ncell] (same order), then the L2 norm of every step.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/cffc_euler2d.h"
#include "ref_types.h"

int Euler2D_HIP_Attach(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &List, Euler2D_Input_Parameters &IP);
int Euler2D_HIP_Detach(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &List);
template <class Quad_Soln_Block>
int Send_All_Messages(Quad_Soln_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List,
                      const int Number_of_Solution_Variables, const int Send_Mesh_Geometry_Only);
void BCs(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List, Euler2D_Input_Parameters &Input_Parameters);
double CFL(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List, Euler2D_Input_Parameters &Input_Parameters);
void Set_Global_TimeStep(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List, const double &Dt_min);
double L2_Norm_Residual(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List);
int dUdt_Multistage_Explicit(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlockResourceList &Global_Soln_Block_List,
                             AdaptiveBlock2D_List &Soln_Block_List, Euler2D_Input_Parameters &Input_Parameters,
                             const int I_Stage);
void Apply_Boundary_Flux_Corrections_Multistage_Explicit(Euler2D_Quad_Block *Soln_ptr,
                                                         AdaptiveBlock2D_List &Soln_Block_List,
                                                         Euler2D_Input_Parameters &Input_Parameters,
                                                         const int I_Stage);
int Update_Solution_Multistage_Explicit(Euler2D_Quad_Block *Soln_ptr, AdaptiveBlock2D_List &Soln_Block_List,
                                        Euler2D_Input_Parameters &Input_Parameters, const int I_Stage);

template <class T> static bool rd(FILE *f, T *p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
template <class T> static T **rows(int ni, int nj) {
    T **r = new T *[ni];
    T *d = new T[(size_t)ni * nj];
    for (int i = 0; i < ni; ++i) r[i] = d + (size_t)i * nj;
    return r;
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: shim_c1 case.bin nsteps out.bin\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t nblk = 0, n_stage = 0;
    double cfl = 0.0;
    if (!rd(f, &nblk, 1) || !rd(f, &n_stage, 1) || !rd(f, &cfl, 1)) return 2;
    std::vector<int32_t> ni(nblk), nj(nblk), nbr(8 * (size_t)nblk), level(nblk);
    if (!rd(f, ni.data(), nblk) || !rd(f, nj.data(), nblk) || !rd(f, nbr.data(), 8 * (size_t)nblk) ||
        !rd(f, level.data(), nblk))
        return 2;
    size_t nbc = 0, nnode = 0, ncell = 0;
    for (int b = 0; b < nblk; ++b) {
        nbc += 2 * (size_t)(nj[b] + 4) + 2 * (size_t)(ni[b] + 4);
        nnode += (size_t)(ni[b] + 5) * (nj[b] + 5);
        ncell += (size_t)ni[b] * nj[b];
    }
    std::vector<int32_t> bc(nbc);
    std::vector<double> xn(nnode), yn(nnode), wo(4 * nbc), W(4 * ncell);
    if (!rd(f, bc.data(), nbc) || !rd(f, xn.data(), nnode) || !rd(f, yn.data(), nnode) || !rd(f, wo.data(), 4 * nbc) ||
        !rd(f, W.data(), 4 * ncell))
        return 2;
    fclose(f);

    // the blocks, as the reference holds them after its ICs
    std::vector<Euler2D_Quad_Block> blk(nblk);
    std::vector<AdaptiveBlock2D> ab(nblk);
    size_t oN = 0, oB = 0, oC = 0;
    for (int b = 0; b < nblk; ++b) {
        Euler2D_Quad_Block &B = blk[b];
        B.Nghost = 2;
        B.NCi = ni[b] + 4; B.ICl = 2; B.ICu = ni[b] + 1;
        B.NCj = nj[b] + 4; B.JCl = 2; B.JCu = nj[b] + 1;
        B.W = rows<Euler2D_pState>(B.NCi, B.NCj);
        B.U = rows<Euler2D_cState>(B.NCi, B.NCj);
        Grid2D_Quad_Block_HO &G = B.Grid;
        G.Nghost = 2;
        G.NCi = B.NCi; G.ICl = B.ICl; G.ICu = B.ICu; G.NCj = B.NCj; G.JCl = B.JCl; G.JCu = B.JCu;
        G.NNi = B.NCi + 1; G.INl = 2; G.INu = ni[b] + 2; G.NNj = B.NCj + 1; G.JNl = 2; G.JNu = nj[b] + 2;
        G.Node = rows<Node2D_HO>(G.NNi, G.NNj);
        for (int j = 0; j < G.NNj; ++j)
            for (int i = 0; i < G.NNi; ++i) G.Node[i][j].X = Vector2D{xn[oN + (size_t)j * G.NNi + i], yn[oN + (size_t)j * G.NNi + i]};
        oN += (size_t)G.NNi * G.NNj;
        int **bct[4] = {&G.BCtypeW, &G.BCtypeE, &G.BCtypeS, &G.BCtypeN};
        Euler2D_pState **wos[4] = {&B.WoW, &B.WoE, &B.WoS, &B.WoN};
        for (int s = 0; s < 4; ++s) {
            const int len = s < 2 ? B.NCj : B.NCi;
            *bct[s] = new int[len];
            *wos[s] = new Euler2D_pState[len];
            for (int k = 0; k < len; ++k, ++oB) {
                (*bct[s])[k] = bc[oB];
                (*wos[s])[k] = Euler2D_pState{wo[4 * oB], {wo[4 * oB + 1], wo[4 * oB + 2]}, wo[4 * oB + 3]};
            }
        }
        for (int j = B.JCl; j <= B.JCu; ++j)
            for (int i = B.ICl; i <= B.ICu; ++i, oC += 4)
                B.W[i][j] = Euler2D_pState{W[oC], {W[oC + 1], W[oC + 2]}, W[oC + 3]};
        AdaptiveBlock2D &A = ab[b];
        A.used = ADAPTIVEBLOCK2D_USED;
        A.info = AdaptiveBlock2D_Info{0, b, level[b]};
        int *cnt[8] = {&A.nN, &A.nS, &A.nE, &A.nW, &A.nNE, &A.nNW, &A.nSE, &A.nSW};
        AdaptiveBlock2D_Info *inf[8] = {A.infoN, A.infoS, A.infoE, A.infoW, A.infoNE, A.infoNW, A.infoSE, A.infoSW};
        for (int s = 0; s < 8; ++s) {
            const int o = nbr[8 * (size_t)b + s];
            *cnt[s] = o >= 0 ? 1 : 0;
            inf[s][0] = AdaptiveBlock2D_Info{0, o, o >= 0 ? level[o] : 0};
        }
    }
    AdaptiveBlock2D_List List{0, nblk, ab.data()};
    AdaptiveBlockResourceList Global{nblk};
    Euler2D_Input_Parameters IP{CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA, n_stage, cfl, CFFC_RECONSTRUCTION_LEAST_SQUARES,
                                CFFC_LIMITER_VENKATAKRISHNAN, CFFC_FLUX_FUNCTION_ROE, CFFC_GLOBAL_TIME_STEPPING, 0};
    Euler2D_Quad_Block *Soln_ptr = blk.data();
    if (int e = Euler2D_HIP_Attach(Soln_ptr, List, IP)) {
        fprintf(stderr, "attach failed: %d (%s)\n", e, cffc_last_error());
        return 1;
    }
    // Euler2DQuadSolvers.cc:549-898, explicit branch, steady-state form
    const int nsteps = atoi(argv[2]);
    std::vector<double> l2;
    for (int n = 0; n <= nsteps; ++n) {
        const double dTime = CFL(Soln_ptr, List, IP);
        Set_Global_TimeStep(Soln_ptr, List, dTime);
        if (n > 0) l2.push_back(L2_Norm_Residual(Soln_ptr, List));
        if (n == nsteps) break;
        for (int i_stage = 1; i_stage <= IP.N_Stage; ++i_stage) {
            int err = Send_All_Messages(Soln_ptr, List, 4, 0);
            BCs(Soln_ptr, List, IP);
            if (!err) err = dUdt_Multistage_Explicit(Soln_ptr, Global, List, IP, i_stage);
            Apply_Boundary_Flux_Corrections_Multistage_Explicit(Soln_ptr, List, IP, i_stage);
            if (!err) err = Update_Solution_Multistage_Explicit(Soln_ptr, List, IP, i_stage);
            if (err) {
                fprintf(stderr, "step %d stage %d: error %d (%s)\n", n, i_stage, err, cffc_last_error());
                return 1;
            }
        }
    }
    if (int e = Euler2D_HIP_Detach(Soln_ptr, List)) {
        fprintf(stderr, "detach failed: %d\n", e);
        return 1;
    }
    FILE *o = fopen(argv[3], "wb");
    if (!o) return 2;
    for (int b = 0; b < nblk; ++b) {
        const Euler2D_Quad_Block &B = blk[b];
        for (int j = B.JCl; j <= B.JCu; ++j)
            for (int i = B.ICl; i <= B.ICu; ++i) {
                const double w[4] = {B.W[i][j].d, B.W[i][j].v.x, B.W[i][j].v.y, B.W[i][j].p};
                fwrite(w, sizeof(double), 4, o);
            }
    }
    fwrite(l2.data(), sizeof(double), l2.size(), o);
    fclose(o);
    printf("shim_c1: %d blocks, %zu cells, %d steps through the reference-named calls\n", nblk, ncell, nsteps);
    return 0;
}
