// ref_kat_driver.cc -- TEST INFRASTRUCTURE ONLY.
//
// Known-answer driver for the reference's own state and CFD routines.  It is
// compiled against src_2D/Euler2D/Euler2DState.cc and src_2D/CFD/CFD.cc from
// the reference tree (see oracle/Makefile, target "ref") and evaluates them
// on input vectors read from stdin.  Each input line is
//     <tag> <hex-double> ...
// and each output line is
//     <tag> <hex-double> ...   (the inputs followed by the results)
// with doubles printed in C99 hex-float form so no bit is lost.
// Tags: ROE_N (Wl[4] Wr[4] nx ny -> F[4]), ROE_AVG (Wl[4] Wr[4] -> Wa[4]),
// VENKAT (uQuad[4] u0 umin umax -> phi), HARTEN (la ll lr -> pos neg),
// REFLECT (W[4] nx ny -> W[4]), BCCHARP (Wi[4] Wo[4] nx ny -> W[4]),
// W_OF_U (U[4] -> W[4]), U_OF_W (W[4] -> U[4]), FX (W[4] -> F[4]),
// RK (i n -> beta), OPTSM (i n limiter -> beta), HLLE_N (Wl[4] Wr[4] nx ny ->
// F[4]), BILIN (U1..U4 x1 y1 .. x4 y4 x0 y0 -> flag U0, scalar states),
// AXISRC (W[4] x y -> S[4], the axisymmetric source S(W, X)).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Euler2D/Euler2DState.h"

static std::vector<double> read_vals(char *p) {
    std::vector<double> v;
    char *tok = strtok(p, " \t\r\n");
    while (tok) {
        v.push_back(strtod(tok, 0));
        tok = strtok(0, " \t\r\n");
    }
    return v;
}

static void emit(const char *tag, const std::vector<double> &in, const double *out, int nout) {
    printf("%s", tag);
    for (size_t i = 0; i < in.size(); ++i) printf(" %a", in[i]);
    for (int i = 0; i < nout; ++i) printf(" %a", out[i]);
    printf("\n");
}

static Euler2D_pState P(const std::vector<double> &v, int o) { return Euler2D_pState(v[o], v[o + 1], v[o + 2], v[o + 3]); }

int main() {
    char line[4096];
    while (fgets(line, sizeof line, stdin)) {
        char tag[32];
        if (sscanf(line, "%31s", tag) != 1) continue;
        std::vector<double> v = read_vals(line + strlen(tag));
        double out[8];
        std::string t(tag);
        if (t == "ROE_N") {
            Euler2D_cState F = FluxRoe_n(P(v, 0), P(v, 4), Vector2D(v[8], v[9]));
            for (int k = 0; k < 4; ++k) out[k] = F[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "ROE_AVG") {
            Euler2D_pState Wa = RoeAverage(P(v, 0), P(v, 4));
            for (int k = 0; k < 4; ++k) out[k] = Wa[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "VENKAT") {
            double uq[4] = {v[0], v[1], v[2], v[3]};
            out[0] = Limiter_Venkatakrishnan(uq, v[4], v[5], v[6], 4);
            emit(tag, v, out, 1);
        } else if (t == "HARTEN") {
            out[0] = HartenFixPos(v[0], v[1], v[2]);
            out[1] = HartenFixNeg(v[0], v[1], v[2]);
            emit(tag, v, out, 2);
        } else if (t == "REFLECT") {
            Euler2D_pState W = Reflect(P(v, 0), Vector2D(v[4], v[5]));
            for (int k = 0; k < 4; ++k) out[k] = W[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "BCCHARP") {
            Euler2D_pState W = BC_Characteristic_Pressure(P(v, 0), P(v, 4), Vector2D(v[8], v[9]));
            for (int k = 0; k < 4; ++k) out[k] = W[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "W_OF_U") {
            Euler2D_cState U(v[0], v[1], v[2], v[3]);
            Euler2D_pState W = U.W();
            for (int k = 0; k < 4; ++k) out[k] = W[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "U_OF_W") {
            Euler2D_cState U = P(v, 0).U();
            for (int k = 0; k < 4; ++k) out[k] = U[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "FX") {
            Euler2D_cState F = P(v, 0).Fx();
            for (int k = 0; k < 4; ++k) out[k] = F[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "AXISRC") {
            Euler2D_cState Sx = S(P(v, 0), Vector2D(v[4], v[5]));
            for (int k = 0; k < 4; ++k) out[k] = Sx[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "HLLE_N") {
            Euler2D_cState F = FluxHLLE_n(P(v, 0), P(v, 4), Vector2D(v[8], v[9]));
            for (int k = 0; k < 4; ++k) out[k] = F[k + 1];
            emit(tag, v, out, 4);
        } else if (t == "BILIN") {
            double U0 = 0.0;
            int e = Bilinear_Interpolation(v[0], Vector2D(v[4], v[5]), v[1], Vector2D(v[6], v[7]), v[2],
                                           Vector2D(v[8], v[9]), v[3], Vector2D(v[10], v[11]),
                                           Vector2D(v[12], v[13]), U0);
            out[0] = (double)e;
            out[1] = e ? 0.0 : U0;
            emit(tag, v, out, 2);
        } else if (t == "RK") {
            out[0] = Runge_Kutta((int)v[0], (int)v[1]);
            emit(tag, v, out, 1);
        } else if (t == "OPTSM") {
            out[0] = MultiStage_Optimally_Smoothing((int)v[0], (int)v[1], (int)v[2]);
            emit(tag, v, out, 1);
        } else {
            fprintf(stderr, "unknown tag %s\n", tag);
            return 1;
        }
    }
    return 0;
}
