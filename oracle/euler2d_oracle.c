/*
 * euler2d_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement, in plain C, of CFFC's explicit finite-volume residual and
 * multistage update for 2D Euler on quadrilateral multi-block meshes.  It is
 * the checker for the HIP implementation and nothing else: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path never calls it.
 *
 * Every function follows the reference's operation order so that, compiled
 * with -O2 -ffp-contract=off (strict IEEE double, no FMA contraction), it
 * reproduces the reference's arithmetic bit for bit.  Each function cites the
 * reference file:line it restates.  Pinning (see tests/):
 *   - state conversions, Roe flux, Harten fix, Venkatakrishnan limiter,
 *     Reflect, BC_Characteristic_Pressure, Runge_Kutta: bitwise against
 *     known-answer vectors written by the reference's own Euler2DState.cc /
 *     CFD.cc compiled from source (oracle/_ref) and committed under
 *     tests/golden/;
 *   - block-level functions (LSQ reconstruction, residual, update, CFL,
 *     norms, BCs, ghost exchange): against closed forms and invariants
 *     (linear-field exactness of the LSQ gradient, free-stream preservation
 *     on jittered meshes, discrete conservation, the diagonal mirror symmetry
 *     of the 4-quadrant problem, the uniform-mesh CFL closed form, the RK4
 *     combination, brute-force ghost copies, multi-block vs single-block
 *     equivalence).  The full reference binary needs LAPACK/Fortran
 *     libraries the image lacks, so no end-to-end reference trajectory pins
 *     these; the residual loop's accumulation order is pinned only by the
 *     single-block equivalence and conservation tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#include "../include/cffc_euler2d.h"

/* ---- constants: src_2D/Math/Math.h:40-53, Physics/GasConstants.h:66 ---- */
#define ZERO 0.0
#define ONE 1.0
#define TWO 2.0
#define FOUR 4.0
#define SIX 6.0
#define HALF 0.50
#define MILLION 1000000.00
#define TOLER 0.0000001
#define NANO 0.000000001
#define GAMMA_AIR 1.40

/* Euler2DState.cc:13-15 */
static const double g = GAMMA_AIR;
static const double gm1 = GAMMA_AIR - ONE;
static const double gm1i = ONE / (GAMMA_AIR - ONE);

/* Math.h:73-92 */
static inline double dmax(double x, double y) { return (x > y) ? x : y; }
static inline double dmin(double x, double y) { return (x < y) ? x : y; }
static inline double sqr(double x) { return x * x; }

typedef struct { double d, vx, vy, p; } pst; /* Euler2D_pState */
typedef struct { double d, mx, my, E; } cst; /* Euler2D_cState */

static inline double pget(const pst *w, int n) { return ((const double *)w)[n - 1]; }
static inline double *pref(pst *w, int n) { return ((double *)w) + (n - 1); }

/* ---- state functions: Euler2DState.h ------------------------------------ */
/* Euler2D_cState::W() Euler2DState.h:3429-3435, v() :1511, p() :1530 */
static inline pst W_of_U(cst U) {
    pst W;
    W.d = U.d;
    W.vx = U.mx / U.d;
    W.vy = U.my / U.d;
    W.p = gm1 * (U.E - HALF * (U.mx * U.mx + U.my * U.my) / U.d);
    return W;
}
/* Euler2D_pState::U() Euler2DState.h:1748-1762, dv() :1132, E() :1059 */
static inline cst U_of_W(pst W) {
    cst U;
    U.d = W.d;
    U.mx = W.d * W.vx;
    U.my = W.d * W.vy;
    U.E = W.p * gm1i + HALF * W.d * (W.vx * W.vx + W.vy * W.vy);
    return U;
}
/* h() Euler2DState.h:1070 */
static inline double W_h(pst W) { return g * W.p / (gm1 * W.d) + HALF * (W.vx * W.vx + W.vy * W.vy); }
/* H() :1081 */
static inline double W_H(pst W) { return g * gm1i * W.p + HALF * W.d * (W.vx * W.vx + W.vy * W.vy); }
/* axisymmetric source S(W, X), Euler2DState.h:3179-3182 (y the radius) */
static inline cst axisym_source(pst W, double y) {
    cst S = {-W.d * W.vy / y, -W.d * W.vx * W.vy / y, -W.d * (W.vy * W.vy) / y, -W.vy * W_H(W) / y};
    return S;
}
/* a() :1092, a2() :1103 */
static inline double W_a(pst W) { return sqrt(g * W.p / W.d); }
static inline double W_a2(pst W) { return g * W.p / W.d; }
/* cState e() :1556 */
static inline double U_e(cst U) { return U.E / U.d - HALF * (U.mx * U.mx + U.my * U.my) / sqr(U.d); }

/* Fx() Euler2DState.h:1845-1860 */
static inline cst W_Fx(pst W) {
    cst F;
    F.d = W.d * W.vx;
    F.mx = W.d * sqr(W.vx) + W.p;
    F.my = W.d * W.vx * W.vy;
    F.E = W.vx * W_H(W);
    return F;
}
/* lambda_x() Euler2DState.h:2158-2161 */
static inline pst W_lambda_x(pst W) {
    double c = W_a(W);
    pst l = {W.vx - c, W.vx, W.vx, W.vx + c};
    return l;
}
/* lp_x(index) Euler2DState.h:2968-2986 */
static inline pst W_lp_x(pst W, int index) {
    double c;
    pst r;
    switch (index) {
    case 2: r.d = ONE; r.vx = ZERO; r.vy = ZERO; r.p = -ONE / W_a2(W); return r;
    case 3: r.d = ZERO; r.vx = ZERO; r.vy = ONE; r.p = ZERO; return r;
    case 4: c = W_a(W); r.d = ZERO; r.vx = HALF * W.d / c; r.vy = ZERO; r.p = HALF / sqr(c); return r;
    default: c = W_a(W); r.d = ZERO; r.vx = -HALF * W.d / c; r.vy = ZERO; r.p = HALF / sqr(c); return r;
    }
}
/* rc_x(index) Euler2DState.h:2702-2720 */
static inline cst W_rc_x(pst W, int index) {
    double c;
    cst r;
    switch (index) {
    case 2: r.d = ONE; r.mx = W.vx; r.my = W.vy; r.E = HALF * (W.vx * W.vx + W.vy * W.vy); return r;
    case 3: r.d = ZERO; r.mx = ZERO; r.my = W.d; r.E = W.d * W.vy; return r;
    case 4: c = W_a(W); r.d = ONE; r.mx = W.vx + c; r.my = W.vy; r.E = W_h(W) + W.vx * c; return r;
    default: c = W_a(W); r.d = ONE; r.mx = W.vx - c; r.my = W.vy; r.E = W_h(W) - W.vx * c; return r;
    }
}
/* pState inner product, Euler2DState.h:1222-1224 (Vector2D * :362) */
static inline double pdot(pst a, pst b) { return a.d * b.d + (a.vx * b.vx + a.vy * b.vy) + a.p * b.p; }

/* ---- CFD.cc ------------------------------------------------------------- */
/* HartenFixPos CFD.cc:605-619 */
static double harten_fix_pos(double la, double ll, double lr) {
    double delta = dmax(ZERO, TWO * (lr - ll));
    if (fabs(la) > delta || delta < TOLER) {
        delta = ZERO;
    } else {
        delta = HALF * (sqr(la) / delta + delta) - fabs(la);
    }
    return HALF * (la + fabs(la) + delta);
}
/* HartenFixNeg CFD.cc:629-643 */
static double harten_fix_neg(double la, double ll, double lr) {
    double delta = dmax(ZERO, TWO * (lr - ll));
    if (fabs(la) > delta || delta < TOLER) {
        delta = ZERO;
    } else {
        delta = HALF * (sqr(la) / delta + delta) - fabs(la);
    }
    return HALF * (la - fabs(la) - delta);
}
/* Limiter_Venkatakrishnan CFD.cc:728-752 */
static double limiter_venkatakrishnan(const double *uQuad, double u0, double u0Min, double u0Max, int nQuad) {
    double phi = ONE, y;
    for (int i = 0; i <= nQuad - 1; i++) {
        if (uQuad[i] - u0 > TOLER) {
            y = (u0Max - u0) / (uQuad[i] - u0);
            phi = dmin(phi, (sqr(y) + TWO * y) / (sqr(y) + y + TWO));
        } else if (uQuad[i] - u0 < -TOLER) {
            y = (u0Min - u0) / (uQuad[i] - u0);
            phi = dmin(phi, (sqr(y) + TWO * y) / (sqr(y) + y + TWO));
        } else {
            phi = dmin(phi, ONE);
        }
    }
    return phi;
}
/* Runge_Kutta CFD.cc:901-939 */
static double runge_kutta(int I_stage, int N_stage) {
    double beta;
    switch (N_stage) {
    case 1: beta = ONE; break;
    case 2: beta = ONE; if (I_stage == 2) beta = HALF * beta; break;
    case 4:
        beta = ONE;
        if (I_stage == 1) beta = HALF * beta;
        if (I_stage == 2) beta = HALF * beta;
        if (I_stage == 4) beta = beta / SIX;
        break;
    case 5:
        beta = ONE;
        if (I_stage == 1) beta *= 0.059;
        if (I_stage == 2) beta *= 0.14;
        if (I_stage == 3) beta *= 0.273;
        if (I_stage == 4) beta *= HALF;
        break;
    default: beta = ONE; break;
    }
    return beta;
}
/* MultiStage_Optimally_Smoothing CFD.cc:953-1030: van Leer, Tai & Powell
   (1989) coefficients for LIMITER_ZERO (first order) and limited schemes. */
static double multistage_optimal_smoothing(int I_stage, int N_stage, int Limiter_Type) {
    double beta;
    switch (N_stage) {
    case 1: beta = ONE; break;
    case 2:
        if (Limiter_Type == CFFC_LIMITER_ZERO) { beta = ONE; if (I_stage == 1) beta = 0.3333 * beta; }
        else { beta = 0.4693; if (I_stage == 1) beta = 0.4242 * beta; }
        break;
    case 3:
        if (Limiter_Type == CFFC_LIMITER_ZERO) {
            beta = 1.50; if (I_stage == 1) beta = 0.1481 * beta; if (I_stage == 2) beta = 0.4000 * beta;
        } else {
            beta = 0.6936; if (I_stage == 1) beta = 0.1918 * beta; if (I_stage == 2) beta = 0.4929 * beta;
        }
        break;
    case 4:
        if (Limiter_Type == CFFC_LIMITER_ZERO) {
            beta = 2.00; if (I_stage == 1) beta = 0.0833 * beta; if (I_stage == 2) beta = 0.2069 * beta;
            if (I_stage == 3) beta = 0.4265 * beta;
        } else {
            beta = 0.9214; if (I_stage == 1) beta = 0.1084 * beta; if (I_stage == 2) beta = 0.2602 * beta;
            if (I_stage == 3) beta = 0.5052 * beta;
        }
        break;
    case 5:
        if (Limiter_Type == CFFC_LIMITER_ZERO) {
            beta = 2.50; if (I_stage == 1) beta = 0.0533 * beta; if (I_stage == 2) beta = 0.1263 * beta;
            if (I_stage == 3) beta = 0.2375 * beta; if (I_stage == 4) beta = 0.4414 * beta;
        } else {
            beta = 1.1508; if (I_stage == 1) beta = 0.0695 * beta; if (I_stage == 2) beta = 0.1602 * beta;
            if (I_stage == 3) beta = 0.2898 * beta; if (I_stage == 4) beta = 0.5060 * beta;
        }
        break;
    default: beta = ONE; break;
    }
    return beta;
}

/* ---- Euler2DState.cc ---------------------------------------------------- */
/* RoeAverage Euler2DState.cc:654-683 */
static pst roe_average(pst Wl, pst Wr) {
    double hl = W_h(Wl), hr = W_h(Wr);
    double sdl = sqrt(Wl.d), sdr = sqrt(Wr.d);
    double da = sdl * sdr;
    double ua = (sdl * Wl.vx + sdr * Wr.vx) / (sdl + sdr);
    double va = (sdl * Wl.vy + sdr * Wr.vy) / (sdl + sdr);
    double ha = (sdl * hl + sdr * hr) / (sdl + sdr);
    double aa2 = gm1 * (ha - HALF * (sqr(ua) + sqr(va)));
    double pa = da * aa2 / g;
    pst Wa = {da, ua, va, pa};
    return Wa;
}
/* HartenFixNeg/Pos (state versions) Euler2DState.cc:1872-1905 */
static pst harten_fix_neg_state(pst la, pst ll, pst lr) {
    pst r = {harten_fix_neg(la.d, ll.d, lr.d), HALF * (la.vx - fabs(la.vx)), HALF * (la.vy - fabs(la.vy)),
             harten_fix_neg(la.p, ll.p, lr.p)};
    return r;
}
static pst harten_fix_pos_state(pst la, pst ll, pst lr) {
    pst r = {harten_fix_pos(la.d, ll.d, lr.d), HALF * (la.vx + fabs(la.vx)), HALF * (la.vy + fabs(la.vy)),
             harten_fix_pos(la.p, ll.p, lr.p)};
    return r;
}
/* FluxRoe_x Euler2DState.cc:2311-2353 */
static cst flux_roe_x(pst Wl, pst Wr) {
    pst Wa = roe_average(Wl, Wr);
    pst dWrl = {Wr.d - Wl.d, Wr.vx - Wl.vx, Wr.vy - Wl.vy, Wr.p - Wl.p};
    pst lambdas_l = W_lambda_x(Wl), lambdas_r = W_lambda_x(Wr), lambdas_a = W_lambda_x(Wa);
    cst Flux;
    if (Wa.vx >= ZERO) {
        Flux = W_Fx(Wl);
        pst ws = harten_fix_neg_state(lambdas_a, lambdas_l, lambdas_r);
        for (int i = 1; i <= 4; ++i) {
            double w = pget(&ws, i);
            if (w < ZERO) {
                double s = w * pdot(W_lp_x(Wa, i), dWrl);
                cst rc = W_rc_x(Wa, i);
                Flux.d += s * rc.d; Flux.mx += s * rc.mx; Flux.my += s * rc.my; Flux.E += s * rc.E;
            }
        }
    } else {
        Flux = W_Fx(Wr);
        pst ws = harten_fix_pos_state(lambdas_a, lambdas_l, lambdas_r);
        for (int i = 1; i <= 4; ++i) {
            double w = pget(&ws, i);
            if (w > ZERO) {
                double s = w * pdot(W_lp_x(Wa, i), dWrl);
                cst rc = W_rc_x(Wa, i);
                Flux.d -= s * rc.d; Flux.mx -= s * rc.mx; Flux.my -= s * rc.my; Flux.E -= s * rc.E;
            }
        }
    }
    return Flux;
}
/* FluxRoe_n Euler2DState.cc:2446-2489 */
static cst flux_roe_n(pst Wl, pst Wr, double nx, double ny) {
    double cos_angle = nx, sin_angle = ny;
    pst Wlr = {Wl.d, Wl.vx * cos_angle + Wl.vy * sin_angle, -Wl.vx * sin_angle + Wl.vy * cos_angle, Wl.p};
    pst Wrr = {Wr.d, Wr.vx * cos_angle + Wr.vy * sin_angle, -Wr.vx * sin_angle + Wr.vy * cos_angle, Wr.p};
    cst Fr = flux_roe_x(Wlr, Wrr);
    cst F = {Fr.d, Fr.mx * cos_angle - Fr.my * sin_angle, Fr.mx * sin_angle + Fr.my * cos_angle, Fr.E};
    return F;
}
/* FluxHLLE_x Euler2DState.cc:2921-2963: wave speeds from the left, right and
   Roe-average eigenvalues, (sR*Fl - sL*Fr + sL*sR*(Ur - Ul)) / (sR - sL) */
static cst flux_hlle_x(pst Wl, pst Wr) {
    pst Wa = roe_average(Wl, Wr);
    cst Ul = U_of_W(Wl), Ur = U_of_W(Wr);
    cst dU = {Ur.d - Ul.d, Ur.mx - Ul.mx, Ur.my - Ul.my, Ur.E - Ul.E};
    pst ll = W_lambda_x(Wl), lr = W_lambda_x(Wr), la = W_lambda_x(Wa);
    double sl = dmin(ll.d, la.d), sr = dmax(lr.p, la.p);
    if (sl >= ZERO) return W_Fx(Wl);
    if (sr <= ZERO) return W_Fx(Wr);
    cst Fl = W_Fx(Wl), Fr = W_Fx(Wr), F;
    double slr = sl * sr, den = sr - sl;
    F.d = ((sr * Fl.d - sl * Fr.d) + slr * dU.d) / den;
    F.mx = ((sr * Fl.mx - sl * Fr.mx) + slr * dU.mx) / den;
    F.my = ((sr * Fl.my - sl * Fr.my) + slr * dU.my) / den;
    F.E = ((sr * Fl.E - sl * Fr.E) + slr * dU.E) / den;
    return F;
}
/* FluxHLLE_n Euler2DState.cc:3046-3095 (the rotation of FluxRoe_n) */
static cst flux_hlle_n(pst Wl, pst Wr, double nx, double ny) {
    double cos_angle = nx, sin_angle = ny;
    pst Wlr = {Wl.d, Wl.vx * cos_angle + Wl.vy * sin_angle, -Wl.vx * sin_angle + Wl.vy * cos_angle, Wl.p};
    pst Wrr = {Wr.d, Wr.vx * cos_angle + Wr.vy * sin_angle, -Wr.vx * sin_angle + Wr.vy * cos_angle, Wr.p};
    cst Fr = flux_hlle_x(Wlr, Wrr);
    cst F = {Fr.d, Fr.mx * cos_angle - Fr.my * sin_angle, Fr.mx * sin_angle + Fr.my * cos_angle, Fr.E};
    return F;
}
/* the flux-function switch of dUdt_Multistage_Explicit / dUdt_Residual_
   Evaluation (Euler2DQuad.cc:958-992) for the two functions built here */
static cst flux_n(const cffc_scheme_desc *s, pst Wl, pst Wr, double nx, double ny) {
    if (s->i_flux_function == CFFC_FLUX_FUNCTION_HLLE) return flux_hlle_n(Wl, Wr, nx, ny);
    return flux_roe_n(Wl, Wr, nx, ny);
}
/* Reflect Euler2DState.cc:751-787 */
static pst reflect(pst W, double nx, double ny) {
    double cos_angle = nx, sin_angle = ny;
    double dr = W.d;
    double ur = W.vx * cos_angle + W.vy * sin_angle;
    double vr = -W.vx * sin_angle + W.vy * cos_angle;
    double pr = W.p;
    ur = -ur;
    pst r = {dr, ur * cos_angle - vr * sin_angle, ur * sin_angle + vr * cos_angle, pr};
    return r;
}
/* BC_Characteristic_Pressure Euler2DState.cc:1383-1465 */
static pst bc_characteristic_pressure(pst Wi, pst Wo, double nx, double ny) {
    double cos_angle = nx, sin_angle = ny;
    pst Wir = {Wi.d, Wi.vx * cos_angle + Wi.vy * sin_angle, -Wi.vx * sin_angle + Wi.vy * cos_angle, Wi.p};
    pst Wor = {Wo.d, Wo.vx * cos_angle + Wo.vy * sin_angle, -Wo.vx * sin_angle + Wo.vy * cos_angle, Wo.p};
    double mi = Wir.vx / W_a(Wir);
    double db, ub, vb, pb, ab, ubr, vbr;
    if (mi >= ONE) {
        db = Wir.d; ubr = Wir.vx; vbr = Wir.vy; pb = Wir.p;
    } else if (mi >= ZERO) {
        pb = Wor.p;
        db = Wir.d * pow(pb / Wir.p, ONE / g);
        ab = sqrt(g * pb / db);
        ubr = Wir.vx + TWO * (W_a(Wir) - ab) * gm1i;
        vbr = Wir.vy;
    } else if (mi >= -ONE) {
        pb = Wor.p;
        db = Wor.d;
        vbr = Wor.vy;
        ab = sqrt(g * pb / db);
        ubr = Wir.vx + TWO * (W_a(Wir) - ab) * gm1i;
    } else {
        db = Wor.d; ubr = Wor.vx; vbr = Wor.vy; pb = Wor.p;
    }
    ub = ubr * cos_angle - vbr * sin_angle;
    vb = ubr * sin_angle + vbr * cos_angle;
    pst r = {db, ub, vb, pb};
    return r;
}

/* ====================================================================== */
/* Multi-block data                                                        */
/* ====================================================================== */
typedef struct {
    int ni, nj, ng, nci, ncj;
    int ICl, ICu, JCl, JCu;
    const double *xn, *yn; /* (nci+1)*(ncj+1) nodes */
    double *xc, *yc, *A;   /* per cell (Cell[i][j].Xc, .A) */
    int *bcW, *bcE, *bcS, *bcN;
    pst *WoW, *WoE, *WoS, *WoN;
    pst *W, *dWdx, *dWdy, *phi;
    cst *U, *Uo, *dUdt[3];
    double *dt;
    int nbr[8];
    int freeze_limiter;
    int level;             /* AdaptiveBlock2D_List Block[].info.level */
    int nbr2[4];           /* second neighbour of a face with two finer neighbours (infoX[1]) */
    cst *FluxN, *FluxS, *FluxE, *FluxW; /* block-edge fluxes, Euler2DQuad.h:140-205 */
    cst *P, *uo;           /* FAS_Multigrid_Quad_Block forcing term and uo_MG (multigrid levels) */
} oblk;

typedef struct orc_ctx {
    int nblk;
    oblk *b;
    cffc_scheme_desc s;
    int64_t ncell;
    double *xnode, *ynode; /* owned copies */
    int amr;               /* some neighbours differ in level */
    int max_level;
    double rs_epsilon;     /* Residual_Smoothing_Epsilon */
    int rs_iterations;     /* Residual_Smoothing_Gauss_Seidel_Iterations (0: off) */
} orc_ctx;

#define CI(B, i, j) ((j) * (B)->nci + (i))
#define NI(B, i, j) ((j) * ((B)->nci + 1) + (i))

/* HO_Grid2DQuad.h: face midpoints :1820-1840, lengths :1884-1900,
   normals :1972-2000 (computed from the nodes on every access) */
static inline void nodeX(const oblk *B, int i, int j, double *x, double *y) {
    *x = B->xn[NI(B, i, j)];
    *y = B->yn[NI(B, i, j)];
}
static inline void xfaceE(const oblk *B, int i, int j, double *x, double *y) {
    double x1, y1, x2, y2;
    nodeX(B, i + 1, j, &x1, &y1); nodeX(B, i + 1, j + 1, &x2, &y2);
    *x = HALF * (x1 + x2); *y = HALF * (y1 + y2);
}
static inline void xfaceW(const oblk *B, int i, int j, double *x, double *y) {
    double x1, y1, x2, y2;
    nodeX(B, i, j, &x1, &y1); nodeX(B, i, j + 1, &x2, &y2);
    *x = HALF * (x1 + x2); *y = HALF * (y1 + y2);
}
static inline void xfaceN(const oblk *B, int i, int j, double *x, double *y) {
    double x1, y1, x2, y2;
    nodeX(B, i, j + 1, &x1, &y1); nodeX(B, i + 1, j + 1, &x2, &y2);
    *x = HALF * (x1 + x2); *y = HALF * (y1 + y2);
}
static inline void xfaceS(const oblk *B, int i, int j, double *x, double *y) {
    double x1, y1, x2, y2;
    nodeX(B, i, j, &x1, &y1); nodeX(B, i + 1, j, &x2, &y2);
    *x = HALF * (x1 + x2); *y = HALF * (y1 + y2);
}
static inline double vabs(double x, double y) { return sqrt(x * x + y * y); }
/* lfaceN = abs(Node[i][j+1]-Node[i+1][j+1]), S = abs(Node[i][j]-Node[i+1][j]),
   E = abs(Node[i+1][j]-Node[i+1][j+1]), W = abs(Node[i][j]-Node[i][j+1]) */
static inline double lfaceE(const oblk *B, int i, int j) {
    double x1, y1, x2, y2;
    nodeX(B, i + 1, j, &x1, &y1); nodeX(B, i + 1, j + 1, &x2, &y2);
    return vabs(x1 - x2, y1 - y2);
}
static inline double lfaceW(const oblk *B, int i, int j) {
    double x1, y1, x2, y2;
    nodeX(B, i, j, &x1, &y1); nodeX(B, i, j + 1, &x2, &y2);
    return vabs(x1 - x2, y1 - y2);
}
static inline double lfaceN(const oblk *B, int i, int j) {
    double x1, y1, x2, y2;
    nodeX(B, i, j + 1, &x1, &y1); nodeX(B, i + 1, j + 1, &x2, &y2);
    return vabs(x1 - x2, y1 - y2);
}
static inline double lfaceS(const oblk *B, int i, int j) {
    double x1, y1, x2, y2;
    nodeX(B, i, j, &x1, &y1); nodeX(B, i + 1, j, &x2, &y2);
    return vabs(x1 - x2, y1 - y2);
}
/* cell_perimeter HO_Grid2DQuad.h:1909-1911 */
static inline double cell_perimeter(const oblk *B, int i, int j) {
    return lfaceE(B, i, j) + lfaceN(B, i, j) + lfaceW(B, i, j) + lfaceS(B, i, j);
}
static inline void nfaceE(const oblk *B, int i, int j, double *nx, double *ny) {
    if (lfaceE(B, i, j) > NANO * cell_perimeter(B, i, j)) {
        double x1, y1, x2, y2;
        nodeX(B, i + 1, j + 1, &x2, &y2); nodeX(B, i + 1, j, &x1, &y1);
        double l = vabs(x2 - x1, y2 - y1);
        *nx = (y2 - y1) / l; *ny = -(x2 - x1) / l;
    } else { *nx = ONE; *ny = ZERO; }
}
static inline void nfaceW(const oblk *B, int i, int j, double *nx, double *ny) {
    if (lfaceW(B, i, j) > NANO * cell_perimeter(B, i, j)) {
        double x1, y1, x2, y2;
        nodeX(B, i, j, &x1, &y1); nodeX(B, i, j + 1, &x2, &y2);
        double l = vabs(x1 - x2, y1 - y2);
        *nx = (y1 - y2) / l; *ny = -(x1 - x2) / l;
    } else { *nx = ONE; *ny = ZERO; }
}
static inline void nfaceN(const oblk *B, int i, int j, double *nx, double *ny) {
    if (lfaceN(B, i, j) > NANO * cell_perimeter(B, i, j)) {
        double x1, y1, x2, y2;
        nodeX(B, i, j + 1, &x1, &y1); nodeX(B, i + 1, j + 1, &x2, &y2);
        double l = vabs(x1 - x2, y1 - y2);
        *nx = (y1 - y2) / l; *ny = -(x1 - x2) / l;
    } else { *nx = ONE; *ny = ZERO; }
}
static inline void nfaceS(const oblk *B, int i, int j, double *nx, double *ny) {
    if (lfaceS(B, i, j) > NANO * cell_perimeter(B, i, j)) {
        double x1, y1, x2, y2;
        nodeX(B, i + 1, j, &x2, &y2); nodeX(B, i, j, &x1, &y1);
        double l = vabs(x2 - x1, y2 - y1);
        *nx = (y2 - y1) / l; *ny = -(x2 - x1) / l;
    } else { *nx = ONE; *ny = ZERO; }
}
/* quadCentroid Math/Math.h:311-350 with vertices SW, SE, NE, NW
   (quadAreaAndCentroid HO_Grid2DQuad.h:1403-1410, global coordinates since
   Minimize_Error_Calculation_Of_Geometric_Properties defaults to OFF,
   HO_Grid2DQuad.cc:80) */
static void quad_centroid(const oblk *B, int i, int j, double *xc, double *yc, double *A) {
    double X0x, X0y, X1x, X1y, X2x, X2y, X3x, X3y;
    nodeX(B, i, j, &X0x, &X0y); nodeX(B, i + 1, j, &X1x, &X1y);
    nodeX(B, i + 1, j + 1, &X2x, &X2y); nodeX(B, i, j + 1, &X3x, &X3y);
    double ai, atmp = 0, xtmp = 0, ytmp = 0;
    ai = X3x * X0y - X0x * X3y; atmp += ai; xtmp += (X0x + X3x) * ai; ytmp += (X0y + X3y) * ai;
    ai = X0x * X1y - X1x * X0y; atmp += ai; xtmp += (X1x + X0x) * ai; ytmp += (X1y + X0y) * ai;
    ai = X1x * X2y - X2x * X1y; atmp += ai; xtmp += (X2x + X1x) * ai; ytmp += (X2y + X1y) * ai;
    ai = X2x * X3y - X3x * X2y; atmp += ai; xtmp += (X3x + X2x) * ai; ytmp += (X3y + X2y) * ai;
    *A = 0.5 * atmp;
    if (atmp != 0) { *xc = xtmp / (3 * atmp); *yc = ytmp / (3 * atmp); }
    else { *xc = 0; *yc = 0; }
}

/* ---- BCs: Euler2DQuadSingleBlock.cc:2354-3010 (supported types only) ---- */
static int bc_supported(int bc) {
    return bc == CFFC_BC_NONE || bc == CFFC_BC_FIXED || bc == CFFC_BC_CONSTANT_EXTRAPOLATION ||
           bc == CFFC_BC_REFLECTION || bc == CFFC_BC_PERIODIC || bc == CFFC_BC_CHARACTERISTIC;
}

static void set_ghost_W(oblk *B, int i, int j, pst W) {
    B->W[CI(B, i, j)] = W;
    B->U[CI(B, i, j)] = U_of_W(W);
}
static void copy_cell(oblk *B, int id, int jd, int is, int js) {
    B->W[CI(B, id, jd)] = B->W[CI(B, is, js)];
    B->U[CI(B, id, jd)] = B->U[CI(B, is, js)];
}

static void block_bcs(oblk *B) {
    int ICl = B->ICl, ICu = B->ICu, JCl = B->JCl, JCu = B->JCu, Ng = B->ng;
    double nx, ny;
    for (int j = JCl - Ng; j <= JCu + Ng; ++j) {
        /* West :2363-2519 */
        if ((j >= JCl && j <= JCu) || (j < JCl && B->bcS[ICl - 1] == CFFC_BC_NONE) ||
            (j > JCu && B->bcN[ICl - 1] == CFFC_BC_NONE)) {
            switch (B->bcW[j]) {
            case CFFC_BC_NONE: break;
            case CFFC_BC_FIXED:
                for (int gh = 1; gh <= Ng; ++gh) set_ghost_W(B, ICl - gh, j, B->WoW[j]);
                break;
            case CFFC_BC_CONSTANT_EXTRAPOLATION:
                for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, ICl - gh, j, ICl, j);
                break;
            case CFFC_BC_REFLECTION:
                nfaceW(B, ICl, j, &nx, &ny);
                for (int gh = 1; gh <= Ng; ++gh)
                    set_ghost_W(B, ICl - gh, j, reflect(B->W[CI(B, ICl + gh - 1, j)], nx, ny));
                break;
            case CFFC_BC_PERIODIC:
                for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, ICl - gh, j, ICu - gh + 1, j);
                break;
            case CFFC_BC_CHARACTERISTIC:
                nfaceW(B, ICl, j, &nx, &ny);
                set_ghost_W(B, ICl - 1, j, bc_characteristic_pressure(B->W[CI(B, ICl, j)], B->WoW[j], nx, ny));
                for (int gh = 2; gh <= Ng; ++gh) copy_cell(B, ICl - gh, j, ICl - 1, j);
                break;
            default:
                for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, ICl - gh, j, ICl, j);
                break;
            }
        }
        /* East :2520-2675 */
        if ((j >= JCl && j <= JCu) || (j < JCl && B->bcS[ICu + 1] == CFFC_BC_NONE) ||
            (j > JCu && B->bcN[ICu + 1] == CFFC_BC_NONE)) {
            switch (B->bcE[j]) {
            case CFFC_BC_NONE: break;
            case CFFC_BC_FIXED:
                for (int gh = 1; gh <= Ng; ++gh) set_ghost_W(B, ICu + gh, j, B->WoE[j]);
                break;
            case CFFC_BC_CONSTANT_EXTRAPOLATION:
                for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, ICu + gh, j, ICu, j);
                break;
            case CFFC_BC_REFLECTION:
                nfaceE(B, ICu, j, &nx, &ny);
                for (int gh = 1; gh <= Ng; ++gh)
                    set_ghost_W(B, ICu + gh, j, reflect(B->W[CI(B, ICu - gh + 1, j)], nx, ny));
                break;
            case CFFC_BC_PERIODIC:
                for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, ICu + gh, j, ICl + gh - 1, j);
                break;
            case CFFC_BC_CHARACTERISTIC:
                nfaceE(B, ICu, j, &nx, &ny);
                set_ghost_W(B, ICu + 1, j, bc_characteristic_pressure(B->W[CI(B, ICu, j)], B->WoE[j], nx, ny));
                for (int gh = 2; gh <= Ng; ++gh) copy_cell(B, ICu + gh, j, ICu + 1, j);
                break;
            default:
                for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, ICu + gh, j, ICu, j);
                break;
            }
        }
    }
    for (int i = ICl - Ng; i <= ICu + Ng; ++i) {
        /* South :2681-2837 */
        switch (B->bcS[i]) {
        case CFFC_BC_NONE: break;
        case CFFC_BC_FIXED:
            for (int gh = 1; gh <= Ng; ++gh) set_ghost_W(B, i, JCl - gh, B->WoS[i]);
            break;
        case CFFC_BC_CONSTANT_EXTRAPOLATION:
            for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, i, JCl - gh, i, JCl);
            break;
        case CFFC_BC_REFLECTION:
            nfaceS(B, i, JCl, &nx, &ny);
            for (int gh = 1; gh <= Ng; ++gh)
                set_ghost_W(B, i, JCl - gh, reflect(B->W[CI(B, i, JCl + gh - 1)], nx, ny));
            break;
        case CFFC_BC_PERIODIC:
            for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, i, JCl - gh, i, JCu - gh + 1);
            break;
        case CFFC_BC_CHARACTERISTIC:
            nfaceS(B, i, JCl, &nx, &ny);
            set_ghost_W(B, i, JCl - 1, bc_characteristic_pressure(B->W[CI(B, i, JCl)], B->WoS[i], nx, ny));
            for (int gh = 2; gh <= Ng; ++gh) copy_cell(B, i, JCl - gh, i, JCl - 1);
            break;
        default:
            for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, i, JCl - gh, i, JCl);
            break;
        }
        /* North :2838-2989 */
        switch (B->bcN[i]) {
        case CFFC_BC_NONE: break;
        case CFFC_BC_FIXED:
            for (int gh = 1; gh <= Ng; ++gh) set_ghost_W(B, i, JCu + gh, B->WoN[i]);
            break;
        case CFFC_BC_CONSTANT_EXTRAPOLATION:
            for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, i, JCu + gh, i, JCu);
            break;
        case CFFC_BC_REFLECTION:
            nfaceN(B, i, JCu, &nx, &ny);
            for (int gh = 1; gh <= Ng; ++gh)
                set_ghost_W(B, i, JCu + gh, reflect(B->W[CI(B, i, JCu - gh + 1)], nx, ny));
            break;
        case CFFC_BC_PERIODIC:
            for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, i, JCu + gh, i, JCl + gh - 1);
            break;
        case CFFC_BC_CHARACTERISTIC:
            nfaceN(B, i, JCu, &nx, &ny);
            set_ghost_W(B, i, JCu + 1, bc_characteristic_pressure(B->W[CI(B, i, JCu)], B->WoN[i], nx, ny));
            for (int gh = 2; gh <= Ng; ++gh) copy_cell(B, i, JCu + gh, i, JCu + 1);
            break;
        default:
            for (int gh = 1; gh <= Ng; ++gh) copy_cell(B, i, JCu + gh, i, JCu);
            break;
        }
    }
}

/* ---- same-level ghost exchange ----------------------------------------
   Load_Send_Message_Buffers_NoResChange (AdaptiveBlock2D_MessagePassing.h:32),
   Exchange_Messages_NoResChange (AdaptiveBlock_NoMPI.cc:28),
   Unload_Receive_Message_Buffers_NoResChange (:5680); LoadSendBuffer /
   UnloadReceiveBuffer Euler2DQuad.h:1570-1651 (U transferred, then W=W(U)).
   Faces carry Nghost layers over the interior span, corners Nghost x Nghost.
   Neighbours share orientation (dimen.i, dimen.j > 0). */
static void exchange_block(oblk *all, int b) {
    oblk *B = &all[b];
    int Ng = B->ng;
    for (int slot = 0; slot < 8; ++slot) {
        int nb = B->nbr[slot];
        if (nb < 0 || all[nb].level != B->level) continue;
        const oblk *S = &all[nb];
        int i0, i1, j0, j1; /* destination ranges in B */
        int di, dj;         /* source index = dest index + (di,dj) */
        switch (slot) {
        case CFFC_NBR_N: i0 = B->ICl; i1 = B->ICu; j0 = B->JCu + 1; j1 = B->JCu + Ng; di = S->ICl - B->ICl; dj = S->JCl - (B->JCu + 1); break;
        case CFFC_NBR_S: i0 = B->ICl; i1 = B->ICu; j0 = B->JCl - Ng; j1 = B->JCl - 1; di = S->ICl - B->ICl; dj = S->JCu - (B->JCl - 1); break;
        case CFFC_NBR_E: i0 = B->ICu + 1; i1 = B->ICu + Ng; j0 = B->JCl; j1 = B->JCu; di = S->ICl - (B->ICu + 1); dj = S->JCl - B->JCl; break;
        case CFFC_NBR_W: i0 = B->ICl - Ng; i1 = B->ICl - 1; j0 = B->JCl; j1 = B->JCu; di = S->ICu - (B->ICl - 1); dj = S->JCl - B->JCl; break;
        case CFFC_NBR_NE: i0 = B->ICu + 1; i1 = B->ICu + Ng; j0 = B->JCu + 1; j1 = B->JCu + Ng; di = S->ICl - (B->ICu + 1); dj = S->JCl - (B->JCu + 1); break;
        case CFFC_NBR_NW: i0 = B->ICl - Ng; i1 = B->ICl - 1; j0 = B->JCu + 1; j1 = B->JCu + Ng; di = S->ICu - (B->ICl - 1); dj = S->JCl - (B->JCu + 1); break;
        case CFFC_NBR_SE: i0 = B->ICu + 1; i1 = B->ICu + Ng; j0 = B->JCl - Ng; j1 = B->JCl - 1; di = S->ICl - (B->ICu + 1); dj = S->JCu - (B->JCl - 1); break;
        default: /* SW */ i0 = B->ICl - Ng; i1 = B->ICl - 1; j0 = B->JCl - Ng; j1 = B->JCl - 1; di = S->ICu - (B->ICl - 1); dj = S->JCu - (B->JCl - 1); break;
        }
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
                cst U = S->U[CI(S, i + di, j + dj)];
                B->U[CI(B, i, j)] = U;
                B->W[CI(B, i, j)] = W_of_U(U);
            }
    }
}

/* ---- LSQ reconstruction: Euler2DQuadSingleBlock.cc:3678-4087 ---------- */
static void lsq_reconstruct(oblk *B, int i, int j, int Limiter) {
    int n_pts, ii[8], jj[8];
    int Ng = B->ng, ICl = B->ICl, ICu = B->ICu, JCl = B->JCl, JCu = B->JCu;
#define SET(k, a, c) do { ii[k] = (a); jj[k] = (c); } while (0)
#define FULL8() do { n_pts = 8; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i+1, j-1); SET(3, i-1, j); \
                     SET(4, i+1, j); SET(5, i-1, j+1); SET(6, i, j+1); SET(7, i+1, j+1); } while (0)
    int bcw = B->bcW[j], bce = B->bcE[j], bcs = B->bcS[i], bcn = B->bcN[i];
#define EXTRAP(bc) ((bc) == CFFC_BC_PERIODIC || (bc) == CFFC_BC_CONSTANT_EXTRAPOLATION || \
                    (bc) == 4002 /* LINEAR_EXTRAPOLATION */ || (bc) == CFFC_BC_CHARACTERISTIC)
    if (i == ICl - Ng || i == ICu + Ng || j == JCl - Ng || j == JCu + Ng) {
        n_pts = 0;
    } else if (i == ICl - Ng + 1 && bcw != CFFC_BC_NONE) {
        if (j == JCl - Ng + 1 || j == JCu + Ng - 1) n_pts = 0;
        else if (EXTRAP(bcw)) {
            if (j == JCl) { n_pts = 5; SET(0, i-1, j); SET(1, i+1, j); SET(2, i-1, j+1); SET(3, i, j+1); SET(4, i+1, j+1); }
            else if (j == JCu) { n_pts = 5; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i+1, j-1); SET(3, i-1, j); SET(4, i+1, j); }
            else FULL8();
        } else {
            if (j == JCl) { n_pts = 3; SET(0, i+1, j); SET(1, i, j+1); SET(2, i+1, j+1); }
            else if (j == JCu) { n_pts = 3; SET(0, i, j-1); SET(1, i+1, j-1); SET(2, i+1, j); }
            else { n_pts = 5; SET(0, i, j-1); SET(1, i+1, j-1); SET(2, i+1, j); SET(3, i, j+1); SET(4, i+1, j+1); }
        }
    } else if (i == ICu + Ng - 1 && bce != CFFC_BC_NONE) {
        if (j == JCl - Ng + 1 || j == JCu + Ng - 1) n_pts = 0;
        else if (EXTRAP(bce)) {
            if (j == JCl) { n_pts = 5; SET(0, i-1, j); SET(1, i+1, j); SET(2, i-1, j+1); SET(3, i, j+1); SET(4, i+1, j+1); }
            else if (j == JCu) { n_pts = 5; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i+1, j-1); SET(3, i-1, j); SET(4, i+1, j); }
            else FULL8();
        } else {
            if (j == JCl) { n_pts = 3; SET(0, i-1, j); SET(1, i-1, j+1); SET(2, i, j+1); }
            else if (j == JCu) { n_pts = 3; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i-1, j); }
            else { n_pts = 5; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i-1, j); SET(3, i-1, j+1); SET(4, i, j+1); }
        }
    } else if (j == JCl - Ng + 1 && bcs != CFFC_BC_NONE) {
        if (i == ICl - Ng + 1 || i == ICu + Ng - 1) n_pts = 0;
        else if (EXTRAP(bcs)) {
            if (i == ICl) { n_pts = 5; SET(0, i, j-1); SET(1, i+1, j-1); SET(2, i+1, j); SET(3, i, j+1); SET(4, i+1, j+1); }
            else if (i == ICu) { n_pts = 5; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i-1, j); SET(3, i-1, j+1); SET(4, i, j+1); }
            else FULL8();
        } else {
            if (i == ICl) { n_pts = 3; SET(0, i+1, j); SET(1, i, j+1); SET(2, i+1, j+1); }
            else if (i == ICu) { n_pts = 3; SET(0, i-1, j); SET(1, i-1, j+1); SET(2, i, j+1); }
            else { n_pts = 5; SET(0, i-1, j); SET(1, i+1, j); SET(2, i-1, j+1); SET(3, i, j+1); SET(4, i+1, j+1); }
        }
    } else if (j == JCu + Ng - 1 && bcn != CFFC_BC_NONE) {
        if (i == ICl - Ng + 1 || i == ICu + Ng - 1) n_pts = 0;
        else if (EXTRAP(bcn)) {
            if (i == ICl) { n_pts = 5; SET(0, i, j-1); SET(1, i+1, j-1); SET(2, i+1, j); SET(3, i, j+1); SET(4, i+1, j+1); }
            else if (i == ICu) { n_pts = 5; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i-1, j); SET(3, i-1, j+1); SET(4, i, j+1); }
            else FULL8();
        } else {
            if (i == ICl) { n_pts = 3; SET(0, i, j-1); SET(1, i+1, j-1); SET(2, i+1, j); }
            else if (i == ICu) { n_pts = 3; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i-1, j); }
            else { n_pts = 5; SET(0, i-1, j-1); SET(1, i, j-1); SET(2, i+1, j-1); SET(3, i-1, j); SET(4, i+1, j); }
        }
    } else {
        FULL8();
    }
#undef SET
#undef FULL8
#undef EXTRAP
    int c = CI(B, i, j);
    if (n_pts > 0) {
        pst DUDx = {0, 0, 0, 0}, DUDy = {0, 0, 0, 0};
        double DxDx = ZERO, DxDy = ZERO, DyDy = ZERO;
        pst W0 = B->W[c];
        for (int n2 = 0; n2 <= n_pts - 1; ++n2) {
            int cn = CI(B, ii[n2], jj[n2]);
            double dXx = B->xc[cn] - B->xc[c], dXy = B->yc[cn] - B->yc[c];
            pst Wn = B->W[cn];
            pst DU = {Wn.d - W0.d, Wn.vx - W0.vx, Wn.vy - W0.vy, Wn.p - W0.p};
            DUDx.d += DU.d * dXx; DUDx.vx += DU.vx * dXx; DUDx.vy += DU.vy * dXx; DUDx.p += DU.p * dXx;
            DUDy.d += DU.d * dXy; DUDy.vx += DU.vx * dXy; DUDy.vy += DU.vy * dXy; DUDy.p += DU.p * dXy;
            DxDx += dXx * dXx; DxDy += dXx * dXy; DyDy += dXy * dXy;
        }
        double np = (double)n_pts;
        DUDx.d /= np; DUDx.vx /= np; DUDx.vy /= np; DUDx.p /= np;
        DUDy.d /= np; DUDy.vx /= np; DUDy.vy /= np; DUDy.p /= np;
        DxDx /= np; DxDy /= np; DyDy /= np;
        double det = DxDx * DyDy - DxDy * DxDy;
        pst gx, gy;
        /* (DUDx*DyDy - DUDy*DxDy)/det : pState*double, pState-pState, pState/double */
        gx.d = (DUDx.d * DyDy - DUDy.d * DxDy) / det;
        gx.vx = (DUDx.vx * DyDy - DUDy.vx * DxDy) / det;
        gx.vy = (DUDx.vy * DyDy - DUDy.vy * DxDy) / det;
        gx.p = (DUDx.p * DyDy - DUDy.p * DxDy) / det;
        gy.d = (DUDy.d * DxDx - DUDx.d * DxDy) / det;
        gy.vx = (DUDy.vx * DxDx - DUDx.vx * DxDy) / det;
        gy.vy = (DUDy.vy * DxDx - DUDx.vy * DxDy) / det;
        gy.p = (DUDy.p * DxDx - DUDx.p * DxDy) / det;
        B->dWdx[c] = gx;
        B->dWdy[c] = gy;
        if (!B->freeze_limiter) {
            double xE, yE, xW, yW, xN, yN, xS, yS;
            xfaceE(B, i, j, &xE, &yE); xfaceW(B, i, j, &xW, &yW);
            xfaceN(B, i, j, &xN, &yN); xfaceS(B, i, j, &xS, &yS);
            double xc0 = B->xc[c], yc0 = B->yc[c];
            for (int n = 1; n <= 4; ++n) {
                double u0 = pget(&W0, n), u0Min = u0, u0Max = u0;
                for (int n2 = 0; n2 <= n_pts - 1; ++n2) {
                    double v = pget(&B->W[CI(B, ii[n2], jj[n2])], n);
                    u0Min = dmin(u0Min, v);
                    u0Max = dmax(u0Max, v);
                }
                double gxn = pget(&gx, n), gyn = pget(&gy, n), uQuad[4], phi;
                uQuad[0] = u0 + gxn * (xE - xc0) + gyn * (yE - yc0);
                uQuad[1] = u0 + gxn * (xW - xc0) + gyn * (yW - yc0);
                uQuad[2] = u0 + gxn * (xN - xc0) + gyn * (yN - yc0);
                uQuad[3] = u0 + gxn * (xS - xc0) + gyn * (yS - yc0);
                switch (Limiter) {
                case CFFC_LIMITER_ONE: phi = ONE; break;
                case CFFC_LIMITER_ZERO: phi = ZERO; break;
                default: phi = limiter_venkatakrishnan(uQuad, u0, u0Min, u0Max, 4); break;
                }
                *pref(&B->phi[c], n) = phi;
            }
        }
    } else {
        pst z = {0, 0, 0, 0};
        B->dWdx[c] = z;
        B->dWdy[c] = z;
        B->phi[c] = z;
    }
}

/* W + (phi^dWdx)*dX.x + (phi^dWdy)*dX.y  (Euler2DQuad.cc:1338-1340) */
static inline pst face_state(const oblk *B, int c, double dXx, double dXy) {
    pst W = B->W[c], P = B->phi[c], gx = B->dWdx[c], gy = B->dWdy[c], r;
    r.d = W.d + (P.d * gx.d) * dXx + (P.d * gy.d) * dXy;
    r.vx = W.vx + (P.vx * gx.vx) * dXx + (P.vx * gy.vx) * dXy;
    r.vy = W.vy + (P.vy * gx.vy) * dXx + (P.vy * gy.vy) * dXy;
    r.p = W.p + (P.p * gx.p) * dXx + (P.p * gy.p) * dXy;
    return r;
}

static int stage_k_residual(const cffc_scheme_desc *s, int i_stage) {
    if (s->i_time_integration == CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA && s->n_stage == 4)
        return (i_stage == 4) ? 0 : i_stage - 1;
    return 0;
}
static double stage_omega(const cffc_scheme_desc *s, int i_stage) {
    if (s->i_time_integration == CFFC_TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING)
        return multistage_optimal_smoothing(i_stage, s->n_stage, s->i_limiter);
    return runge_kutta(i_stage, s->n_stage);
}

/* dUdt -= ((CFL*dt)*Flux*l)/A  and  += variant (Euler2DQuad.cc:1502-1509) */
static inline void acc_sub(cst *d, double cdt, cst F, double l, double A) {
    d->d -= cdt * F.d * l / A; d->mx -= cdt * F.mx * l / A; d->my -= cdt * F.my * l / A; d->E -= cdt * F.E * l / A;
}
static inline void acc_add(cst *d, double cdt, cst F, double l, double A) {
    d->d += cdt * F.d * l / A; d->mx += cdt * F.mx * l / A; d->my += cdt * F.my * l / A; d->E += cdt * F.E * l / A;
}

static int is_face_bc(int bc) { return bc == CFFC_BC_REFLECTION || bc == CFFC_BC_CHARACTERISTIC; }

/* sign*Flux*l as the reference writes the saved boundary fluxes: FluxE/N =
   Flux*lface, FluxW/S = -Flux*lface (the unary minus first, exact) */
static inline cst flux_times(cst F, double sign, double l) {
    cst r;
    r.d = (sign * F.d) * l; r.mx = (sign * F.mx) * l; r.my = (sign * F.my) * l; r.E = (sign * F.E) * l;
    return r;
}

/* Euler2D_Quad_Block::dUdt_Multistage_Explicit, Euler2DQuad.cc:1235-1727 */
static void block_dudt(oblk *B, const cffc_scheme_desc *s, int i_stage) {
    int ICl = B->ICl, ICu = B->ICu, JCl = B->JCl, JCu = B->JCu, Ng = B->ng;
    int k = stage_k_residual(s, i_stage);
    const cst Z = {0, 0, 0, 0};
    double CFL = s->cfl_number;
    /* Linear_Reconstruction_LeastSquares, SingleBlock.cc:4467-4482 */
    for (int j = JCl - Ng + 1; j <= JCu + Ng - 1; ++j)
        for (int i = ICl - Ng + 1; i <= ICu + Ng - 1; ++i) lsq_reconstruct(B, i, j, s->i_limiter);
    /* i-direction fluxes :1304-1534 */
    for (int j = JCl - 1; j <= JCu + 1; ++j) {
        if (i_stage == 1) B->Uo[CI(B, ICl - 1, j)] = B->U[CI(B, ICl - 1, j)];
        B->dUdt[k][CI(B, ICl - 1, j)] = Z;
        for (int i = ICl - 1; i <= ICu; ++i) {
            int cL = CI(B, i, j), cR = CI(B, i + 1, j);
            if (i_stage == 1) {
                B->Uo[cR] = B->U[cR];
                B->dUdt[k][cR] = Z;
            } else if (j > JCl - 1 && j < JCu + 1) {
                int ti = s->i_time_integration;
                if (ti == CFFC_TIME_STEPPING_EXPLICIT_PREDICTOR_CORRECTOR) {
                } else if (ti == CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA) {
                    if (s->n_stage == 2) {
                    } else if (s->n_stage == 4 && i_stage == 4) {
                        cst a = B->dUdt[0][cR], b = B->dUdt[1][cR], c2 = B->dUdt[2][cR], r;
                        r.d = a.d + TWO * b.d + TWO * c2.d;
                        r.mx = a.mx + TWO * b.mx + TWO * c2.mx;
                        r.my = a.my + TWO * b.my + TWO * c2.my;
                        r.E = a.E + TWO * b.E + TWO * c2.E;
                        B->dUdt[k][cR] = r;
                    } else {
                        B->dUdt[k][cR] = Z;
                    }
                } else {
                    B->dUdt[k][cR] = Z;
                }
            }
            if (j > JCl - 1 && j < JCu + 1) {
                pst Wl, Wr;
                double nx, ny, xf, yf;
                if (i == ICl - 1 && is_face_bc(B->bcW[j])) {
                    xfaceW(B, i + 1, j, &xf, &yf);
                    Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
                    nfaceW(B, i + 1, j, &nx, &ny);
                    if (B->bcW[j] == CFFC_BC_REFLECTION) Wl = reflect(Wr, nx, ny);
                    else Wl = bc_characteristic_pressure(Wr, B->WoW[j], nx, ny);
                } else if (i == ICu && is_face_bc(B->bcE[j])) {
                    xfaceE(B, i, j, &xf, &yf);
                    Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                    nfaceE(B, i, j, &nx, &ny);
                    if (B->bcE[j] == CFFC_BC_REFLECTION) Wr = reflect(Wl, nx, ny);
                    else Wr = bc_characteristic_pressure(Wl, B->WoE[j], nx, ny);
                } else {
                    xfaceE(B, i, j, &xf, &yf);
                    Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                    xfaceW(B, i + 1, j, &xf, &yf);
                    Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
                }
                nfaceE(B, i, j, &nx, &ny);
                cst F = flux_n(s, Wl, Wr, nx, ny);
                acc_sub(&B->dUdt[k][cL], CFL * B->dt[cL], F, lfaceE(B, i, j), B->A[cL]);
                acc_add(&B->dUdt[k][cR], CFL * B->dt[cR], F, lfaceW(B, i + 1, j), B->A[cR]);
                /* axisymmetric source terms :1511-1516, after cell i's east
                   face (the ghost cell ICl-1 is cleared below anyway) */
                if (s->axisymmetric && i >= ICl) {
                    const double cdt = CFL * B->dt[cL];
                    const cst Sx = axisym_source(B->W[cL], B->yc[cL]);
                    cst *d = &B->dUdt[k][cL];
                    d->d += cdt * Sx.d; d->mx += cdt * Sx.mx; d->my += cdt * Sx.my; d->E += cdt * Sx.E;
                }
                /* west and east boundary fluxes :1518-1523 */
                if (i == ICl - 1) B->FluxW[j - JCl] = flux_times(F, -ONE, lfaceW(B, i + 1, j));
                else if (i == ICu) B->FluxE[j - JCl] = flux_times(F, ONE, lfaceE(B, i, j));
            }
        }
        if (j > JCl - 1 && j < JCu + 1) {
            B->dUdt[k][CI(B, ICl - 1, j)] = Z;
            B->dUdt[k][CI(B, ICu + 1, j)] = Z;
        }
    }
    /* j-direction fluxes :1537-1721 */
    for (int i = ICl; i <= ICu; ++i) {
        for (int j = JCl - 1; j <= JCu; ++j) {
            int cL = CI(B, i, j), cR = CI(B, i, j + 1);
            pst Wl, Wr;
            double nx, ny, xf, yf;
            if (j == JCl - 1 && is_face_bc(B->bcS[i])) {
                xfaceS(B, i, j + 1, &xf, &yf);
                Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
                nfaceS(B, i, j + 1, &nx, &ny);
                if (B->bcS[i] == CFFC_BC_REFLECTION) Wl = reflect(Wr, nx, ny);
                else Wl = bc_characteristic_pressure(Wr, B->WoS[i], nx, ny);
            } else if (j == JCu && is_face_bc(B->bcN[i])) {
                xfaceN(B, i, j, &xf, &yf);
                Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                nfaceN(B, i, j, &nx, &ny);
                if (B->bcN[i] == CFFC_BC_REFLECTION) Wr = reflect(Wl, nx, ny);
                else Wr = bc_characteristic_pressure(Wl, B->WoN[i], nx, ny);
            } else {
                xfaceN(B, i, j, &xf, &yf);
                Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                xfaceS(B, i, j + 1, &xf, &yf);
                Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
            }
            nfaceN(B, i, j, &nx, &ny);
            cst F = flux_n(s, Wl, Wr, nx, ny);
            acc_sub(&B->dUdt[k][cL], CFL * B->dt[cL], F, lfaceN(B, i, j), B->A[cL]);
            acc_add(&B->dUdt[k][cR], CFL * B->dt[cR], F, lfaceS(B, i, j + 1), B->A[cR]);
            /* south and north boundary fluxes :1709-1715 */
            if (j == JCl - 1) B->FluxS[i - ICl] = flux_times(F, -ONE, lfaceS(B, i, j + 1));
            else if (j == JCu) B->FluxN[i - ICl] = flux_times(F, ONE, lfaceN(B, i, j));
        }
        B->dUdt[k][CI(B, i, JCl - 1)] = Z;
        B->dUdt[k][CI(B, i, JCu + 1)] = Z;
    }
}

/* (Flux*l)/A of the residual evaluation (Euler2DQuad.cc:1001-1006) */
static inline void res_sub(cst *d, cst F, double l, double A) {
    d->d -= F.d * l / A; d->mx -= F.mx * l / A; d->my -= F.my * l / A; d->E -= F.E * l / A;
}
static inline void res_add(cst *d, cst F, double l, double A) {
    d->d += F.d * l / A; d->mx += F.mx * l / A; d->my += F.my * l / A; d->E += F.E * l / A;
}

/* Euler2D_Quad_Block::dUdt_Residual_Evaluation, Euler2DQuad.cc:814-1217:
   the steady residual R(U) = -sum(Flux*l)/A into dUdt[0] (no CFL*dt factor),
   with the faces, BC states and accumulation order of the multistage
   residual. */
static void block_residual(oblk *B, const cffc_scheme_desc *s) {
    int ICl = B->ICl, ICu = B->ICu, JCl = B->JCl, JCu = B->JCu, Ng = B->ng;
    const cst Z = {0, 0, 0, 0};
    cst *R = B->dUdt[0];
    for (int j = JCl - Ng + 1; j <= JCu + Ng - 1; ++j)
        for (int i = ICl - Ng + 1; i <= ICu + Ng - 1; ++i) lsq_reconstruct(B, i, j, s->i_limiter);
    for (int j = JCl - 1; j <= JCu + 1; ++j) {
        R[CI(B, ICl - 1, j)] = Z;
        for (int i = ICl - 1; i <= ICu; ++i) {
            int cL = CI(B, i, j), cR = CI(B, i + 1, j);
            R[cR] = Z;
            if (j > JCl - 1 && j < JCu + 1) {
                pst Wl, Wr;
                double nx, ny, xf, yf;
                if (i == ICl - 1 && is_face_bc(B->bcW[j])) {
                    xfaceW(B, i + 1, j, &xf, &yf);
                    Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
                    nfaceW(B, i + 1, j, &nx, &ny);
                    if (B->bcW[j] == CFFC_BC_REFLECTION) Wl = reflect(Wr, nx, ny);
                    else Wl = bc_characteristic_pressure(Wr, B->WoW[j], nx, ny);
                } else if (i == ICu && is_face_bc(B->bcE[j])) {
                    xfaceE(B, i, j, &xf, &yf);
                    Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                    nfaceE(B, i, j, &nx, &ny);
                    if (B->bcE[j] == CFFC_BC_REFLECTION) Wr = reflect(Wl, nx, ny);
                    else Wr = bc_characteristic_pressure(Wl, B->WoE[j], nx, ny);
                } else {
                    xfaceE(B, i, j, &xf, &yf);
                    Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                    xfaceW(B, i + 1, j, &xf, &yf);
                    Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
                }
                nfaceE(B, i, j, &nx, &ny);
                cst F = flux_n(s, Wl, Wr, nx, ny);
                res_sub(&R[cL], F, lfaceE(B, i, j), B->A[cL]);
                res_add(&R[cR], F, lfaceW(B, i + 1, j), B->A[cR]);
                /* axisymmetric source terms :1016-1020 */
                if (s->axisymmetric && i >= ICl) {
                    const cst Sx = axisym_source(B->W[cL], B->yc[cL]);
                    cst *d = &R[cL];
                    d->d += Sx.d; d->mx += Sx.mx; d->my += Sx.my; d->E += Sx.E;
                }
                if (i == ICl - 1) B->FluxW[j - JCl] = flux_times(F, -ONE, lfaceW(B, i + 1, j));
                else if (i == ICu) B->FluxE[j - JCl] = flux_times(F, ONE, lfaceE(B, i, j));
            }
        }
        if (j > JCl - 1 && j < JCu + 1) {
            R[CI(B, ICl - 1, j)] = Z;
            R[CI(B, ICu + 1, j)] = Z;
        }
    }
    for (int i = ICl; i <= ICu; ++i) {
        for (int j = JCl - 1; j <= JCu; ++j) {
            int cL = CI(B, i, j), cR = CI(B, i, j + 1);
            pst Wl, Wr;
            double nx, ny, xf, yf;
            if (j == JCl - 1 && is_face_bc(B->bcS[i])) {
                xfaceS(B, i, j + 1, &xf, &yf);
                Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
                nfaceS(B, i, j + 1, &nx, &ny);
                if (B->bcS[i] == CFFC_BC_REFLECTION) Wl = reflect(Wr, nx, ny);
                else Wl = bc_characteristic_pressure(Wr, B->WoS[i], nx, ny);
            } else if (j == JCu && is_face_bc(B->bcN[i])) {
                xfaceN(B, i, j, &xf, &yf);
                Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                nfaceN(B, i, j, &nx, &ny);
                if (B->bcN[i] == CFFC_BC_REFLECTION) Wr = reflect(Wl, nx, ny);
                else Wr = bc_characteristic_pressure(Wl, B->WoN[i], nx, ny);
            } else {
                xfaceN(B, i, j, &xf, &yf);
                Wl = face_state(B, cL, xf - B->xc[cL], yf - B->yc[cL]);
                xfaceS(B, i, j + 1, &xf, &yf);
                Wr = face_state(B, cR, xf - B->xc[cR], yf - B->yc[cR]);
            }
            nfaceN(B, i, j, &nx, &ny);
            cst F = flux_n(s, Wl, Wr, nx, ny);
            res_sub(&R[cL], F, lfaceN(B, i, j), B->A[cL]);
            res_add(&R[cR], F, lfaceS(B, i, j + 1), B->A[cR]);
            if (j == JCl - 1) B->FluxS[i - ICl] = flux_times(F, -ONE, lfaceS(B, i, j + 1));
            else if (j == JCu) B->FluxN[i - ICl] = flux_times(F, ONE, lfaceN(B, i, j));
        }
        R[CI(B, i, JCl - 1)] = Z;
        R[CI(B, i, JCu + 1)] = Z;
    }
}

/* ====================================================================== */
/* Resolution changes (quadtree AMR): Send_All_Messages' fine-to-coarse and */
/* coarse-to-fine legs (AdaptiveBlock2D_MessagePassing.h:1185, 2399, 6284,  */
/* 7140, in the order of :7883-8020) and the conservative flux correction   */
/* (Send_Conservative_Flux_Corrections :8328, Apply_Boundary_Flux_          */
/* Corrections_Multistage_Explicit Euler2DQuadSingleBlock.cc:4992-5080).    */
/* Neighbours share orientation and, across a resolution change, the cell   */
/* counts ni, nj (both even).  A face with two finer neighbours lists them  */
/* in nbr[f] (the west / south one, infoX[0]) and nbr2[f] (infoX[1]).       */
/* ====================================================================== */

/* sector of fine block b on the face of its coarser neighbour nbr[f]:
   0 = west/south half (infoX[0] of the coarse block), 1 = east/north half */
static int face_sector(const oblk *all, int b, int f) {
    const oblk *C = &all[all[b].nbr[f]];
    if (C->nbr[f ^ 1] == b) return 0;
    if (C->nbr2[f ^ 1] == b) return 1;
    return -1;
}
static inline int floor_half(int k) { return k >= 0 ? k / 2 : -((1 - k) / 2); }

static void set_ghost_U(oblk *B, int i, int j, cst U) {
    B->U[CI(B, i, j)] = U;
    B->W[CI(B, i, j)] = W_of_U(U);
}

/* LoadSendBuffer_F2C, Euler2DQuad.h:1597-1624: area-weighted average of the
   fine cells (i,j), (i+1,j), (i,j+1), (i+1,j+1); the coarse ghost cell then
   takes U and W = W(U) (UnloadReceiveBuffer_F2C :1658-1684) */
static cst restrict4(const oblk *F, int i, int j) {
    int c00 = CI(F, i, j), c10 = CI(F, i + 1, j), c01 = CI(F, i, j + 1), c11 = CI(F, i + 1, j + 1);
    double A00 = F->A[c00], A10 = F->A[c10], A01 = F->A[c01], A11 = F->A[c11];
    double den = A00 + A10 + A01 + A11;
    cst u00 = F->U[c00], u10 = F->U[c10], u01 = F->U[c01], u11 = F->U[c11], r;
    r.d = (A00 * u00.d + A10 * u10.d + A01 * u01.d + A11 * u11.d) / den;
    r.mx = (A00 * u00.mx + A10 * u10.mx + A01 * u01.mx + A11 * u11.mx) / den;
    r.my = (A00 * u00.my + A10 * u10.my + A01 * u01.my + A11 * u11.my) / den;
    r.E = (A00 * u00.E + A10 * u10.E + A01 * u01.E + A11 * u11.E) / den;
    return r;
}

/* fine-to-coarse leg for coarse block b: its ghost cells facing finer
   neighbours (faces: two blocks, each half the face; corners: one block)
   take the restriction of the fine cells they cover */
static void f2c_block(oblk *all, int b) {
    oblk *B = &all[b];
    int ICl = B->ICl, ICu = B->ICu, JCl = B->JCl, JCu = B->JCu;
    int hi = B->ni / 2, hj = B->nj / 2;
    for (int f = 0; f < 4; ++f) {
        int nb = B->nbr[f];
        if (nb < 0 || all[nb].level != B->level + 1) continue;
        for (int g = 1; g <= 2; ++g) {
            if (f == CFFC_NBR_N || f == CFFC_NBR_S) {
                for (int i = ICl; i <= ICu; ++i) {
                    int k = i - ICl;
                    const oblk *F = &all[k < hi ? B->nbr[f] : B->nbr2[f]];
                    int i_f = F->ICl + 2 * (k % hi);
                    int j_f = (f == CFFC_NBR_N) ? F->JCl + 2 * (g - 1) : F->JCu - 2 * g + 1;
                    set_ghost_U(B, i, (f == CFFC_NBR_N) ? JCu + g : JCl - g, restrict4(F, i_f, j_f));
                }
            } else {
                for (int j = JCl; j <= JCu; ++j) {
                    int k = j - JCl;
                    const oblk *F = &all[k < hj ? B->nbr[f] : B->nbr2[f]];
                    int j_f = F->JCl + 2 * (k % hj);
                    int i_f = (f == CFFC_NBR_E) ? F->ICl + 2 * (g - 1) : F->ICu - 2 * g + 1;
                    set_ghost_U(B, (f == CFFC_NBR_E) ? ICu + g : ICl - g, j, restrict4(F, i_f, j_f));
                }
            }
        }
    }
    for (int s = CFFC_NBR_NE; s <= CFFC_NBR_SW; ++s) {
        int nb = B->nbr[s];
        if (nb < 0 || all[nb].level != B->level + 1) continue;
        const oblk *F = &all[nb];
        int east = (s == CFFC_NBR_NE || s == CFFC_NBR_SE), north = (s == CFFC_NBR_NE || s == CFFC_NBR_NW);
        for (int gj = 1; gj <= 2; ++gj)
            for (int gi = 1; gi <= 2; ++gi) {
                int i_f = east ? F->ICl + 2 * (gi - 1) : F->ICu - 2 * gi + 1;
                int j_f = north ? F->JCl + 2 * (gj - 1) : F->JCu - 2 * gj + 1;
                set_ghost_U(B, east ? ICu + gi : ICl - gi, north ? JCu + gj : JCl - gj, restrict4(F, i_f, j_f));
            }
    }
}

/* quadCentroid (Math.h:311-350) of the vertices (x[q], y[q]), q = SW, SE, NE, NW */
static void quad_centroid_v(const double *x, const double *y, double *xc, double *yc) {
    double ai, atmp = 0, xtmp = 0, ytmp = 0;
    ai = x[3] * y[0] - x[0] * y[3]; atmp += ai; xtmp += (x[0] + x[3]) * ai; ytmp += (y[0] + y[3]) * ai;
    ai = x[0] * y[1] - x[1] * y[0]; atmp += ai; xtmp += (x[1] + x[0]) * ai; ytmp += (y[1] + y[0]) * ai;
    ai = x[1] * y[2] - x[2] * y[1]; atmp += ai; xtmp += (x[2] + x[1]) * ai; ytmp += (y[2] + y[1]) * ai;
    ai = x[2] * y[3] - x[3] * y[2]; atmp += ai; xtmp += (x[3] + x[2]) * ai; ytmp += (y[3] + y[2]) * ai;
    if (atmp != 0) { *xc = xtmp / (3 * atmp); *yc = ytmp / (3 * atmp); }
    else { *xc = 0; *yc = 0; }
}

/* Grid2D_Quad_Block_HO::centroidSW/SE/NW/NE (HO_Grid2DQuad.h:1486-1660):
   centroid of quarter q (0 SW, 1 SE, 2 NW, 3 NE) of cell (i, j), the
   quarter spanned by a corner node, the two adjacent edge midpoints and
   the average of the four nodes */
static void subcell_centroid(const oblk *C, int i, int j, int q, double *xc, double *yc) {
    double x00, y00, x10, y10, x01, y01, x11, y11;
    nodeX(C, i, j, &x00, &y00); nodeX(C, i + 1, j, &x10, &y10);
    nodeX(C, i, j + 1, &x01, &y01); nodeX(C, i + 1, j + 1, &x11, &y11);
    double cx = 0.25 * (x00 + x10 + x01 + x11), cy = 0.25 * (y00 + y10 + y01 + y11);
    double sx = HALF * (x00 + x10), sy = HALF * (y00 + y10);   /* south edge midpoint */
    double ex = HALF * (x10 + x11), ey = HALF * (y10 + y11);   /* east */
    double nx = HALF * (x01 + x11), ny = HALF * (y01 + y11);   /* north */
    double wx = HALF * (x00 + x01), wy = HALF * (y00 + y01);   /* west */
    double X[4], Y[4];
    switch (q) {
    case 0: X[0] = x00; Y[0] = y00; X[1] = sx; Y[1] = sy; X[2] = cx; Y[2] = cy; X[3] = wx; Y[3] = wy; break;
    case 1: X[0] = sx; Y[0] = sy; X[1] = x10; Y[1] = y10; X[2] = ex; Y[2] = ey; X[3] = cx; Y[3] = cy; break;
    case 2: X[0] = wx; Y[0] = wy; X[1] = cx; Y[1] = cy; X[2] = nx; Y[2] = ny; X[3] = x01; Y[3] = y01; break;
    default: X[0] = cx; Y[0] = cy; X[1] = ex; Y[1] = ey; X[2] = x11; Y[2] = y11; X[3] = nx; Y[3] = ny; break;
    }
    quad_centroid_v(X, Y, xc, yc);
}

/* SubcellReconstruction (Euler2DQuad.h:1721-1950) of an interior cell (i, j)
   with LIMITER_VENKATAKRISHNAN and the 8-point stencil, into (gx, gy, ph)
   rather than the block's arrays (the main reconstruction of the stage
   recomputes the same values there); frozen limiters keep the stored phi */
static void subcell_reconstruct(const oblk *C, int i, int j, pst *gx, pst *gy, pst *ph) {
    static const int di[8] = {-1, 0, 1, -1, 1, -1, 0, 1}, dj[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    int c = CI(C, i, j);
    pst W0 = C->W[c];
    pst DUDx = {0, 0, 0, 0}, DUDy = {0, 0, 0, 0};
    double DxDx = ZERO, DxDy = ZERO, DyDy = ZERO;
    for (int n2 = 0; n2 < 8; ++n2) {
        int cn = CI(C, i + di[n2], j + dj[n2]);
        double dXx = C->xc[cn] - C->xc[c], dXy = C->yc[cn] - C->yc[c];
        pst Wn = C->W[cn];
        pst DU = {Wn.d - W0.d, Wn.vx - W0.vx, Wn.vy - W0.vy, Wn.p - W0.p};
        DUDx.d += DU.d * dXx; DUDx.vx += DU.vx * dXx; DUDx.vy += DU.vy * dXx; DUDx.p += DU.p * dXx;
        DUDy.d += DU.d * dXy; DUDy.vx += DU.vx * dXy; DUDy.vy += DU.vy * dXy; DUDy.p += DU.p * dXy;
        DxDx += dXx * dXx; DxDy += dXx * dXy; DyDy += dXy * dXy;
    }
    double np = 8.0;
    DUDx.d /= np; DUDx.vx /= np; DUDx.vy /= np; DUDx.p /= np;
    DUDy.d /= np; DUDy.vx /= np; DUDy.vy /= np; DUDy.p /= np;
    DxDx /= np; DxDy /= np; DyDy /= np;
    double det = DxDx * DyDy - DxDy * DxDy;
    gx->d = (DUDx.d * DyDy - DUDy.d * DxDy) / det;
    gx->vx = (DUDx.vx * DyDy - DUDy.vx * DxDy) / det;
    gx->vy = (DUDx.vy * DyDy - DUDy.vy * DxDy) / det;
    gx->p = (DUDx.p * DyDy - DUDy.p * DxDy) / det;
    gy->d = (DUDy.d * DxDx - DUDx.d * DxDy) / det;
    gy->vx = (DUDy.vx * DxDx - DUDx.vx * DxDy) / det;
    gy->vy = (DUDy.vy * DxDx - DUDx.vy * DxDy) / det;
    gy->p = (DUDy.p * DxDx - DUDx.p * DxDy) / det;
    if (C->freeze_limiter) {
        *ph = C->phi[c];
        return;
    }
    double xE, yE, xW, yW, xN, yN, xS, yS;
    xfaceE(C, i, j, &xE, &yE); xfaceW(C, i, j, &xW, &yW);
    xfaceN(C, i, j, &xN, &yN); xfaceS(C, i, j, &xS, &yS);
    double xc0 = C->xc[c], yc0 = C->yc[c];
    for (int n = 1; n <= 4; ++n) {
        double u0 = pget(&W0, n), u0Min = u0, u0Max = u0;
        for (int n2 = 0; n2 < 8; ++n2) {
            double v = pget(&C->W[CI(C, i + di[n2], j + dj[n2])], n);
            u0Min = dmin(u0Min, v);
            u0Max = dmax(u0Max, v);
        }
        double gxn = pget(gx, n), gyn = pget(gy, n), uQuad[4];
        uQuad[0] = u0 + gxn * (xE - xc0) + gyn * (yE - yc0);
        uQuad[1] = u0 + gxn * (xW - xc0) + gyn * (yW - yc0);
        uQuad[2] = u0 + gxn * (xN - xc0) + gyn * (yN - yc0);
        uQuad[3] = u0 + gxn * (xS - xc0) + gyn * (yS - yc0);
        *pref(ph, n) = limiter_venkatakrishnan(uQuad, u0, u0Min, u0Max, 4);
    }
}

/* LoadSendBuffer_C2F (Euler2DQuad.cc:266-360): the limited linear state of
   coarse cell (i, j) at the centroid of its quarter q, sent as U; the fine
   ghost cell takes U and W = W(U) (UnloadReceiveBuffer_C2F Euler2DQuad.h:1691) */
static cst prolong(const oblk *C, int i, int j, int q, const pst *gx, const pst *gy, const pst *ph) {
    double xq, yq;
    subcell_centroid(C, i, j, q, &xq, &yq);
    int c = CI(C, i, j);
    double dXx = xq - C->xc[c], dXy = yq - C->yc[c];
    pst W = C->W[c], r;
    r.d = W.d + (ph->d * gx->d) * dXx + (ph->d * gy->d) * dXy;
    r.vx = W.vx + (ph->vx * gx->vx) * dXx + (ph->vx * gy->vx) * dXy;
    r.vy = W.vy + (ph->vy * gx->vy) * dXx + (ph->vy * gy->vy) * dXy;
    r.p = W.p + (ph->p * gx->p) * dXx + (ph->p * gy->p) * dXy;
    return U_of_W(r);
}

/* coarse-to-fine leg for fine block b.  A face with a coarser neighbour
   receives the quarters of ni/2 + 1 (nj/2 + 1) coarse cells: its own span
   plus the corner over the middle of the coarse face
   (Load_Send_Message_Buffers_ResChange_CoarseToFine, :2399-2470); other
   corners with a coarser corner neighbour receive the quarters of that
   block's facing corner cell. */
static void c2f_block(oblk *all, int b) {
    oblk *B = &all[b];
    int ICl = B->ICl, ICu = B->ICu, JCl = B->JCl, JCu = B->JCu;
    int covered[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    pst gx, gy, ph;
    for (int f = 0; f < 4; ++f) {
        int nb = B->nbr[f];
        if (nb < 0 || all[nb].level != B->level - 1) continue;
        const oblk *C = &all[nb];
        int sig = face_sector(all, b, f);
        if (f == CFFC_NBR_N || f == CFFC_NBR_S) {
            int i0 = sig == 0 ? ICl : ICl - 2, i1 = sig == 0 ? ICu + 2 : ICu;
            int jc = (f == CFFC_NBR_N) ? C->JCl : C->JCu;
            int ic_prev = -1000;
            for (int i = i0; i <= i1; ++i) {
                int k = i - ICl, kf = floor_half(k), qx = k - 2 * kf;
                int ic = C->ICl + sig * (B->ni / 2) + kf;
                if (ic != ic_prev) { subcell_reconstruct(C, ic, jc, &gx, &gy, &ph); ic_prev = ic; }
                for (int g = 1; g <= 2; ++g) {
                    int qy = (f == CFFC_NBR_N) ? g - 1 : 2 - g;
                    set_ghost_U(B, i, (f == CFFC_NBR_N) ? JCu + g : JCl - g, prolong(C, ic, jc, 2 * qy + qx, &gx, &gy, &ph));
                }
            }
            if (f == CFFC_NBR_N) covered[sig == 0 ? CFFC_NBR_NE : CFFC_NBR_NW] = 1;
            else covered[sig == 0 ? CFFC_NBR_SE : CFFC_NBR_SW] = 1;
        } else {
            int j0 = sig == 0 ? JCl : JCl - 2, j1 = sig == 0 ? JCu + 2 : JCu;
            int ic = (f == CFFC_NBR_E) ? C->ICl : C->ICu;
            int jc_prev = -1000;
            for (int j = j0; j <= j1; ++j) {
                int k = j - JCl, kf = floor_half(k), qy = k - 2 * kf;
                int jc = C->JCl + sig * (B->nj / 2) + kf;
                if (jc != jc_prev) { subcell_reconstruct(C, ic, jc, &gx, &gy, &ph); jc_prev = jc; }
                for (int g = 1; g <= 2; ++g) {
                    int qx = (f == CFFC_NBR_E) ? g - 1 : 2 - g;
                    set_ghost_U(B, (f == CFFC_NBR_E) ? ICu + g : ICl - g, j, prolong(C, ic, jc, 2 * qy + qx, &gx, &gy, &ph));
                }
            }
            if (f == CFFC_NBR_E) covered[sig == 0 ? CFFC_NBR_NE : CFFC_NBR_SE] = 1;
            else covered[sig == 0 ? CFFC_NBR_NW : CFFC_NBR_SW] = 1;
        }
    }
    for (int s = CFFC_NBR_NE; s <= CFFC_NBR_SW; ++s) {
        int nb = B->nbr[s];
        if (covered[s] || nb < 0 || all[nb].level != B->level - 1) continue;
        const oblk *D = &all[nb];
        int east = (s == CFFC_NBR_NE || s == CFFC_NBR_SE), north = (s == CFFC_NBR_NE || s == CFFC_NBR_NW);
        int ic = east ? D->ICl : D->ICu, jc = north ? D->JCl : D->JCu;
        int ig0 = east ? ICu + 1 : ICl - 2, jg0 = north ? JCu + 1 : JCl - 2;
        subcell_reconstruct(D, ic, jc, &gx, &gy, &ph);
        for (int qy = 0; qy < 2; ++qy)
            for (int qx = 0; qx < 2; ++qx) set_ghost_U(B, ig0 + qx, jg0 + qy, prolong(D, ic, jc, 2 * qy + qx, &gx, &gy, &ph));
    }
}

/* Send_Conservative_Flux_Corrections + Apply_Boundary_Flux_Corrections_
   Multistage_Explicit for coarse block b: on a face with two finer
   neighbours, the saved coarse flux becomes -(sum of the two fine fluxes
   it covers) - own (UnloadReceiveBuffer_Flux_F2C Euler2DQuad.h:2120-2170,
   fine pairs from LoadSendBuffer_Flux_F2C :2064-2112), and
   dUdt[k] -= (CFL*dt)*Flux/A on the face's cells, faces in the order N, S, E, W */
static void flux_correction_block(oblk *all, int b, const cffc_scheme_desc *s, int i_stage) {
    oblk *B = &all[b];
    int k = stage_k_residual(s, i_stage);
    double CFL = s->cfl_number;
    int hi = B->ni / 2, hj = B->nj / 2;
    for (int f = 0; f < 4; ++f) {
        int nb = B->nbr[f];
        if (nb < 0 || all[nb].level != B->level + 1) continue;
        int n = (f == CFFC_NBR_N || f == CFFC_NBR_S) ? B->ni : B->nj, h = (n == B->ni) ? hi : hj;
        for (int q = 0; q < n; ++q) {
            const oblk *F = &all[q < h ? B->nbr[f] : B->nbr2[f]];
            int qf = 2 * (q % h);
            const cst *fine = f == CFFC_NBR_N ? F->FluxS : f == CFFC_NBR_S ? F->FluxN : f == CFFC_NBR_E ? F->FluxW : F->FluxE;
            cst *own = f == CFFC_NBR_N ? B->FluxN : f == CFFC_NBR_S ? B->FluxS : f == CFFC_NBR_E ? B->FluxE : B->FluxW;
            cst a = fine[qf], c2 = fine[qf + 1], o = own[q], r;
            r.d = -(a.d + c2.d) - o.d; r.mx = -(a.mx + c2.mx) - o.mx;
            r.my = -(a.my + c2.my) - o.my; r.E = -(a.E + c2.E) - o.E;
            own[q] = r;
            int i = (f == CFFC_NBR_N || f == CFFC_NBR_S) ? B->ICl + q : (f == CFFC_NBR_E ? B->ICu : B->ICl);
            int j = (f == CFFC_NBR_E || f == CFFC_NBR_W) ? B->JCl + q : (f == CFFC_NBR_N ? B->JCu : B->JCl);
            int c = CI(B, i, j);
            double cdt = CFL * B->dt[c], A = B->A[c];
            cst *d = &B->dUdt[k][c];
            d->d -= cdt * r.d / A; d->mx -= cdt * r.mx / A; d->my -= cdt * r.my / A; d->E -= cdt * r.E / A;
        }
    }
}

/* Update_Solution_Multistage_Explicit, Euler2DQuadSingleBlock.cc:5300-5507
   (GLOBAL and SCALAR_LOCAL time stepping).  Returns 0 or the i-index of the
   first failing cell; *bad_j receives its j-index. */
static int block_update(oblk *B, const cffc_scheme_desc *s, int i_stage, int *bad_j) {
    int k = stage_k_residual(s, i_stage);
    double omega = stage_omega(s, i_stage);
    for (int j = B->JCl; j <= B->JCu; ++j)
        for (int i = B->ICl; i <= B->ICu; ++i) {
            int c = CI(B, i, j);
            cst *U = &B->U[c], Uo = B->Uo[c], *dU = &B->dUdt[k][c];
            U->d = Uo.d + omega * dU->d; U->mx = Uo.mx + omega * dU->mx;
            U->my = Uo.my + omega * dU->my; U->E = Uo.E + omega * dU->E;
            int bad = (U->d <= ZERO || U->E <= ZERO || U_e(*U) <= ZERO);
            if (s->local_time_stepping == CFFC_GLOBAL_TIME_STEPPING && bad) {
                *bad_j = j;
                return i;
            } else if (s->local_time_stepping == CFFC_SCALAR_LOCAL_TIME_STEPPING && bad) {
                double f = ONE;
                for (int n = 1; n <= 10; ++n) {
                    f = HALF * f;
                    B->dt[c] = f * B->dt[c];
                    dU->d = f * dU->d; dU->mx = f * dU->mx; dU->my = f * dU->my; dU->E = f * dU->E;
                    U->d = Uo.d + omega * dU->d; U->mx = Uo.mx + omega * dU->mx;
                    U->my = Uo.my + omega * dU->my; U->E = Uo.E + omega * dU->E;
                    if (U->d > ZERO && U->E > ZERO && U_e(*U) > ZERO) break;
                }
                if (U->d <= ZERO || U->E <= ZERO || U_e(*U) <= ZERO) {
                    *bad_j = j;
                    return i;
                }
            }
            B->W[c] = W_of_U(*U);
        }
    return 0;
}

/* CFL, Euler2DQuadSingleBlock.cc:3019-3080 (non-Weiss-Smith branch) */
static double block_cfl(oblk *B) {
    double dtMin = MILLION;
    int Ng = B->ng;
    for (int j = B->JCl - Ng; j <= B->JCu + Ng; ++j)
        for (int i = B->ICl - Ng; i <= B->ICu + Ng; ++i) {
            int c = CI(B, i, j);
            if (i < B->ICl || i > B->ICu || j < B->JCl || j > B->JCu) {
                B->dt[c] = ZERO;
            } else {
                double d_i = TWO * (B->A[c] / (lfaceE(B, i, j) + lfaceW(B, i, j)));
                double d_j = TWO * (B->A[c] / (lfaceN(B, i, j) + lfaceS(B, i, j)));
                double nEx, nEy, nWx, nWy, nNx, nNy, nSx, nSy;
                nfaceE(B, i, j, &nEx, &nEy); nfaceW(B, i, j, &nWx, &nWy);
                nfaceN(B, i, j, &nNx, &nNy); nfaceS(B, i, j, &nSx, &nSy);
                pst W = B->W[c];
                double v_i = HALF * (W.vx * (nEx - nWx) + W.vy * (nEy - nWy));
                double v_j = HALF * (W.vx * (nNx - nSx) + W.vy * (nNy - nSy));
                double a = W_a(W);
                B->dt[c] = dmin(d_i / (a + fabs(v_i)), d_j / (a + fabs(v_j)));
                dtMin = dmin(dtMin, B->dt[c]);
            }
        }
    for (int j = B->JCl - Ng; j <= B->JCu + Ng; ++j)
        for (int i = B->ICl - Ng; i <= B->ICu + Ng; ++i)
            if (i < B->ICl || i > B->ICu || j < B->JCl || j > B->JCu) B->dt[CI(B, i, j)] = dtMin;
    return dtMin;
}
/* Set_Global_TimeStep :3089-3100 */
static void block_set_global_dt(oblk *B, double dt) {
    for (int j = B->JCl; j <= B->JCu; ++j)
        for (int i = B->ICl; i <= B->ICu; ++i) B->dt[CI(B, i, j)] = dt;
}

/* ====================================================================== */
/* exported API (ctypes)                                                   */
/* ====================================================================== */
static void blk_free(oblk *B) {
    free(B->xc); free(B->yc); free(B->A);
    free(B->bcW); free(B->bcE); free(B->bcS); free(B->bcN);
    free(B->WoW); free(B->WoE); free(B->WoS); free(B->WoN);
    free(B->W); free(B->dWdx); free(B->dWdy); free(B->phi);
    free(B->U); free(B->Uo); free(B->dUdt[0]); free(B->dUdt[1]); free(B->dUdt[2]);
    free(B->dt);
    free(B->FluxN); free(B->FluxS); free(B->FluxE); free(B->FluxW);
    free(B->P); free(B->uo);
}

void orc_destroy(orc_ctx *ctx) {
    if (!ctx) return;
    for (int b = 0; b < ctx->nblk; ++b) blk_free(&ctx->b[b]);
    free(ctx->b);
    free(ctx->xnode);
    free(ctx->ynode);
    free(ctx);
}

int orc_create(const cffc_mesh_desc *m, const cffc_scheme_desc *s, orc_ctx **out) {
    if (!m || !s || !out || m->nblk <= 0 || m->nghost != 2) return CFFC_E_ARG;
    if (s->i_time_integration != CFFC_TIME_STEPPING_EXPLICIT_RUNGE_KUTTA &&
        s->i_time_integration != CFFC_TIME_STEPPING_EXPLICIT_EULER &&
        s->i_time_integration != CFFC_TIME_STEPPING_EXPLICIT_PREDICTOR_CORRECTOR &&
        s->i_time_integration != CFFC_TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING)
        return CFFC_E_ARG;
    orc_ctx *ctx = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    ctx->nblk = m->nblk;
    ctx->s = *s;
    ctx->b = (oblk *)calloc(m->nblk, sizeof(oblk));
    int64_t nn = 0, nbc = 0;
    for (int b = 0; b < m->nblk; ++b) {
        nn += (int64_t)(m->ni[b] + 5) * (m->nj[b] + 5);
        nbc += 2 * (m->nj[b] + 4) + 2 * (m->ni[b] + 4);
    }
    ctx->xnode = (double *)malloc(nn * sizeof(double));
    ctx->ynode = (double *)malloc(nn * sizeof(double));
    memcpy(ctx->xnode, m->xnode, nn * sizeof(double));
    memcpy(ctx->ynode, m->ynode, nn * sizeof(double));
    int64_t noff = 0, boff = 0;
    ctx->ncell = 0;
    for (int b = 0; b < m->nblk; ++b) {
        oblk *B = &ctx->b[b];
        B->ni = m->ni[b]; B->nj = m->nj[b]; B->ng = 2;
        B->nci = B->ni + 4; B->ncj = B->nj + 4;
        B->ICl = 2; B->ICu = B->ni + 1; B->JCl = 2; B->JCu = B->nj + 1;
        B->xn = ctx->xnode + noff; B->yn = ctx->ynode + noff;
        noff += (int64_t)(B->nci + 1) * (B->ncj + 1);
        int nc = B->nci * B->ncj;
        B->xc = (double *)malloc(nc * sizeof(double));
        B->yc = (double *)malloc(nc * sizeof(double));
        B->A = (double *)malloc(nc * sizeof(double));
        for (int j = 0; j < B->ncj; ++j)
            for (int i = 0; i < B->nci; ++i)
                quad_centroid(B, i, j, &B->xc[CI(B, i, j)], &B->yc[CI(B, i, j)], &B->A[CI(B, i, j)]);
        B->bcW = (int *)malloc(B->ncj * sizeof(int)); B->bcE = (int *)malloc(B->ncj * sizeof(int));
        B->bcS = (int *)malloc(B->nci * sizeof(int)); B->bcN = (int *)malloc(B->nci * sizeof(int));
        B->WoW = (pst *)malloc(B->ncj * sizeof(pst)); B->WoE = (pst *)malloc(B->ncj * sizeof(pst));
        B->WoS = (pst *)malloc(B->nci * sizeof(pst)); B->WoN = (pst *)malloc(B->nci * sizeof(pst));
        int *dst[4] = {B->bcW, B->bcE, B->bcS, B->bcN};
        pst *wdst[4] = {B->WoW, B->WoE, B->WoS, B->WoN};
        int len[4] = {B->ncj, B->ncj, B->nci, B->nci};
        for (int f = 0; f < 4; ++f) {
            for (int q = 0; q < len[f]; ++q) {
                dst[f][q] = m->bc[boff + q];
                if (!bc_supported(dst[f][q])) { orc_destroy(ctx); return CFFC_E_ARG; }
                memcpy(&wdst[f][q], &m->wo[4 * (boff + q)], sizeof(pst));
            }
            boff += len[f];
        }
        for (int q = 0; q < 8; ++q) B->nbr[q] = m->nbr[8 * b + q];
        B->level = m->level ? m->level[b] : 0;
        for (int q = 0; q < 4; ++q) B->nbr2[q] = m->nbr2 ? m->nbr2[4 * b + q] : -1;
        B->FluxN = (cst *)calloc(B->ni, sizeof(cst)); B->FluxS = (cst *)calloc(B->ni, sizeof(cst));
        B->FluxE = (cst *)calloc(B->nj, sizeof(cst)); B->FluxW = (cst *)calloc(B->nj, sizeof(cst));
        B->W = (pst *)calloc(nc, sizeof(pst)); B->dWdx = (pst *)calloc(nc, sizeof(pst));
        B->dWdy = (pst *)calloc(nc, sizeof(pst)); B->phi = (pst *)calloc(nc, sizeof(pst));
        B->U = (cst *)calloc(nc, sizeof(cst)); B->Uo = (cst *)calloc(nc, sizeof(cst));
        for (int q = 0; q < 3; ++q) B->dUdt[q] = (cst *)calloc(nc, sizeof(cst));
        B->dt = (double *)calloc(nc, sizeof(double));
        ctx->ncell += (int64_t)B->ni * B->nj;
        if (B->level > ctx->max_level) ctx->max_level = B->level;
    }
    for (int b = 0; b < m->nblk; ++b)
        for (int q = 0; q < 8; ++q) {
            int nb = ctx->b[b].nbr[q];
            if (nb >= 0 && ctx->b[nb].level != ctx->b[b].level) ctx->amr = 1;
        }
    *out = ctx;
    return 0;
}

int64_t orc_num_cells(const orc_ctx *ctx) { return ctx->ncell; }

/* interior cells, blocks in order, i fastest, 4 doubles per cell */
int orc_set_W(orc_ctx *ctx, const double *W) {
    int64_t q = 0;
    for (int b = 0; b < ctx->nblk; ++b) {
        oblk *B = &ctx->b[b];
        for (int j = B->JCl; j <= B->JCu; ++j)
            for (int i = B->ICl; i <= B->ICu; ++i, ++q) {
                pst w; memcpy(&w, W + 4 * q, sizeof(pst));
                B->W[CI(B, i, j)] = w;
                B->U[CI(B, i, j)] = U_of_W(w);
            }
    }
    return 0;
}
int orc_set_U(orc_ctx *ctx, const double *U) {
    int64_t q = 0;
    for (int b = 0; b < ctx->nblk; ++b) {
        oblk *B = &ctx->b[b];
        for (int j = B->JCl; j <= B->JCu; ++j)
            for (int i = B->ICl; i <= B->ICu; ++i, ++q) {
                cst u; memcpy(&u, U + 4 * q, sizeof(cst));
                B->U[CI(B, i, j)] = u;
                B->W[CI(B, i, j)] = W_of_U(u);
            }
    }
    return 0;
}
static int get_field(orc_ctx *ctx, int which, int k, double *out) {
    int64_t q = 0;
    for (int b = 0; b < ctx->nblk; ++b) {
        oblk *B = &ctx->b[b];
        for (int j = B->JCl; j <= B->JCu; ++j)
            for (int i = B->ICl; i <= B->ICu; ++i, ++q) {
                int c = CI(B, i, j);
                const void *src;
                switch (which) {
                case 0: src = &B->W[c]; break;
                case 1: src = &B->U[c]; break;
                case 2: src = &B->dUdt[k][c]; break;
                case 3: src = &B->dWdx[c]; break;
                case 4: src = &B->dWdy[c]; break;
                default: src = &B->phi[c]; break;
                }
                memcpy(out + 4 * q, src, 4 * sizeof(double));
            }
    }
    return 0;
}
int orc_get_W(orc_ctx *ctx, double *W) { return get_field(ctx, 0, 0, W); }
int orc_get_U(orc_ctx *ctx, double *U) { return get_field(ctx, 1, 0, U); }
int orc_get_dUdt(orc_ctx *ctx, int k, double *d) { return get_field(ctx, 2, k, d); }
int orc_get_recon(orc_ctx *ctx, int which, double *d) { return get_field(ctx, 3 + which, 0, d); }
int orc_get_dt(orc_ctx *ctx, double *dt) {
    int64_t q = 0;
    for (int b = 0; b < ctx->nblk; ++b) {
        oblk *B = &ctx->b[b];
        for (int j = B->JCl; j <= B->JCu; ++j)
            for (int i = B->ICl; i <= B->ICu; ++i, ++q) dt[q] = B->dt[CI(B, i, j)];
    }
    return 0;
}
/* the whole padded W array of one block (ghosts included), i fastest */
int orc_get_block_W_padded(orc_ctx *ctx, int b, double *W) {
    oblk *B = &ctx->b[b];
    memcpy(W, B->W, sizeof(pst) * B->nci * B->ncj);
    return 0;
}
/* Test access to one block's reconstruction: set its padded W (ghosts
   included, no BCs or messages), run Linear_Reconstruction_LeastSquares over
   the cells block_dudt reconstructs (SingleBlock.cc:4467-4482), and read
   the padded dWdx / dWdy / phi back (which = 0, 1, 2). */
int orc_set_block_W_padded(orc_ctx *ctx, int b, const double *W) {
    oblk *B = &ctx->b[b];
    memcpy(B->W, W, sizeof(pst) * B->nci * B->ncj);
    return 0;
}
int orc_reconstruct_block(orc_ctx *ctx, int b) {
    oblk *B = &ctx->b[b];
    for (int j = B->JCl - B->ng + 1; j <= B->JCu + B->ng - 1; ++j)
        for (int i = B->ICl - B->ng + 1; i <= B->ICu + B->ng - 1; ++i) lsq_reconstruct(B, i, j, ctx->s.i_limiter);
    return 0;
}
int orc_get_block_recon_padded(orc_ctx *ctx, int b, int which, double *out) {
    oblk *B = &ctx->b[b];
    const pst *src = which == 0 ? B->dWdx : which == 1 ? B->dWdy : B->phi;
    memcpy(out, src, sizeof(pst) * B->nci * B->ncj);
    return 0;
}
/* cell geometry of one block (padded): xc, yc, A */
int orc_get_block_geometry(orc_ctx *ctx, int b, double *xc, double *yc, double *A) {
    oblk *B = &ctx->b[b];
    int nc = B->nci * B->ncj;
    memcpy(xc, B->xc, nc * sizeof(double)); memcpy(yc, B->yc, nc * sizeof(double));
    memcpy(A, B->A, nc * sizeof(double));
    return 0;
}

/* ---- multi-block wrappers (Euler2DQuadMultiBlock.cc) -------------------- */
typedef struct {
    orc_ctx *ctx;
    int b0, b1;
} range_t;

int orc_exchange_and_bcs(orc_ctx *ctx) {
    for (int b = 0; b < ctx->nblk; ++b) exchange_block(ctx->b, b);
    for (int b = 0; b < ctx->nblk; ++b) f2c_block(ctx->b, b);
    for (int b = 0; b < ctx->nblk; ++b) c2f_block(ctx->b, b);
    for (int b = 0; b < ctx->nblk; ++b) block_bcs(&ctx->b[b]);
    return 0;
}
/* Ghost cells of a new interior state on a mesh with resolution changes.
   The coarse-to-fine leg reads coarse ghosts that the physical BCs fill
   after it, whose corner rules in turn read ghosts from that leg one level
   coarser; max_level + 2 rounds of messages + BCs leave every ghost a
   function of the interior state alone (on one level, one round). */
int orc_prime_ghosts(orc_ctx *ctx) {
    int rounds = ctx->amr ? ctx->max_level + 2 : 1;
    for (int r = 0; r < rounds; ++r) orc_exchange_and_bcs(ctx);
    return rounds;
}
/* == Send_Conservative_Flux_Corrections + Apply_Boundary_Flux_Corrections_
      Multistage_Explicit (Euler2DQuadMultiBlock.cc:1307) */
int orc_apply_flux_corrections(orc_ctx *ctx, int i_stage) {
    for (int b = 0; b < ctx->nblk; ++b) flux_correction_block(ctx->b, b, &ctx->s, i_stage);
    return 0;
}
/* the saved boundary fluxes FluxN/S/E/W of one block (4 doubles each, in
   that order: ni, ni, nj, nj entries) */
int orc_get_block_fluxes(orc_ctx *ctx, int b, double *out) {
    oblk *B = &ctx->b[b];
    memcpy(out, B->FluxN, sizeof(cst) * B->ni); out += 4 * B->ni;
    memcpy(out, B->FluxS, sizeof(cst) * B->ni); out += 4 * B->ni;
    memcpy(out, B->FluxE, sizeof(cst) * B->nj); out += 4 * B->nj;
    memcpy(out, B->FluxW, sizeof(cst) * B->nj);
    return 0;
}
int orc_cfl(orc_ctx *ctx, double *dt_min, int set_global) {
    double dtMin = MILLION;
    for (int b = 0; b < ctx->nblk; ++b) dtMin = dmin(dtMin, block_cfl(&ctx->b[b]));
    if (set_global)
        for (int b = 0; b < ctx->nblk; ++b) block_set_global_dt(&ctx->b[b], dtMin);
    *dt_min = dtMin;
    return 0;
}
int orc_dudt_stage(orc_ctx *ctx, int i_stage) {
    for (int b = 0; b < ctx->nblk; ++b) block_dudt(&ctx->b[b], &ctx->s, i_stage);
    return 0;
}
int orc_update_stage(orc_ctx *ctx, int i_stage, int64_t *bad_cell) {
    for (int b = 0; b < ctx->nblk; ++b) {
        int bj = 0;
        int e = block_update(&ctx->b[b], &ctx->s, i_stage, &bj);
        if (e) {
            if (bad_cell) {
                int bi = e;
                *bad_cell = ((int64_t)b << 32) | ((int64_t)bj << 16) | bi;
            }
            return e;
        }
    }
    return 0;
}
/* Residual_Smoothing, Euler2DQuadSingleBlock.cc:4521-4564: implicit
   smoothing of dUdt[k] on the cells one away from the block edge, by
   lexicographic Gauss-Seidel sweeps (j outer, i inner) on dUdt[k+1], whose
   values on the edge ring are whatever that array holds (zero unless a
   scheme writes it).  k follows Residual_Smoothing, MultiBlock.cc:1227-1257. */
static void block_residual_smoothing(oblk *B, int k, double epsilon, int iters) {
    for (int j = B->JCl + 1; j <= B->JCu - 1; ++j)
        for (int i = B->ICl + 1; i <= B->ICu - 1; ++i) B->dUdt[k + 1][CI(B, i, j)] = B->dUdt[k][CI(B, i, j)];
    for (int n = 1; n <= iters; ++n)
        for (int j = B->JCl + 1; j <= B->JCu - 1; ++j)
            for (int i = B->ICl + 1; i <= B->ICu - 1; ++i) {
                const cst r = B->dUdt[k][CI(B, i, j)];
                const cst a = B->dUdt[k + 1][CI(B, i, j - 1)], b = B->dUdt[k + 1][CI(B, i - 1, j)];
                const cst c = B->dUdt[k + 1][CI(B, i + 1, j)], d = B->dUdt[k + 1][CI(B, i, j + 1)];
                const double den = ONE + FOUR * epsilon;
                cst o;
                o.d = (r.d + epsilon * (a.d + b.d + c.d + d.d)) / den;
                o.mx = (r.mx + epsilon * (a.mx + b.mx + c.mx + d.mx)) / den;
                o.my = (r.my + epsilon * (a.my + b.my + c.my + d.my)) / den;
                o.E = (r.E + epsilon * (a.E + b.E + c.E + d.E)) / den;
                B->dUdt[k + 1][CI(B, i, j)] = o;
            }
    for (int j = B->JCl + 1; j <= B->JCu - 1; ++j)
        for (int i = B->ICl + 1; i <= B->ICu - 1; ++i) B->dUdt[k][CI(B, i, j)] = B->dUdt[k + 1][CI(B, i, j)];
}
int orc_set_residual_smoothing(orc_ctx *ctx, double epsilon, int iterations) {
    ctx->rs_epsilon = epsilon;
    ctx->rs_iterations = iterations;
    return 0;
}
int orc_residual_smoothing(orc_ctx *ctx, int i_stage) {
    const int k = stage_k_residual(&ctx->s, i_stage);
    for (int b = 0; b < ctx->nblk; ++b)
        block_residual_smoothing(&ctx->b[b], k, ctx->rs_epsilon, ctx->rs_iterations);
    return 0;
}

/* L1/L2/Max_Norm_Residual :1017-1104 */
int orc_residual_norms(orc_ctx *ctx, int var, double out[3]) {
    double l1 = ZERO, l2 = ZERO, mx = ZERO;
    for (int b = 0; b < ctx->nblk; ++b) {
        oblk *B = &ctx->b[b];
        double bl1 = ZERO, bl2 = ZERO, bmx = ZERO;
        for (int j = B->JCl; j <= B->JCu; ++j)
            for (int i = B->ICl; i <= B->ICu; ++i) {
                double v = ((double *)&B->dUdt[0][CI(B, i, j)])[var - 1];
                bl1 += fabs(v);
                bl2 += sqr(v);
                bmx = dmax(bmx, fabs(v));
            }
        l1 += bl1;
        l2 += sqr(sqrt(bl2));
        mx = dmax(mx, bmx);
    }
    out[0] = l1;
    out[1] = sqrt(l2);
    out[2] = mx;
    return 0;
}
int orc_freeze_limiters(orc_ctx *ctx, int freeze) {
    for (int b = 0; b < ctx->nblk; ++b) ctx->b[b].freeze_limiter = freeze;
    return 0;
}

/* ---- the explicit time loop, Euler2DQuadSolvers.cc:549-898 --------------
   Steady (Time_Accurate = 0) loop over nsteps steps.  Blocks are split over
   nthreads threads, one contiguous block range per thread, mirroring the
   reference's one-MPI-rank-per-core block decomposition; the reductions
   (dt minimum, norms, error OR) are combined after a barrier. */
typedef struct {
    orc_ctx *ctx;
    int nthreads, tid, nsteps;
    double tol;
    pthread_barrier_t *bar;
    double *dtmins, *l1s, *l2s, *mxs;
    int *errs;
    double *norms, *dt_out;
    int *steps_done;
    int64_t *bad_cell;
} worker_t;

static void *worker(void *arg) {
    worker_t *w = (worker_t *)arg;
    orc_ctx *ctx = w->ctx;
    int nb = ctx->nblk, T = w->nthreads;
    int b0 = (int)((int64_t)nb * w->tid / T), b1 = (int)((int64_t)nb * (w->tid + 1) / T);
    const cffc_scheme_desc *s = &ctx->s;
    int n;
    for (n = 0; n <= w->nsteps; ++n) {
        /* CFL + CFFC_Minimum_MPI + Set_Global_TimeStep :657-684 */
        double dtm = MILLION;
        for (int b = b0; b < b1; ++b) dtm = dmin(dtm, block_cfl(&ctx->b[b]));
        w->dtmins[w->tid] = dtm;
        pthread_barrier_wait(w->bar);
        double dTime = MILLION;
        for (int t = 0; t < T; ++t) dTime = dmin(dTime, w->dtmins[t]);
        if (s->local_time_stepping == CFFC_GLOBAL_TIME_STEPPING)
            for (int b = b0; b < b1; ++b) block_set_global_dt(&ctx->b[b], dTime);
        /* norms of the previous step's dUdt[0] :687-696 */
        if (n > 0) {
            double l1 = 0, l2 = 0, mx = 0;
            for (int b = b0; b < b1; ++b) {
                oblk *B = &ctx->b[b];
                double bl1 = 0, bl2 = 0, bmx = 0;
                for (int j = B->JCl; j <= B->JCu; ++j)
                    for (int i = B->ICl; i <= B->ICu; ++i) {
                        double v = B->dUdt[0][CI(B, i, j)].d;
                        bl1 += fabs(v); bl2 += sqr(v); bmx = dmax(bmx, fabs(v));
                    }
                l1 += bl1; l2 += sqr(sqrt(bl2)); mx = dmax(mx, bmx);
            }
            w->l1s[w->tid] = l1; w->l2s[w->tid] = l2; w->mxs[w->tid] = mx;
        }
        pthread_barrier_wait(w->bar);
        if (n > 0) {
            double l1 = 0, l2 = 0, mx = 0;
            for (int t = 0; t < T; ++t) { l1 += w->l1s[t]; l2 += w->l2s[t]; mx = dmax(mx, w->mxs[t]); }
            l2 = sqrt(l2);
            if (w->tid == 0 && w->norms) {
                w->norms[3 * (n - 1) + 0] = l1; w->norms[3 * (n - 1) + 1] = l2; w->norms[3 * (n - 1) + 2] = mx;
            }
            if (w->tol > 0 && l2 <= w->tol) break;
        }
        if (n == w->nsteps) break;
        if (w->tid == 0 && w->dt_out) w->dt_out[n] = dTime;
        for (int st = 1; st <= s->n_stage; ++st) {
            /* same-level and fine-to-coarse legs read interior cells only */
            for (int b = b0; b < b1; ++b) { exchange_block(ctx->b, b); f2c_block(ctx->b, b); }
            pthread_barrier_wait(w->bar);
            if (ctx->amr) {
                /* coarse-to-fine leg: reads the coarse blocks' cells and ghosts
                   (those of the physical BCs from the previous BCs call) */
                for (int b = b0; b < b1; ++b) c2f_block(ctx->b, b);
                pthread_barrier_wait(w->bar);
            }
            for (int b = b0; b < b1; ++b) block_bcs(&ctx->b[b]);
            for (int b = b0; b < b1; ++b) block_dudt(&ctx->b[b], s, st);
            if (ctx->amr) {
                /* every block's saved fluxes precede the corrections */
                pthread_barrier_wait(w->bar);
                for (int b = b0; b < b1; ++b) flux_correction_block(ctx->b, b, s, st);
            }
            /* implicit residual smoothing, Euler2DQuadSolvers.cc:861-867 */
            if (ctx->rs_iterations > 0)
                for (int b = b0; b < b1; ++b)
                    block_residual_smoothing(&ctx->b[b], stage_k_residual(s, st), ctx->rs_epsilon, ctx->rs_iterations);
            int err = 0;
            for (int b = b0; b < b1 && !err; ++b) {
                int bj = 0;
                err = block_update(&ctx->b[b], s, st, &bj);
                if (err && w->bad_cell) *w->bad_cell = ((int64_t)b << 32) | ((int64_t)bj << 16) | err;
            }
            w->errs[w->tid] = err;
            pthread_barrier_wait(w->bar);
            int any = 0;
            for (int t = 0; t < T; ++t) if (w->errs[t]) { any = w->errs[t]; break; }
            if (any) { if (w->tid == 0) *w->steps_done = -any; return NULL; }
        }
    }
    if (w->tid == 0) *w->steps_done = n;
    return NULL;
}

/* Returns 0, or the reference's error code (> 0); *steps_done = steps run. */
int orc_run(orc_ctx *ctx, int nsteps, double tol, int nthreads, double *norms, double *dt_out, int *steps_done) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ctx->nblk) nthreads = ctx->nblk;
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, nthreads);
    double *dtm = calloc(nthreads, sizeof(double)), *l1 = calloc(nthreads, sizeof(double));
    double *l2 = calloc(nthreads, sizeof(double)), *mx = calloc(nthreads, sizeof(double));
    int *errs = calloc(nthreads, sizeof(int));
    worker_t *ws = calloc(nthreads, sizeof(worker_t));
    pthread_t *th = calloc(nthreads, sizeof(pthread_t));
    int done = 0;
    int64_t bad = 0;
    if (ctx->amr) orc_prime_ghosts(ctx);
    for (int t = 0; t < nthreads; ++t) {
        worker_t x = {ctx, nthreads, t, nsteps, tol, &bar, dtm, l1, l2, mx, errs, norms, dt_out, &done, &bad};
        ws[t] = x;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &ws[t]);
    worker(&ws[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&bar);
    free(dtm); free(l1); free(l2); free(mx); free(errs); free(ws); free(th);
    if (done < 0) { if (steps_done) *steps_done = 0; return -done; }
    if (steps_done) *steps_done = done;
    return 0;
}

/* ---- scalar entry points for known-answer tests ------------------------ */
void orc_flux_roe_n(const double *Wl, const double *Wr, double nx, double ny, double *F) {
    pst l, r; memcpy(&l, Wl, sizeof l); memcpy(&r, Wr, sizeof r);
    cst f = flux_roe_n(l, r, nx, ny);
    memcpy(F, &f, sizeof f);
}
void orc_flux_roe_x(const double *Wl, const double *Wr, double *F) {
    pst l, r; memcpy(&l, Wl, sizeof l); memcpy(&r, Wr, sizeof r);
    cst f = flux_roe_x(l, r);
    memcpy(F, &f, sizeof f);
}
void orc_roe_average(const double *Wl, const double *Wr, double *Wa) {
    pst l, r; memcpy(&l, Wl, sizeof l); memcpy(&r, Wr, sizeof r);
    pst a = roe_average(l, r);
    memcpy(Wa, &a, sizeof a);
}
double orc_limiter_venkatakrishnan(const double *uQuad, double u0, double u0Min, double u0Max) {
    return limiter_venkatakrishnan(uQuad, u0, u0Min, u0Max, 4);
}
double orc_harten_fix_pos(double la, double ll, double lr) { return harten_fix_pos(la, ll, lr); }
double orc_harten_fix_neg(double la, double ll, double lr) { return harten_fix_neg(la, ll, lr); }
void orc_reflect(const double *W, double nx, double ny, double *out) {
    pst w; memcpy(&w, W, sizeof w);
    pst r = reflect(w, nx, ny);
    memcpy(out, &r, sizeof r);
}
void orc_bc_characteristic_pressure(const double *Wi, const double *Wo, double nx, double ny, double *out) {
    pst a, b; memcpy(&a, Wi, sizeof a); memcpy(&b, Wo, sizeof b);
    pst r = bc_characteristic_pressure(a, b, nx, ny);
    memcpy(out, &r, sizeof r);
}
void orc_W_of_U(const double *U, double *W) {
    cst u; memcpy(&u, U, sizeof u);
    pst w = W_of_U(u);
    memcpy(W, &w, sizeof w);
}
void orc_U_of_W(const double *W, double *U) {
    pst w; memcpy(&w, W, sizeof w);
    cst u = U_of_W(w);
    memcpy(U, &u, sizeof u);
}
void orc_Fx(const double *W, double *F) {
    pst w; memcpy(&w, W, sizeof w);
    cst f = W_Fx(w);
    memcpy(F, &f, sizeof f);
}
double orc_runge_kutta(int i_stage, int n_stage) { return runge_kutta(i_stage, n_stage); }
double orc_multistage_optimal_smoothing(int i_stage, int n_stage, int limiter) {
    return multistage_optimal_smoothing(i_stage, n_stage, limiter);
}
void orc_quad_centroid(const double *x4, const double *y4, double *out3) {
    /* vertices SW, SE, NE, NW */
    double xs[4] = {x4[0], x4[1], x4[3], x4[2]}, ys[4] = {y4[0], y4[1], y4[3], y4[2]};
    oblk B;
    memset(&B, 0, sizeof B);
    B.nci = 1;
    B.xn = xs; B.yn = ys; /* node (i,j) at j*(nci+1)+i: (0,0)=SW (1,0)=SE (0,1)=NW (1,1)=NE */
    quad_centroid(&B, 0, 0, &out3[0], &out3[1], &out3[2]);
}

/* ====================================================================== */
/* FAS multigrid, FASMultigrid2D/FASMultigrid2D.h (regular V and W cycles) */
/* ====================================================================== */
/* One orc_ctx per level (level 0 the finest), each on its own mesh: level
   l+1 halves the resolution of level l block by block (Half_Mesh_Resolution,
   HO_Grid2DQuad.cc:12614), same block order.  Level schemes: the smoother's
   on level 0; below it HLLE and the zero limiter when first_order_coarse is
   set (dUdt_Multistage_Explicit_for_Multigrid :2491-2540 switch them for
   Level > Top_Level), which also selects the first-order optimal-smoothing
   coefficients.  Every level steps with its own per-cell dt (Smooth's CFL
   call, :2877; the global minimum is not applied there). */
typedef struct orc_mg orc_mg;
void orc_mg_destroy(orc_mg *mg);
struct orc_mg {
    int L;
    orc_ctx **lv;
    cffc_mg_desc d;
    int rs_iters;
    double rs_eps;
    cffc_scheme_desc fine_s, coarse_s;   /* top-level and below-top schemes */
};

/* the scheme of every level for a top level: the smoother's on the top
   level, HLLE and the zero limiter below it when first_order_coarse is set
   (dUdt_Multistage_Explicit_for_Multigrid :2502-2509, Level > Top_Level) */
static void mg_set_top(orc_mg *mg, int top) {
    for (int l = 0; l < mg->L; ++l) mg->lv[l]->s = (l > top) ? mg->coarse_s : mg->fine_s;
}

#define FINE_IDX(Ng, ic) (2 * ((ic) - (Ng)) + (Ng))

/* Restrict_Solution_Blocks :996-1037 (coarse interior cells: coarse-mesh BCs
   are applied, so nghost = 0) and Restrict_Residuals :1053-1095 into P */
static void mg_restrict(orc_ctx *F, orc_ctx *C, int residual) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *f = &F->b[b], *c = &C->b[b];
        const int Ng = f->ng;
        const cst *src = residual ? f->dUdt[0] : f->U;
        cst *dst = residual ? c->P : c->U;
        for (int ic = c->ICl; ic <= c->ICu; ++ic)
            for (int jc = c->JCl; jc <= c->JCu; ++jc) {
                const int i0 = FINE_IDX(Ng, ic), j0 = FINE_IDX(Ng, jc);
                const int q0 = CI(f, i0, j0), q1 = CI(f, i0 + 1, j0), q2 = CI(f, i0, j0 + 1), q3 = CI(f, i0 + 1, j0 + 1);
                const cst a = src[q0], bb = src[q1], cc = src[q2], dd = src[q3];
                const double A0 = f->A[q0], A1 = f->A[q1], A2 = f->A[q2], A3 = f->A[q3], Ac = c->A[CI(c, ic, jc)];
                cst r;
                r.d = (a.d * A0 + bb.d * A1 + cc.d * A2 + dd.d * A3) / Ac;
                r.mx = (a.mx * A0 + bb.mx * A1 + cc.mx * A2 + dd.mx * A3) / Ac;
                r.my = (a.my * A0 + bb.my * A1 + cc.my * A2 + dd.my * A3) / Ac;
                r.E = (a.E * A0 + bb.E * A1 + cc.E * A2 + dd.E * A3) / Ac;
                dst[CI(c, ic, jc)] = r;
            }
    }
}
/* Restrict_Boundary_Ref_States :1112-1186: the coarse WoN/S/E/W of the
   interior span from the two fine boundary cells' states, area weighted */
static void mg_restrict_wo(orc_ctx *F, orc_ctx *C) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *f = &F->b[b], *c = &C->b[b];
        const int Ng = f->ng;
#define WAVG(Wa, Wb, Aa, Ab) { ((Wa).d * (Aa) + (Wb).d * (Ab)) / ((Aa) + (Ab)), ((Wa).vx * (Aa) + (Wb).vx * (Ab)) / ((Aa) + (Ab)), \
                               ((Wa).vy * (Aa) + (Wb).vy * (Ab)) / ((Aa) + (Ab)), ((Wa).p * (Aa) + (Wb).p * (Ab)) / ((Aa) + (Ab)) }
        for (int ic = c->ICl; ic <= c->ICu; ++ic) {
            const int i0 = FINE_IDX(Ng, ic);
            double An0 = f->A[CI(f, i0, f->JCu)], An1 = f->A[CI(f, i0 + 1, f->JCu)];
            double As0 = f->A[CI(f, i0, f->JCl)], As1 = f->A[CI(f, i0 + 1, f->JCl)];
            pst wn = WAVG(f->WoN[i0], f->WoN[i0 + 1], An0, An1);
            pst ws = WAVG(f->WoS[i0], f->WoS[i0 + 1], As0, As1);
            c->WoN[ic] = wn; c->WoS[ic] = ws;
        }
        for (int jc = c->JCl; jc <= c->JCu; ++jc) {
            const int j0 = FINE_IDX(Ng, jc);
            double Aw0 = f->A[CI(f, f->ICl, j0)], Aw1 = f->A[CI(f, f->ICl, j0 + 1)];
            double Ae0 = f->A[CI(f, f->ICu, j0)], Ae1 = f->A[CI(f, f->ICu, j0 + 1)];
            pst ww = WAVG(f->WoW[j0], f->WoW[j0 + 1], Aw0, Aw1);
            pst we = WAVG(f->WoE[j0], f->WoE[j0 + 1], Ae0, Ae1);
            c->WoW[jc] = ww; c->WoE[jc] = we;
        }
#undef WAVG
    }
}
/* Update_Primitive_Variables :2423-2445 */
static void mg_update_primitive(orc_ctx *C) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *B = &C->b[b];
        for (int i = B->ICl; i <= B->ICu; ++i)
            for (int j = B->JCl; j <= B->JCu; ++j) B->W[CI(B, i, j)] = W_of_U(B->U[CI(B, i, j)]);
    }
}
/* Subtract_dUdt_from_P :2331-2352 */
static void mg_subtract_dudt(orc_ctx *C) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *B = &C->b[b];
        for (int i = B->ICl; i <= B->ICu; ++i)
            for (int j = B->JCl; j <= B->JCu; ++j) {
                cst *P = &B->P[CI(B, i, j)];
                const cst r = B->dUdt[0][CI(B, i, j)];
                P->d -= r.d; P->mx -= r.mx; P->my -= r.my; P->E -= r.E;
            }
    }
}
/* CFL_Multigrid :2364-2407: a coarse cell's dt no larger than its fine
   cells' dt scaled by sqrt(A_coarse/A_fine) */
static void mg_cfl_multigrid(orc_ctx *F, orc_ctx *C) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *f = &F->b[b], *c = &C->b[b];
        const int Ng = f->ng;
        for (int ic = c->ICl; ic <= c->ICu; ++ic)
            for (int jc = c->JCl; jc <= c->JCu; ++jc) {
                const int i0 = FINE_IDX(Ng, ic), j0 = FINE_IDX(Ng, jc);
                const double Ac = c->A[CI(c, ic, jc)];
                const int ne = CI(f, i0 + 1, j0 + 1), se = CI(f, i0 + 1, j0), nw = CI(f, i0, j0 + 1), sw = CI(f, i0, j0);
                const double dt_NE = sqrt(Ac / f->A[ne]) * f->dt[ne];
                const double dt_SE = sqrt(Ac / f->A[se]) * f->dt[se];
                const double dt_NW = sqrt(Ac / f->A[nw]) * f->dt[nw];
                const double dt_SW = sqrt(Ac / f->A[sw]) * f->dt[sw];
                double *dt = &c->dt[CI(c, ic, jc)];
                *dt = dmin(*dt, dmin(dt_NE, dmin(dt_SE, dmin(dt_NW, dt_SW))));
            }
    }
}
/* Store_Current_Solution_in_uo_MG :2294-2316 and Evaluate_Solution_Changes
   :2454-2478 (all cells, ghosts included) */
static void mg_store_uo(orc_ctx *C) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *B = &C->b[b];
        memcpy(B->uo, B->U, sizeof(cst) * B->nci * B->ncj);
    }
}
static void mg_solution_changes(orc_ctx *C) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *B = &C->b[b];
        for (int q = 0; q < B->nci * B->ncj; ++q) {
            B->U[q].d -= B->uo[q].d; B->U[q].mx -= B->uo[q].mx;
            B->U[q].my -= B->uo[q].my; B->U[q].E -= B->uo[q].E;
        }
    }
}

/* Bilinear_Interpolation, CFD.h:1367-1416 (Zingg and Yarrow): the point P0
   in the quadrilateral P1..P4 (clockwise from P1), 1 if the solve fails or
   zeta falls outside [0, 1] */
static int bilinear(const cst *U1, double x1, double y1, const cst *U2, double x2, double y2, const cst *U3, double x3,
                    double y3, const cst *U4, double x4, double y4, double x0, double y0, cst *U0) {
    double a0 = x1, a1 = x4 - x1, a2 = x2 - x1, a3 = x1 + x3 - x2 - x4;
    double b0 = y1, b1 = y4 - y1, b2 = y2 - y1, b3 = y1 + y3 - y2 - y4;
    double a = a1 * b3 - a3 * b1;
    double b = a0 * b3 - b3 * x0 + a1 * b2 - a2 * b1 + a3 * y0 - a3 * b0;
    double c = a0 * b2 - b2 * x0 + a2 * y0 - a2 * b0;
    double zeta0, eta0;
    if (a > TOLER || a < -TOLER) {
        if (b * b - 4 * a * c < 0) return 1;
        zeta0 = (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
        if (zeta0 < 0 || zeta0 > 1) {
            zeta0 = (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
            if (zeta0 < 0 || zeta0 > 1) return 1;
        }
    } else if (b > TOLER || b < -TOLER) {
        zeta0 = -c / b;
        if (zeta0 < 0 || zeta0 > 1) return 1;
    } else {
        return 1;
    }
    if (b2 + b3 * zeta0 < TOLER && b2 + b3 * zeta0 > -TOLER) return 1;
    eta0 = (y0 - b0 - b1 * zeta0) / (b2 + b3 * zeta0);
#define BIL(m) U0->m = U1->m + (U4->m - U1->m) * zeta0 + (U2->m - U1->m) * eta0 + (U1->m + U3->m - U2->m - U4->m) * zeta0 * eta0
    BIL(d); BIL(mx); BIL(my); BIL(E);
#undef BIL
    return 0;
}
int orc_bilinear(const double *U, const double *xy, double x0, double y0, double *out) {
    cst u[4], r;
    memcpy(u, U, sizeof u);
    int e = bilinear(&u[0], xy[0], xy[1], &u[1], xy[2], xy[3], &u[2], xy[4], xy[5], &u[3], xy[6], xy[7], x0, y0, &r);
    memcpy(out, &r, sizeof r);
    return e;
}

/* the interpolation quadrilaterals each fine cell tries, by the offset of
   the quadrilateral's SW coarse cell from the parent: fine SW, SE, NE, NW in
   the order of Prolong_and_Update_Solution_Blocks :1718-2211 */
static const int MG_QUAD[4][4][2] = {
    {{-1, -1}, {-1, 0}, {0, -1}, {0, 0}},   /* SW */
    {{0, -1}, {0, 0}, {-1, -1}, {-1, 0}},   /* SE */
    {{0, 0}, {0, -1}, {-1, 0}, {-1, -1}},   /* NE */
    {{-1, 0}, {0, 0}, {-1, -1}, {0, -1}},   /* NW */
};
static const int MG_FINE[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

static int unphysical(cst U) { return U.d <= ZERO || U.E <= ZERO || U_e(U) <= ZERO; }

/* near a Dirichlet-type boundary (BC_FIXED is the only one on this path):
   prolong by injection, :1675-1716 */
static int next_to_dirichlet(const oblk *c, int ic, int jc) {
    return (ic == c->ICl && c->bcW[jc] == CFFC_BC_FIXED) || (ic == c->ICu && c->bcE[jc] == CFFC_BC_FIXED) ||
           (jc == c->JCl && c->bcS[ic] == CFFC_BC_FIXED) || (jc == c->JCu && c->bcN[ic] == CFFC_BC_FIXED);
}

/* Prolong_and_Update_Solution_Blocks :1627-2213: coarse U holds the solution
   change (Evaluate_Solution_Changes); add it to the fine cells, bilinear or
   injected, halving an update that leaves the state unphysical.  A failed
   interpolation restarts the coarse cell with injection from its SW fine
   cell, which then receives a second update, as in the reference. */
static void mg_prolong_update(orc_ctx *C, orc_ctx *F, const cffc_mg_desc *d) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *c = &C->b[b], *f = &F->b[b];
        const int Ng = c->ng;
        if (!d->prolong_using_injection) {
            for (int jc = c->JCl; jc <= c->JCu; ++jc) {
                if (c->bcE[jc] == CFFC_BC_CONSTANT_EXTRAPOLATION) c->U[CI(c, c->ICu + 1, jc)] = c->U[CI(c, c->ICu, jc)];
                if (c->bcW[jc] == CFFC_BC_CONSTANT_EXTRAPOLATION) c->U[CI(c, c->ICl - 1, jc)] = c->U[CI(c, c->ICl, jc)];
            }
            for (int ic = c->ICl; ic <= c->ICu; ++ic) {
                if (c->bcS[ic] == CFFC_BC_CONSTANT_EXTRAPOLATION) c->U[CI(c, ic, c->JCl - 1)] = c->U[CI(c, ic, c->JCl)];
                if (c->bcN[ic] == CFFC_BC_CONSTANT_EXTRAPOLATION) c->U[CI(c, ic, c->JCu + 1)] = c->U[CI(c, ic, c->JCu)];
            }
        }
        for (int ic = c->ICl; ic <= c->ICu; ++ic)
            for (int jc = c->JCl; jc <= c->JCu; ++jc) {
                const int i0 = FINE_IDX(Ng, ic), j0 = FINE_IDX(Ng, jc);
                const cst Uc = c->U[CI(c, ic, jc)];
                if (d->injection_at_dirichlet && next_to_dirichlet(c, ic, jc)) {
                    for (int q = 0; q < 4; ++q) {
                        cst *Uf = &f->U[CI(f, i0 + MG_FINE[q][0], j0 + MG_FINE[q][1])];
                        Uf->d += Uc.d; Uf->mx += Uc.mx; Uf->my += Uc.my; Uf->E += Uc.E;
                    }
                    continue;
                }
                int inject = d->prolong_using_injection;
                for (int q = 0; q < 4; ++q) {
                    const int fi = i0 + MG_FINE[q][0], fj = j0 + MG_FINE[q][1];
                    cst dU;
                    if (!inject) {
                        int err = 1;
                        for (int t = 0; t < 4 && err; ++t) {
                            const int a = ic + MG_QUAD[q][t][0], e = jc + MG_QUAD[q][t][1];
                            const int p1 = CI(c, a, e), p2 = CI(c, a, e + 1), p3 = CI(c, a + 1, e + 1), p4 = CI(c, a + 1, e);
                            err = bilinear(&c->U[p1], c->xc[p1], c->yc[p1], &c->U[p2], c->xc[p2], c->yc[p2], &c->U[p3],
                                           c->xc[p3], c->yc[p3], &c->U[p4], c->xc[p4], c->yc[p4], f->xc[CI(f, fi, fj)],
                                           f->yc[CI(f, fi, fj)], &dU);
                        }
                        if (err) {
                            inject = 1;
                            q = -1;   /* goto restart */
                            continue;
                        }
                    } else {
                        dU = Uc;
                    }
                    cst *Uf = &f->U[CI(f, fi, fj)];
                    if (d->update_stability_switch) {
                        cst Un = {Uf->d + dU.d, Uf->mx + dU.mx, Uf->my + dU.my, Uf->E + dU.E};
                        if (unphysical(Un)) {
                            int n = 1;
                            while (n <= d->max_update_reductions) {
                                dU.d *= HALF; dU.mx *= HALF; dU.my *= HALF; dU.E *= HALF;
                                cst Um = {Uf->d + dU.d, Uf->mx + dU.mx, Uf->my + dU.my, Uf->E + dU.E};
                                if (!unphysical(Um)) break;
                                n++;
                            }
                            if (n > d->max_update_reductions) { dU.d = 0; dU.mx = 0; dU.my = 0; dU.E = 0; }
                        }
                    }
                    Uf->d += dU.d; Uf->mx += dU.mx; Uf->my += dU.my; Uf->E += dU.E;
                }
            }
    }
}

/* dUdt[0] of the whole level: Residual_Evaluation_for_Multigrid :2755-2858
   (no resolution changes within a level, so no flux corrections), plus the
   forcing term P on coarse levels */
static void mg_residual(orc_ctx *C, int add_P) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *B = &C->b[b];
        block_residual(B, &C->s);
        if (add_P)
            for (int j = B->JCl; j <= B->JCu; ++j)
                for (int i = B->ICl; i <= B->ICu; ++i) {
                    cst *r = &B->dUdt[0][CI(B, i, j)];
                    const cst P = B->P[CI(B, i, j)];
                    r->d += P.d; r->mx += P.mx; r->my += P.my; r->E += P.E;
                }
    }
}

/* Smooth :2865-3110 */
static int mg_smooth(orc_mg *mg, int top, int l) {
    orc_ctx *C = mg->lv[l];
    double dtmin;
    orc_cfl(C, &dtmin, 0);
    if (l != top) mg_cfl_multigrid(mg->lv[l - 1], C);
    for (int st = 1; st <= C->s.n_stage; ++st) {
        orc_exchange_and_bcs(C);
        orc_dudt_stage(C, st);
        if (l > top) {
            /* Apply_the_FAS_Multigrid_Forcing_Term :2678-2745 */
            const int k = stage_k_residual(&C->s, st);
            for (int b = 0; b < C->nblk; ++b) {
                oblk *B = &C->b[b];
                for (int j = B->JCl; j <= B->JCu; ++j)
                    for (int i = B->ICl; i <= B->ICu; ++i) {
                        const int q = CI(B, i, j);
                        const double cdt = C->s.cfl_number * B->dt[q];
                        cst *r = &B->dUdt[k][q];
                        r->d += cdt * B->P[q].d; r->mx += cdt * B->P[q].mx;
                        r->my += cdt * B->P[q].my; r->E += cdt * B->P[q].E;
                    }
            }
        }
        /* Residual_Smoothing under TIME_STEPPING_MULTIGRID smooths dUdt[0] */
        if (mg->rs_iters > 0)
            for (int b = 0; b < C->nblk; ++b) block_residual_smoothing(&C->b[b], 0, mg->rs_eps, mg->rs_iters);
        int64_t bad;
        int e = orc_update_stage(C, st, &bad);
        if (e) return e;
    }
    orc_exchange_and_bcs(C);
    mg_residual(C, l > top);
    return 0;
}

/* Coarse_Grid_Correction :3118-3370 */
static int mg_cgc(orc_mg *mg, int top, int l) {
    const cffc_mg_desc *d = &mg->d;
    int e;
    if (l == mg->L - 1) {
        mg_store_uo(mg->lv[l]);
        for (int nu = 1; nu <= d->smooths_on_coarsest_level; ++nu)
            if ((e = mg_smooth(mg, top, l))) return e;
    } else {
        /* the W cycle's two passes apply below the finest level, whatever
           the top level (:3151-3158) */
        const int m_cycle = (d->i_cycle == CFFC_MULTIGRID_V_CYCLE || l == 0) ? 1 : 2;
        for (int m = 1; m <= m_cycle; ++m) {
            mg_store_uo(mg->lv[l]);
            const int n_pre = (l == top) ? d->smooths_on_finest_level : d->pre_smooths;
            for (int nu = 0; nu < n_pre; ++nu)
                if ((e = mg_smooth(mg, top, l))) return e;
            mg_restrict(mg->lv[l], mg->lv[l + 1], 0);
            mg_update_primitive(mg->lv[l + 1]);
            orc_exchange_and_bcs(mg->lv[l + 1]);
            if (d->defect_correction) {
                mg_restrict(mg->lv[l], mg->lv[l + 1], 1);
                mg_residual(mg->lv[l + 1], 0);
                mg_subtract_dudt(mg->lv[l + 1]);
            }
            if ((e = mg_cgc(mg, top, l + 1))) return e;
            mg_prolong_update(mg->lv[l + 1], mg->lv[l], d);
            mg_update_primitive(mg->lv[l]);
            orc_exchange_and_bcs(mg->lv[l]);
            const int n_post = (l == top) ? 0 : d->post_smooths;
            for (int nu = 0; nu < n_post; ++nu)
                if ((e = mg_smooth(mg, top, l))) return e;
        }
    }
    if (l != top) mg_solution_changes(mg->lv[l]);
    return 0;
}

/* Prolong_Solution_Blocks :1200-1625 (full multigrid): the fine cells take
   the coarse solution, bilinear in the same quadrilateral order as the
   correction (injection when none holds the cell, or always), assigned
   rather than added; no Dirichlet injection and no update reduction.  The
   fine W is left as it was (Execute follows with messages and BCs only,
   :3716-3728). */
static void mg_prolong_solution(orc_ctx *C, orc_ctx *F, const cffc_mg_desc *d) {
    for (int b = 0; b < C->nblk; ++b) {
        oblk *c = &C->b[b], *f = &F->b[b];
        const int Ng = c->ng;
        for (int ic = c->ICl; ic <= c->ICu; ++ic)
            for (int jc = c->JCl; jc <= c->JCu; ++jc) {
                const int i0 = FINE_IDX(Ng, ic), j0 = FINE_IDX(Ng, jc);
                const cst Uc = c->U[CI(c, ic, jc)];
                int inject = d->prolong_using_injection;
                for (int q = 0; q < 4; ++q) {
                    const int fi = i0 + MG_FINE[q][0], fj = j0 + MG_FINE[q][1];
                    cst Uf = Uc;
                    if (!inject) {
                        int err = 1;
                        for (int t = 0; t < 4 && err; ++t) {
                            const int a = ic + MG_QUAD[q][t][0], e = jc + MG_QUAD[q][t][1];
                            const int p1 = CI(c, a, e), p2 = CI(c, a, e + 1), p3 = CI(c, a + 1, e + 1), p4 = CI(c, a + 1, e);
                            err = bilinear(&c->U[p1], c->xc[p1], c->yc[p1], &c->U[p2], c->xc[p2], c->yc[p2], &c->U[p3],
                                           c->xc[p3], c->yc[p3], &c->U[p4], c->xc[p4], c->yc[p4], f->xc[CI(f, fi, fj)],
                                           f->yc[CI(f, fi, fj)], &Uf);
                        }
                        if (err) {
                            inject = 1;
                            q = -1;
                            continue;
                        }
                    }
                    f->U[CI(f, fi, fj)] = Uf;
                }
            }
    }
}

int orc_mg_create(const cffc_mesh_desc *meshes, int nlevels, const cffc_scheme_desc *smoother,
                  const cffc_mg_desc *d, orc_mg **out) {
    if (nlevels < 2 || nlevels != d->levels) return -1;
    orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
    mg->L = nlevels;
    mg->d = *d;
    mg->fine_s = *smoother;
    mg->coarse_s = *smoother;
    if (d->first_order_coarse) {
        mg->coarse_s.i_flux_function = CFFC_FLUX_FUNCTION_HLLE;
        mg->coarse_s.i_limiter = CFFC_LIMITER_ZERO;
    }
    mg->lv = (orc_ctx **)calloc(nlevels, sizeof(orc_ctx *));
    for (int l = 0; l < nlevels; ++l) {
        cffc_scheme_desc s = *smoother;
        if (l > 0 && d->first_order_coarse) {
            s.i_flux_function = CFFC_FLUX_FUNCTION_HLLE;
            s.i_limiter = CFFC_LIMITER_ZERO;
        }
        int e = orc_create(&meshes[l], &s, &mg->lv[l]);
        if (e) { orc_mg_destroy(mg); return e; }
        if (l > 0 && (mg->lv[l]->nblk != mg->lv[0]->nblk)) { orc_mg_destroy(mg); return -1; }
        for (int b = 0; b < mg->lv[l]->nblk; ++b) {
            oblk *B = &mg->lv[l]->b[b];
            if (l > 0) {
                const oblk *Fb = &mg->lv[l - 1]->b[b];
                if (Fb->ni != 2 * B->ni || Fb->nj != 2 * B->nj) { orc_mg_destroy(mg); return -1; }
            }
            B->P = (cst *)calloc(B->nci * B->ncj, sizeof(cst));
            B->uo = (cst *)calloc(B->nci * B->ncj, sizeof(cst));
        }
    }
    *out = mg;
    return 0;
}
void orc_mg_destroy(orc_mg *mg) {
    if (!mg) return;
    for (int l = 0; l < mg->L; ++l) orc_destroy(mg->lv[l]);
    free(mg->lv);
    free(mg);
}
orc_ctx *orc_mg_level(orc_mg *mg, int l) { return (l >= 0 && l < mg->L) ? mg->lv[l] : 0; }
int orc_mg_set_residual_smoothing(orc_mg *mg, double epsilon, int iterations) {
    mg->rs_eps = epsilon;
    mg->rs_iters = iterations;
    return 0;
}
/* the start of Execute :3395-3420 for a state set on level 0: every coarse
   level restricted from the one above (the IC_RESTART path of Apply_ICs
   :719-733) with its boundary reference states, then messages and BCs */
int orc_mg_initialize(orc_mg *mg) {
    orc_exchange_and_bcs(mg->lv[0]);
    for (int l = 1; l < mg->L; ++l) {
        mg_restrict(mg->lv[l - 1], mg->lv[l], 0);
        mg_update_primitive(mg->lv[l]);
        mg_restrict_wo(mg->lv[l - 1], mg->lv[l]);
        orc_exchange_and_bcs(mg->lv[l]);
    }
    return 0;
}
/* one FAS cycle from level 0 */
int orc_mg_cycle(orc_mg *mg) {
    mg_set_top(mg, 0);
    return mg_cgc(mg, 0, 0);
}
/* Execute :3375-3745: with fmg_cycles > 0 the full-multigrid start-up
   first (top level Levels-2 up to 1, fmg_cycles cycles each, stopping early
   on the FMG tolerances, then Prolong_Solution_Blocks to the next level
   with messages and BCs), then the regular cycles from level 0.  Before
   each cycle the residual norms of the top level (the dUdt[0] the last
   cycle's smoothing left), the relative norm against the second check of
   that level, limiter freezing at freeze_level (<= 0: never; unfrozen
   again at a new top level, :3462-3476), and the exits: max_cycles cycles
   in all, or the absolute / relative tolerance (FMG ones above level 0).
   norms: 3 per check. */
int orc_mg_execute(orc_mg *mg, int max_cycles, double freeze_level, double *norms, int *cycles_done, int *checks) {
    int first_step = 1, frozen = 0, n = 0, e = 0, nchk = 0;
    const int top0 = mg->d.fmg_cycles > 0 ? mg->L - 2 : 0;
    for (int top = top0; top >= 0 && !e; --top) {
        mg_set_top(mg, top);
        if (!first_step && freeze_level > 0 && frozen && mg->fine_s.i_limiter != CFFC_LIMITER_ZERO) {
            for (int l = top; l < mg->L; ++l) orc_freeze_limiters(mg->lv[l], 0);
            frozen = 0;
        }
        const long long ncyc = top != 0 ? mg->d.fmg_cycles : 0x7fffffffffffffffLL;
        double init_l2 = -1.0, ratio = -1.0;
        for (long long cycles = 1; cycles <= ncyc; ++cycles) {
            double nr[3];
            orc_residual_norms(mg->lv[top], 1, nr);
            if (norms) { norms[3 * nchk] = nr[0]; norms[3 * nchk + 1] = nr[1]; norms[3 * nchk + 2] = nr[2]; }
            nchk++;
            if (cycles == 2) init_l2 = nr[1];
            else if (cycles > 2 && fabs(init_l2) > NANO) ratio = nr[1] / init_l2;
            int stop = n >= max_cycles;
            if (cycles > 2) {
                const double at = top == 0 ? mg->d.absolute_tolerance : mg->d.fmg_absolute_tolerance;
                const double rt = top == 0 ? mg->d.relative_tolerance : mg->d.fmg_relative_tolerance;
                if (nr[1] < at) stop = 1;
                if (ratio < rt) stop = 1;
            }
            if (stop) break;
            if (!first_step && cycles > 1 && freeze_level > 0 && !frozen &&
                mg->fine_s.i_limiter != CFFC_LIMITER_ZERO && nr[1] <= freeze_level) {
                for (int l = top; l < mg->L; ++l) orc_freeze_limiters(mg->lv[l], 1);
                frozen = 1;
            }
            if ((e = mg_cgc(mg, top, top))) break;
            first_step = 0;
            n++;
        }
        /* a spent cycle budget still prolongs through the remaining
           levels, each stopping at its first check, as the reference does */
        if (top != 0 && !e) {
            mg_prolong_solution(mg->lv[top], mg->lv[top - 1], &mg->d);
            orc_exchange_and_bcs(mg->lv[top - 1]);
        }
    }
    mg_set_top(mg, 0);
    if (cycles_done) *cycles_done = n;
    if (checks) *checks = nchk;
    orc_exchange_and_bcs(mg->lv[0]);
    return e;
}
/* P and uo of one level, interior cells in block order (tests) */
int orc_mg_get_P(orc_mg *mg, int l, double *out) {
    orc_ctx *C = mg->lv[l];
    int64_t q = 0;
    for (int b = 0; b < C->nblk; ++b) {
        oblk *B = &C->b[b];
        for (int j = B->JCl; j <= B->JCu; ++j)
            for (int i = B->ICl; i <= B->ICu; ++i, ++q) memcpy(out + 4 * q, &B->P[CI(B, i, j)], sizeof(cst));
    }
    return 0;
}
/* HLLE flux (known-answer tests) */
void orc_flux_hlle_n(const double *Wl, const double *Wr, double nx, double ny, double *F) {
    pst l, r;
    memcpy(&l, Wl, sizeof l); memcpy(&r, Wr, sizeof r);
    cst f = flux_hlle_n(l, r, nx, ny);
    memcpy(F, &f, sizeof f);
}
void orc_axisym_source(const double *W, double x, double y, double *out) {
    (void)x;
    pst w; memcpy(&w, W, sizeof w);
    cst S = axisym_source(w, y);
    memcpy(out, &S, sizeof S);
}
/* Set_Global_TimeStep (Euler2DQuadMultiBlock.cc:994) with a given step */
int orc_set_global_dt(orc_ctx *ctx, double dt) {
    for (int b = 0; b < ctx->nblk; ++b) block_set_global_dt(&ctx->b[b], dt);
    return 0;
}
