/*
 * ocfd_oracle.c - CPU restatement (FP64, plain C) of the OpenCFD-SCU
 * right-hand-side path: Steger-Warming split, WENO7 / WENO7-SYMBO / UP7 /
 * OMP6 / WENO5 / NND2 / hybrid_auto (Jameson style) half-node fluxes with the
 * reference's boundary closures, CD6/CD8 viscous terms, SSP-RK3, cons<->prim,
 * Sutherland viscosity, IBC 108 (compression corner), the shock-capturing
 * filter, periodic ghost fills and the metric computation.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in the
 * product links or calls this file.
 *
 * Every routine names the reference file:line it restates.  Expressions are
 * written in the reference's operand order so that rounding follows it.
 * The reference's launch semantics (which points are written, ghost layers
 * that are never written and therefore stay at their initial value, the order
 * in which terms are added into du) are restated in the drivers.
 *
 * The point loops are parallel over OpenMP threads (each iteration writes
 * its own outputs, so the result does not depend on the thread count).
 *
 * Deliberate choice where the reference is undefined: the hybrid kernels set
 * Hyscheme_flag only for variable 0 (src/OCFD_Schemes_hybrid_auto.cu:874,907,
 * 971,1005) and leave it uninitialised for variables 1..4; this restatement
 * uses the flag for all five variables.
 *
 * Pinned by closed forms and invariants (tests/test_oracle_pins.py,
 * tests/test_oracle_pins_nonlinear.py): the linear schemes and closures,
 * the nonlinear reconstructions' ENO behaviour at a step and the MP limiter's
 * extremum preservation, Steger-Warming, SSP-RK3, the viscous operator and
 * heat flux, metrics, boundary conditions, filter stencil and sensor, the
 * characteristic projection, the Q criterion, and WENO7's face value from
 * first principles (candidate values of the cell-mean cubics, smoothness
 * indicators = 34560 x the integrated squared derivatives of the point-value
 * cubics, optimal weights 1:12:18:4; stencil 2's second-derivative term is
 * taken at node i+1, as src/OCFD_Schemes.cu:632 writes it), WENO5's the
 * same way (Jiang-Shu indicators, weights 1:6:3) and WENO7-SYMBO's
 * nonlinear branch (25920 x the integrals for the cell-mean cubics, the
 * downwind stencil at the largest indicator, first-power weights) and the
 * SYMBO TV limiter's switch between the linear and nonlinear weights.
 * Parity unpinned (restated from the source and checked against the HIP
 * path only): SYMBO's five optimised weights beyond their sum, the
 * rand()-seeded IBC 108 and IBC 124 wall blowing tables, the flag of hybrid
 * variables 1..4 above, the filter's
 * coverage geometry (which rows the reference's launch shapes reach), the
 * IBC 108 top-face start x_begin, and which node-to-node difference the
 * CHARTERIC boundary-node rules take.
 */
#include "ocfd_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define LAP 4
#define NV 5

struct orc {
    ocfd_config c;
    int nx, ny, nz, NX2, NY2, NZ2;
    long nlap, nint;
    double hx, hy, hz;
    double Cv, Cp, Tsb, amu_C0, tmp0, split_C1, split_C3, vis_c;
    /* LAP-box fields */
    double *d, *u, *v, *w, *T, *P, *cc;
    double *A[13]; /* Axx Ayy Azz Akx Aky Akz Aix Aiy Aiz Asx Asy Asz Ajac */
    double *fp[3], *fm[3];
    double *Ev[4];
    double *f_lap;
    /* interior fields */
    double *f, *fn, *du, *Amu;
    double *grad[12]; /* uk ui us vk vi vs wk wi ws Tk Ti Ts */
    double *visd[4];
    int *flag[3];
    /* IBC 108 */
    double *ub1; /* [4][ny] */
    double *pfx, *pgz;
    double TM[64], fait[64];
    int x_begin;
    int have_profile;
    /* IBC 124: inlet [5][nz][ny], upper [5][ny][nx], wall coefficients
     * [3][ny][nx], upper disturbance [ny][nx]; caller tables */
    double *in124, *up124, *coeff124, *udist124;
    double *in124_user, *up124_user;
    int v_dist_need;
    int istep;
    double tt;
};

enum { AXX, AYY, AZZ, AKX, AKY, AKZ, AIX, AIY, AIZ, ASX, ASY, ASZ, AJAC };

static inline long IL(const orc *o, int x, int y, int z) {
    return (long)x + (long)o->NX2 * ((long)y + (long)o->NY2 * (long)z);
}
static inline long I0(const orc *o, int x, int y, int z) {
    return (long)x + (long)o->nx * ((long)y + (long)o->ny * (long)z);
}

/* ======================================================================
 * Scheme kernels, src/OCFD_Schemes.cu
 * ====================================================================== */

/* sign/minmod2/minmod4, src/OCFD_Schemes.cu:767-797 */
static double sgn(double x1, double x2) { return x2 >= 0 ? fabs(x1) : -fabs(x1); }
static double minmod2(double x1, double x2) {
    return 0.5 * (sgn(1.0, x1) + sgn(1.0, x2)) * fmin(fabs(x1), fabs(x2));
}
static double minmod4(double x1, double x2, double x3, double x4) {
    double m = 0.5 * (sgn(1.0, x1) + sgn(1.0, x2));
    m = m * fabs(0.5 * (sgn(1.0, x1) + sgn(1.0, x3)));
    m = m * fabs(0.5 * (sgn(1.0, x1) + sgn(1.0, x4)));
    double t = fmin(x1, x2);
    double t1 = fmin(x3, x4);
    t = fmin(t, t1);
    return m * t;
}

/* WENO7-SYMBO P, src/OCFD_Schemes.cu:391-478 */
static double weno7_symbo_P(int lmt, const double *s, double tvlim, double tvmax) {
    double S0, S1, S2, S3, S4, tmp, tmp1, TVmin = 0, TVmax = 1;
    if (lmt == 1) {
        S0 = fabs(s[1] - s[0]) + fabs(s[2] - s[1]) + fabs(s[3] - s[2]);
        S1 = fabs(s[2] - s[1]) + fabs(s[3] - s[2]) + fabs(s[4] - s[3]);
        S2 = fabs(s[3] - s[2]) + fabs(s[4] - s[3]) + fabs(s[5] - s[4]);
        S3 = fabs(s[4] - s[3]) + fabs(s[5] - s[4]) + fabs(s[6] - s[5]);
        S4 = fabs(s[5] - s[4]) + fabs(s[6] - s[5]) + fabs(s[7] - s[6]);
        tmp = fmin(S0, S1); tmp1 = fmin(S2, S3); tmp = fmin(tmp, tmp1); TVmin = fmin(tmp, S4);
        tmp = fmax(S0, S1); tmp1 = fmax(S2, S3); tmp = fmax(tmp, tmp1); TVmax = fmax(tmp, S4);
    }
    if (TVmax < tvlim * TVmin && TVmax < tvmax) {
        S0 = 0.0401954833730; S1 = 0.2493800006710; S2 = 0.4802686256260;
        S3 = 0.2009775476730; S4 = 0.0291783426580;
    } else {
        S0 = 0.0; S1 = 0.0; S2 = 0.0; S3 = 0.0; S4 = 0.0;
        tmp = -2.0 * s[0] + 9.0 * s[1] - 18.0 * s[2] + 11.0 * s[3]; S0 += 720.0 * tmp * tmp;
        tmp = s[1] - 6.0 * s[2] + 3.0 * s[3] + 2.0 * s[4]; S1 += 720.0 * tmp * tmp;
        tmp = -2.0 * s[2] - 3.0 * s[3] + 6.0 * s[4] - s[5]; S2 += 720.0 * tmp * tmp;
        tmp = -11.0 * s[3] + 18.0 * s[4] - 9.0 * s[5] + 2.0 * s[6]; S3 += 720.0 * tmp * tmp;
        tmp = -26.0 * s[4] + 57.0 * s[5] - 42.0 * s[6] + 11.0 * s[7]; S4 += 720.0 * tmp * tmp;

        tmp = -6.0 * s[0] + 24.0 * s[1] - 30.0 * s[2] + 12.0 * s[3]; S0 += 780.0 * tmp * tmp;
        tmp = 6.0 * s[2] - 12.0 * s[3] + 6.0 * s[4]; S1 += 780.0 * tmp * tmp;
        tmp = 6.0 * s[3] - 12.0 * s[4] + 6.0 * s[5]; S2 += 780.0 * tmp * tmp;
        tmp = 12.0 * s[3] - 30.0 * s[4] + 24.0 * s[5] - 6.0 * s[6]; S3 += 780.0 * tmp * tmp;
        tmp = 18.0 * s[4] - 48.0 * s[5] + 42.0 * s[6] - 12.0 * s[7]; S4 += 780.0 * tmp * tmp;

        tmp = -6.0 * s[0] + 18.0 * (s[1] - s[2]) + 6.0 * s[3]; S0 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[1] + 18.0 * (s[2] - s[3]) + 6.0 * s[4]; S1 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[2] + 18.0 * (s[3] - s[4]) + 6.0 * s[5]; S2 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[3] + 18.0 * (s[4] - s[5]) + 6.0 * s[6]; S3 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[4] + 18.0 * (s[5] - s[6]) + 6.0 * s[7]; S4 += 781.0 * tmp * tmp;
        tmp = fmax(S0, S1); tmp1 = fmax(S2, S3); tmp = fmax(tmp, tmp1); S4 = fmax(tmp, S4);
        {
            double t0, t1, t2, t3;
            t0 = (0.0401954833730) * (2.592e-6 + S1) * (2.592e-6 + S2) * (2.592e-6 + S3) * (2.592e-6 + S4);
            t1 = (0.2493800006710) * (2.592e-6 + S0) * (2.592e-6 + S2) * (2.592e-6 + S3) * (2.592e-6 + S4);
            t2 = (0.4802686256260) * (2.592e-6 + S0) * (2.592e-6 + S1) * (2.592e-6 + S3) * (2.592e-6 + S4);
            t3 = (0.2009775476730) * (2.592e-6 + S0) * (2.592e-6 + S1) * (2.592e-6 + S2) * (2.592e-6 + S4);
            S4 = (0.0291783426580) * (2.592e-6 + S0) * (2.592e-6 + S1) * (2.592e-6 + S2) * (2.592e-6 + S3);
            S3 = t3; S2 = t2; S1 = t1; S0 = t0;
        }
    }
    double am = S0 + S1 + S2 + S3 + S4;
    tmp1 = 0.0;
    tmp = -3.0 * s[0] + 13.0 * s[1] - 23.0 * s[2] + 25.0 * s[3]; tmp1 += S0 * tmp;
    tmp = s[1] - 5.0 * s[2] + 13.0 * s[3] + 3.0 * s[4]; tmp1 += S1 * tmp;
    tmp = -s[2] + 7.0 * s[3] + 7.0 * s[4] - s[5]; tmp1 += S2 * tmp;
    tmp = 3.0 * s[3] + 13.0 * s[4] - 5.0 * s[5] + s[6]; tmp1 += S3 * tmp;
    tmp = 25.0 * s[4] - 23.0 * s[5] + 13.0 * s[6] - 3.0 * s[7]; tmp1 += S4 * tmp;
    tmp1 /= (12.0 * am);
    return tmp1;
}

/* WENO7-SYMBO M, src/OCFD_Schemes.cu:503-589.  TVmin is never computed on
 * this side (src/OCFD_Schemes.cu:513-524), so the linear-weight branch is
 * never taken. */
static double weno7_symbo_M(int lmt, const double *s, double tvlim, double tvmax) {
    double S0, S1, S2, S3, S4, tmp, tmp1, TVmin = 0, TVmax = 1;
    if (lmt == 1) {
        S0 = fabs(s[7] - s[6]) + fabs(s[6] - s[5]) + fabs(s[5] - s[4]);
        S1 = fabs(s[6] - s[5]) + fabs(s[5] - s[4]) + fabs(s[4] - s[3]);
        S2 = fabs(s[5] - s[4]) + fabs(s[4] - s[3]) + fabs(s[3] - s[2]);
        S3 = fabs(s[4] - s[3]) + fabs(s[3] - s[2]) + fabs(s[2] - s[1]);
        S4 = fabs(s[3] - s[2]) + fabs(s[2] - s[1]) + fabs(s[1] - s[0]);
        tmp = fmin(S0, S1); tmp1 = fmin(S2, S3); tmp = fmax(tmp, tmp1); TVmax = fmax(tmp, S4);
    }
    if (TVmax < tvlim * TVmin && TVmax < tvmax) {
        S0 = 0.0401954833730; S1 = 0.2493800006710; S2 = 0.4802686256260;
        S3 = 0.2009775476730; S4 = 0.0291783426580;
    } else {
        S0 = 0.0; S1 = 0.0; S2 = 0.0; S3 = 0.0; S4 = 0.0;
        tmp = -2.0 * s[7] + 9.0 * s[6] - 18.0 * s[5] + 11.0 * s[4]; S0 += 720.0 * tmp * tmp;
        tmp = s[6] - 6.0 * s[5] + 3.0 * s[4] + 2.0 * s[3]; S1 += 720.0 * tmp * tmp;
        tmp = -2.0 * s[5] - 3.0 * s[4] + 6.0 * s[3] - s[2]; S2 += 720.0 * tmp * tmp;
        tmp = -11.0 * s[4] + 18.0 * s[3] - 9.0 * s[2] + 2.0 * s[1]; S3 += 720.0 * tmp * tmp;
        tmp = -26.0 * s[3] + 57.0 * s[2] - 42.0 * s[1] + 11.0 * s[0]; S4 += 720.0 * tmp * tmp;

        tmp = -6.0 * s[7] + 24.0 * s[6] - 30.0 * s[5] + 12.0 * s[4]; S0 += 780.0 * tmp * tmp;
        tmp = 6.0 * s[5] - 12.0 * s[4] + 6.0 * s[3]; S1 += 780.0 * tmp * tmp;
        tmp = 6.0 * s[4] - 12.0 * s[3] + 6.0 * s[2]; S2 += 780.0 * tmp * tmp;
        tmp = 12.0 * s[4] - 30.0 * s[3] + 24.0 * s[2] - 6.0 * s[1]; S3 += 780.0 * tmp * tmp;
        tmp = 18.0 * s[3] - 48.0 * s[2] + 42.0 * s[1] - 12.0 * s[0]; S4 += 780.0 * tmp * tmp;

        tmp = -6.0 * s[7] + 18.0 * (s[6] - s[5]) + 6.0 * s[4]; S0 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[6] + 18.0 * (s[5] - s[4]) + 6.0 * s[3]; S1 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[5] + 18.0 * (s[4] - s[3]) + 6.0 * s[2]; S2 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[4] + 18.0 * (s[3] - s[2]) + 6.0 * s[1]; S3 += 781.0 * tmp * tmp;
        tmp = -6.0 * s[3] + 18.0 * (s[2] - s[1]) + 6.0 * s[0]; S4 += 781.0 * tmp * tmp;
        tmp = fmax(S0, S1); tmp1 = fmax(S2, S3); tmp = fmax(tmp, tmp1); S4 = fmax(tmp, S4);
        {
            double t0, t1, t2, t3;
            t0 = (0.0401954833730) * (2.592e-6 + S1) * (2.592e-6 + S2) * (2.592e-6 + S3) * (2.592e-6 + S4);
            t1 = (0.2493800006710) * (2.592e-6 + S0) * (2.592e-6 + S2) * (2.592e-6 + S3) * (2.592e-6 + S4);
            t2 = (0.4802686256260) * (2.592e-6 + S0) * (2.592e-6 + S1) * (2.592e-6 + S3) * (2.592e-6 + S4);
            t3 = (0.2009775476730) * (2.592e-6 + S0) * (2.592e-6 + S1) * (2.592e-6 + S2) * (2.592e-6 + S4);
            S4 = (0.0291783426580) * (2.592e-6 + S0) * (2.592e-6 + S1) * (2.592e-6 + S2) * (2.592e-6 + S3);
            S3 = t3; S2 = t2; S1 = t1; S0 = t0;
        }
    }
    double am = S0 + S1 + S2 + S3 + S4;
    tmp1 = 0.0;
    tmp = -3.0 * s[7] + 13.0 * s[6] - 23.0 * s[5] + 25.0 * s[4]; tmp1 += S0 * tmp;
    tmp = s[6] - 5.0 * s[5] + 13.0 * s[4] + 3.0 * s[3]; tmp1 += S1 * tmp;
    tmp = -s[5] + 7.0 * s[4] + 7.0 * s[3] - s[2]; tmp1 += S2 * tmp;
    tmp = 3.0 * s[4] + 13.0 * s[3] - 5.0 * s[2] + s[1]; tmp1 += S3 * tmp;
    tmp = 25.0 * s[3] - 23.0 * s[2] + 13.0 * s[1] - 3.0 * s[0]; tmp1 += S4 * tmp;
    tmp1 /= (12.0 * am);
    return tmp1;
}

/* WENO7 P (Balsara-Shu), src/OCFD_Schemes.cu:616-663 */
static double weno7_P(const double *s) {
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, S3 = 0.0, tmp, tmp1, tmp2;
    tmp = -2.0 * s[0] + 9.0 * s[1] - 18.0 * s[2] + 11.0 * s[3]; S0 += 960 * tmp * tmp;
    tmp1 = -6.0 * s[0] + 24.0 * s[1] - 30.0 * s[2] + 12.0 * s[3]; S0 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[0] + 18.0 * (s[1] - s[2]) + 6.0 * s[3]; S0 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S0 += 80.0 * tmp;

    tmp = s[1] - 6.0 * s[2] + 3.0 * s[3] + 2.0 * s[4]; S1 += 960 * tmp * tmp;
    tmp1 = 6.0 * s[2] - 12.0 * s[3] + 6.0 * s[4]; S1 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[1] + 18.0 * (s[2] - s[3]) + 6.0 * s[4]; S1 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S1 += 80.0 * tmp;

    tmp = -2.0 * s[2] - 3.0 * s[3] + 6.0 * s[4] - s[5]; S2 += 960 * tmp * tmp;
    tmp1 = 6.0 * s[3] - 12.0 * s[4] + 6.0 * s[5]; S2 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[2] + 18.0 * (s[3] - s[4]) + 6.0 * s[5]; S2 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S2 += 80.0 * tmp;

    tmp = -11.0 * s[3] + 18.0 * s[4] - 9.0 * s[5] + 2.0 * s[6]; S3 += 960 * tmp * tmp;
    tmp1 = 12.0 * s[3] - 30.0 * s[4] + 24.0 * s[5] - 6.0 * s[6]; S3 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[3] + 18.0 * (s[4] - s[5]) + 6.0 * s[6]; S3 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S3 += 80.0 * tmp;

    tmp = ((3.456e-4 + S1) * (3.456e-4 + S1)) * ((3.456e-4 + S2) * (3.456e-4 + S2)) * ((3.456e-4 + S3) * (3.456e-4 + S3));
    tmp1 = (12.0) * ((3.456e-4 + S0) * (3.456e-4 + S0)) * ((3.456e-4 + S2) * (3.456e-4 + S2)) * ((3.456e-4 + S3) * (3.456e-4 + S3));
    tmp2 = (18.0) * ((3.456e-4 + S0) * (3.456e-4 + S0)) * ((3.456e-4 + S1) * (3.456e-4 + S1)) * ((3.456e-4 + S3) * (3.456e-4 + S3));
    S3 = (4.0) * ((3.456e-4 + S0) * (3.456e-4 + S0)) * ((3.456e-4 + S1) * (3.456e-4 + S1)) * ((3.456e-4 + S2) * (3.456e-4 + S2));
    S2 = tmp2; S1 = tmp1; S0 = tmp;
    double am = S0 + S1 + S2 + S3;
    tmp1 = 0.0;
    tmp = -3.0 * s[0] + 13.0 * s[1] - 23.0 * s[2] + 25.0 * s[3]; tmp1 += S0 * tmp;
    tmp = s[1] - 5.0 * s[2] + 13.0 * s[3] + 3.0 * s[4]; tmp1 += S1 * tmp;
    tmp = -s[2] + 7.0 * s[3] + 7.0 * s[4] - s[5]; tmp1 += S2 * tmp;
    tmp = 3.0 * s[3] + 13.0 * s[4] - 5.0 * s[5] + s[6]; tmp1 += S3 * tmp;
    tmp1 /= (12.0 * am);
    return tmp1;
}

/* WENO7 M, src/OCFD_Schemes.cu:692-741 */
static double weno7_M(const double *s) {
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, S3 = 0.0, tmp, tmp1, tmp2;
    tmp = -2.0 * s[6] + 9.0 * s[5] - 18.0 * s[4] + 11.0 * s[3]; S0 += 960 * tmp * tmp;
    tmp1 = -6.0 * s[6] + 24.0 * s[5] - 30.0 * s[4] + 12.0 * s[3]; S0 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[6] + 18.0 * (s[5] - s[4]) + 6.0 * s[3]; S0 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S0 += (80.0) * tmp;

    tmp = s[5] - 6.0 * s[4] + 3.0 * s[3] + 2.0 * s[2]; S1 += 960 * tmp * tmp;
    tmp1 = 6.0 * s[4] - 12.0 * s[3] + 6.0 * s[2]; S1 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[5] + 18.0 * (s[4] - s[3]) + 6.0 * s[2]; S1 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S1 += (80.0) * tmp;

    tmp = -2.0 * s[4] - 3.0 * s[3] + 6.0 * s[2] - s[1]; S2 += 960 * tmp * tmp;
    tmp1 = 6.0 * s[3] - 12.0 * s[2] + 6.0 * s[1]; S2 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[4] + 18.0 * (s[3] - s[2]) + 6.0 * s[1]; S2 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S2 += (80.0) * tmp;

    tmp = -11.0 * s[3] + 18.0 * s[2] - 9.0 * s[1] + 2.0 * s[0]; S3 += 960 * tmp * tmp;
    tmp1 = 12.0 * s[3] - 30.0 * s[2] + 24.0 * s[1] - 6.0 * s[0]; S3 += (1040.0) * tmp1 * tmp1;
    tmp2 = -6.0 * s[3] + 18.0 * (s[2] - s[1]) + 6.0 * s[0]; S3 += (1043.0) * tmp2 * tmp2;
    tmp = tmp * tmp2; S3 += (80.0) * tmp;

    tmp = ((3.456e-4 + S1) * (3.456e-4 + S1)) * ((3.456e-4 + S2) * (3.456e-4 + S2)) * ((3.456e-4 + S3) * (3.456e-4 + S3));
    tmp1 = (12.0) * ((3.456e-4 + S0) * (3.456e-4 + S0)) * ((3.456e-4 + S2) * (3.456e-4 + S2)) * ((3.456e-4 + S3) * (3.456e-4 + S3));
    tmp2 = (18.0) * ((3.456e-4 + S0) * (3.456e-4 + S0)) * ((3.456e-4 + S1) * (3.456e-4 + S1)) * ((3.456e-4 + S3) * (3.456e-4 + S3));
    S3 = (4.0) * ((3.456e-4 + S0) * (3.456e-4 + S0)) * ((3.456e-4 + S1) * (3.456e-4 + S1)) * ((3.456e-4 + S2) * (3.456e-4 + S2));
    S2 = tmp2; S1 = tmp1; S0 = tmp;
    double am = S0 + S1 + S2 + S3;
    tmp1 = 0.0;
    tmp = -3.0 * s[6] + 13.0 * s[5] - 23.0 * s[4] + 25.0 * s[3]; tmp1 += S0 * tmp;
    tmp = s[5] - 5.0 * s[4] + 13.0 * s[3] + 3.0 * s[2]; tmp1 += S1 * tmp;
    tmp = -s[4] + 7.0 * s[3] + 7.0 * s[2] - s[1]; tmp1 += S2 * tmp;
    tmp = 3.0 * s[3] + 13.0 * s[2] - 5.0 * s[1] + s[0]; tmp1 += S3 * tmp;
    tmp1 /= (12.0 * am);
    return tmp1;
}

/* NND2, src/OCFD_Schemes.cu:800-837 (s points at stencil[1]) */
static double nnd2_P(const double *s) { return s[1] + 0.5 * minmod2(s[2] - s[1], s[1] - s[0]); }
static double nnd2_M(const double *s) { return s[1] + 0.5 * minmod2(s[0] - s[1], s[1] - s[2]); }

/* UP7, src/OCFD_Schemes.cu:865-931 */
static double up7_P(const double *s) {
    return (-3.0 * s[0] + 25.0 * s[1] - 101.0 * s[2] + 319.0 * s[3] + 214.0 * s[4]
            - 38.0 * s[5] + 4.0 * s[6]) / 420.0;
}
static double up7_M(const double *s) {
    return (-3.0 * s[7] + 25.0 * s[6] - 101.0 * s[5] + 319.0 * s[4] + 214.0 * s[3]
            - 38.0 * s[2] + 4.0 * s[1]) / 420.0;
}

/* OMP6 P, src/OCFD_Schemes.cu:936-1001 */
static double omp6_P(int flagomp, const double *s) {
    double m, n;
    if (flagomp == 1) { m = 0.001; n = 0.0; }
    else if (flagomp == 2) { m = 0.0; n = -1.0 / 140.0; }
    else { m = 0.015; n = 0.0; }
    double mid = 0.5 * (m + n); m = 0.5 * (m - n);
    mid = (60.0 * mid * s[7] +
           (1.0 - 60.0 * m - 360.0 * mid) * s[6] +
           (-8.0 + 360.0 * m + 900.0 * mid) * s[5] +
           (37.0 - 900.0 * m - 1200.0 * mid) * s[4] +
           (37.0 + 1200.0 * m + 900.0 * mid) * s[3] +
           (-8.0 - 900.0 * m - 360.0 * mid) * s[2] +
           (1.0 + 360.0 * m + 60.0 * mid) * s[1] -
           60.0 * m * s[0]) / 60.0;
    m = s[3] + minmod2((s[4] - s[3]), (s[3] - s[2]));
    if ((mid - s[3]) * (mid - m) >= 1.e-10) {
        double tmp, tmp1;
        m = s[2] + s[4] - 2.0 * s[3];
        n = s[3] + s[5] - 2.0 * s[4];
        tmp = 4.0 * m - n;
        tmp1 = 4.0 * n - m;
        tmp = 0.5 * (s[3] + s[4]) - 0.5 * minmod4(tmp, tmp1, n, m);
        n = s[1] + s[3] - 2.0 * s[2];
        tmp1 = s[3] + 0.5 * (s[3] - s[2]) + 4.0 * minmod4(4 * n - m, 4 * m - n, m, n) / 3.0;
        {
            m = fmin(s[3], s[4]);
            m = fmin(m, tmp);
            n = s[3] + 4.0 * (s[3] - s[2]);
            n = fmin(s[3], n);
            n = fmin(n, tmp1);
            m = fmax(m, n);
        }
        {
            tmp = fmax(s[3], tmp);
            tmp = fmax(s[4], tmp);
            n = s[3] + 4.0 * (s[3] - s[2]);
            n = fmax(s[3], n);
            n = fmax(n, tmp1);
            n = fmin(tmp, n);
        }
        mid = mid + minmod2(n - mid, m - mid);
    }
    return mid;
}

/* OMP6 M, src/OCFD_Schemes.cu:1029-1092 */
static double omp6_M(int flagomp, const double *s) {
    double m, n;
    if (flagomp == 1) { m = 0.001; n = 0.0; }
    else if (flagomp == 2) { m = 0.0; n = -1.0 / 140.0; }
    else { m = 0.015; n = 0.0; }
    double mid = 0.5 * (m + n); m = 0.5 * (m - n);
    mid = (60.0 * mid * s[0] +
           (1.0 - 60.0 * m - 360.0 * mid) * s[1] +
           (-8.0 + 360.0 * m + 900.0 * mid) * s[2] +
           (37.0 - 900.0 * m - 1200.0 * mid) * s[3] +
           (37.0 + 1200.0 * m + 900.0 * mid) * s[4] +
           (-8.0 - 900.0 * m - 360.0 * mid) * s[5] +
           (1.0 + 360.0 * m + 60.0 * mid) * s[6] -
           60.0 * m * s[7]) / 60.0;
    m = s[4] + minmod2((s[3] - s[4]), (s[4] - s[5]));
    if ((mid - s[4]) * (mid - m) >= 1.e-10) {
        double tmp, tmp1;
        m = s[5] + s[3] - 2.0 * s[4];
        n = s[4] + s[2] - 2.0 * s[3];
        tmp = 4.0 * m - n;
        tmp1 = 4.0 * n - m;
        tmp = 0.5 * (s[4] + s[3]) - 0.5 * minmod4(tmp, tmp1, n, m);
        n = s[6] + s[4] - 2.0 * s[5];
        tmp1 = s[4] + 0.5 * (s[4] - s[3]) + 4.0 * minmod4(4 * n - m, 4 * m - n, m, n) / 3.0;
        {
            m = fmin(s[5], s[3]);
            m = fmin(m, tmp);
            n = s[4] + 4.0 * (s[4] - s[5]);
            n = fmin(s[4], n);
            n = fmin(n, tmp1);
            m = fmax(m, n);
        }
        {
            tmp = fmax(s[4], tmp);
            tmp = fmax(s[3], tmp);
            n = s[4] + 4.0 * (s[4] - s[5]);
            n = fmax(s[4], n);
            n = fmax(n, tmp1);
            n = fmin(tmp, n);
        }
        mid = mid + minmod2(n - mid, m - mid);
    }
    return mid;
}

/* WENO5 (Jiang-Shu), src/OCFD_Schemes.cu:1118-1210 */
static double weno5_P(const double *s) {
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, tmp, ep = 1e-6;
    tmp = s[2] - 2.0 * s[3] + s[4]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[3] + s[4]; S0 += 3 * tmp * tmp;
    double q03 = (2.0 * s[2] + 5.0 * s[3] - s[4]);
    tmp = s[1] - 2.0 * s[2] + s[3]; S1 += 13 * tmp * tmp;
    tmp = s[1] - s[3]; S1 += 3 * tmp * tmp;
    double q13 = (-s[1] + 5.0 * s[2] + 2.0 * s[3]);
    tmp = s[0] - 2.0 * s[1] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[0] - 4.0 * s[1] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    double q23 = (2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2]);
    double a0 = 3.0 * ((12.0 * ep + S1) * (12.0 * ep + S1)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    double a1 = 6.0 * ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    double a2 = ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S1) * (12.0 * ep + S1));
    return (a0 * q03 + a1 * q13 + a2 * q23) / (6.0 * (a0 + a1 + a2));
}
static double weno5_M(const double *s) {
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, tmp, ep = 1e-6;
    tmp = s[2] - 2.0 * s[1] + s[0]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[1] + s[0]; S0 += 3 * tmp * tmp;
    double q03 = (2.0 * s[2] + 5.0 * s[1] - s[0]);
    tmp = s[3] - 2.0 * s[2] + s[1]; S1 += 13 * tmp * tmp;
    tmp = s[3] - s[1]; S1 += 3 * tmp * tmp;
    double q13 = (-s[3] + 5.0 * s[2] + 2.0 * s[1]);
    tmp = s[4] - 2.0 * s[3] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[4] - 4.0 * s[3] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    double q23 = (2.0 * s[4] - 7.0 * s[3] + 11.0 * s[2]);
    double a0 = 3.0 * ((12.0 * ep + S1) * (12.0 * ep + S1)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    double a1 = 6.0 * ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S2) * (12.0 * ep + S2));
    double a2 = ((12.0 * ep + S0) * (12.0 * ep + S0)) * ((12.0 * ep + S1) * (12.0 * ep + S1));
    return (a0 * q03 + a1 * q13 + a2 * q23) / (6.0 * (a0 + a1 + a2));
}

/* One-sided WENO5 subsets for the closures, src/OCFD_bound_Scheme.cu:557-735 */
static double weno5_P_left(const double *s) {
    double S0 = 0.0, tmp, ep = 1e-6;
    tmp = s[2] - 2.0 * s[3] + s[4]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[3] + s[4]; S0 += 3 * tmp * tmp;
    double a0 = 1.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    double q03 = (2.0 * s[2] + 5.0 * s[3] - s[4]);
    return a0 * q03 / (6.0 * a0);
}
static double weno5_M_left(const double *s) {
    double S2 = 0.0, tmp, ep = 1e-6;
    tmp = s[4] - 2.0 * s[3] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[4] - 4.0 * s[3] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    double a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    double q23 = (2.0 * s[4] - 7.0 * s[3] + 11.0 * s[2]);
    return a2 * q23 / (6.0 * a2);
}
static double weno5_P_right_plus(const double *s) {
    double S1 = 0.0, S2 = 0.0, tmp, ep = 1e-6;
    tmp = s[1] - 2.0 * s[2] + s[3]; S1 += 13 * tmp * tmp;
    tmp = s[1] - s[3]; S1 += 3 * tmp * tmp;
    double a1 = 6.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    double q13 = (-s[1] + 5.0 * s[2] + 2.0 * s[3]);
    tmp = s[0] - 2.0 * s[1] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[0] - 4.0 * s[1] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    double a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    double q23 = (2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2]);
    return (a1 * q13 + a2 * q23) / (6.0 * (a1 + a2));
}
static double weno5_P_left_plus(const double *s) {
    double S0 = 0.0, S1 = 0.0, tmp, ep = 1e-6;
    tmp = s[2] - 2.0 * s[3] + s[4]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[3] + s[4]; S0 += 3 * tmp * tmp;
    double a0 = 3.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    double q03 = (2.0 * s[2] + 5.0 * s[3] - s[4]);
    tmp = s[1] - 2.0 * s[2] + s[3]; S1 += 13 * tmp * tmp;
    tmp = s[1] - s[3]; S1 += 3 * tmp * tmp;
    double a1 = 6.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    double q13 = (-s[1] + 5.0 * s[2] + 2.0 * s[3]);
    return (a0 * q03 + a1 * q13) / (6.0 * (a0 + a1));
}
static double weno5_M_right_plus(const double *s) {
    double S0 = 0.0, S1 = 0.0, tmp, ep = 1e-6;
    tmp = s[2] - 2.0 * s[1] + s[0]; S0 += 13 * tmp * tmp;
    tmp = 3.0 * s[2] - 4.0 * s[1] + s[0]; S0 += 3 * tmp * tmp;
    double a0 = 3.0 / ((12.0 * ep + S0) * (12.0 * ep + S0));
    double q03 = (2.0 * s[2] + 5.0 * s[1] - s[0]);
    tmp = s[3] - 2.0 * s[2] + s[1]; S1 += 13 * tmp * tmp;
    tmp = s[3] - s[1]; S1 += 3 * tmp * tmp;
    double a1 = 6.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    double q13 = (-s[3] + 5.0 * s[2] + 2.0 * s[1]);
    return (a0 * q03 + a1 * q13) / (6.0 * (a0 + a1));
}
static double weno5_M_left_plus(const double *s) {
    double S1 = 0.0, S2 = 0.0, tmp, ep = 1e-6;
    tmp = s[3] - 2.0 * s[2] + s[1]; S1 += 13 * tmp * tmp;
    tmp = s[3] - s[1]; S1 += 3 * tmp * tmp;
    double a1 = 6.0 / ((12.0 * ep + S1) * (12.0 * ep + S1));
    double q13 = (-s[3] + 5.0 * s[2] + 2.0 * s[1]);
    tmp = s[4] - 2.0 * s[3] + s[2]; S2 += 13 * tmp * tmp;
    tmp = s[4] - 4.0 * s[3] + 3.0 * s[2]; S2 += 3 * tmp * tmp;
    double a2 = 1.0 / ((12.0 * ep + S2) * (12.0 * ep + S2));
    double q23 = (2.0 * s[4] - 7.0 * s[3] + 11.0 * s[2]);
    return (a1 * q13 + a2 * q23) / (6.0 * (a1 + a2));
}

/* Stencil extent (ka1, kb1) of each launcher's kernel,
 * src/OCFD_Schemes.cu and src/OCFD_Schemes_hybrid_auto.cu (ia1/ib1). */
static void scheme_extent(int scheme, int side, int *ka1, int *kb1) {
    switch (scheme) {
    case OCFD_WENO7: *ka1 = -3; *kb1 = 3; break;
    case OCFD_WENO5: case OCFD_NND2: *ka1 = -2; *kb1 = 2; break;
    default: /* UP7, SYMBO, OMP6, hybrid: 8 points */
        if (side == 0) { *ka1 = -3; *kb1 = 4; } else { *ka1 = -4; *kb1 = 3; }
    }
}

/* Interior half-node flux of `scheme` on the gathered stencil s (indexing as
 * the reference kernel's stencil[]), for side 0 = P, 1 = M. */
static double scheme_flux(int scheme, int side, int hyflag, const double *s,
                          double tvlim, double tvmax) {
    if (side == 0) {
        switch (scheme) {
        case OCFD_UP7: return up7_P(s);
        case OCFD_WENO5: return weno5_P(s);
        case OCFD_WENO7: return weno7_P(s);
        case OCFD_WENO7_SYMBO: return weno7_symbo_P(0, s, tvlim, tvmax);
        case OCFD_WENO7_SYMBO_LIM: return weno7_symbo_P(1, s, tvlim, tvmax);
        case OCFD_NND2: return nnd2_P(&s[1]);
        case OCFD_OMP6_HR: return omp6_P(0, s);
        case OCFD_OMP6_LD: return omp6_P(1, s);
        case OCFD_OMP6_CD8: return omp6_P(2, s);
        case OCFD_HYBRIDAUTO: /* src/OCFD_Schemes_hybrid_auto.cu:911-919 */
            if (hyflag == 1) return up7_P(s);
            else if (hyflag == 2) return weno7_P(s);
            else return weno5_P(&s[1]);
        }
    } else {
        switch (scheme) {
        case OCFD_UP7: return up7_M(s);
        case OCFD_WENO5: return weno5_M(s);
        case OCFD_WENO7: return weno7_M(s);
        case OCFD_WENO7_SYMBO: return weno7_symbo_M(0, s, tvlim, tvmax);
        case OCFD_WENO7_SYMBO_LIM: return weno7_symbo_M(1, s, tvlim, tvmax);
        case OCFD_NND2: return nnd2_M(&s[1]);
        case OCFD_OMP6_HR: return omp6_M(0, s);
        case OCFD_OMP6_LD: return omp6_M(1, s);
        case OCFD_OMP6_CD8: return omp6_M(2, s);
        case OCFD_HYBRIDAUTO: /* src/OCFD_Schemes_hybrid_auto.cu:1008-1016 */
            if (hyflag == 1) return up7_M(s);
            else if (hyflag == 2) return weno7_M(&s[1]);
            else return weno5_M(&s[2]);
        }
    }
    return 0.0;
}

/* Closure of the P flux, src/OCFD_bound_Scheme.cu:739-819.  c = coords along
 * the sweep (face j+1/2 has c=j+1), r = c + start - end = j - n.  Returns 0
 * and sets *h when a closure applies. */
static int closure_P(int bl, int br, int c, int r, const double *s, int ka1, int kb1, double *h) {
    if (bl && c <= -ka1) {
        if (c == 0) *h = s[-ka1 + 1];
        if (c == 1) *h = weno5_P_left(&s[-ka1 - 2]);
        if (c == 2) *h = weno5_P_left_plus(&s[-ka1 - 2]);
        if (c == 3) *h = weno5_P(&s[-ka1 - 2]);
        return 0;
    }
    if (br && r >= -kb1) {
        if (r == -1) *h = s[-ka1] + 0.5 * minmod2(s[-ka1] - s[-ka1 - 1], s[-ka1] - s[-ka1 - 1]);
        if (r == -2) *h = weno5_P_right_plus(&s[-ka1 - 2]);
        if (r <= -3) *h = weno5_P(&s[-ka1 - 2]);
        return 0;
    }
    return 1;
}

/* Closure of the M flux, src/OCFD_bound_Scheme.cu:855-944 */
static int closure_M(int bl, int br, int c, int r, const double *s, int ka1, int kb1, double *h) {
    if (bl && c < -ka1) {
        if (c == 0) *h = weno5_M_left(&s[-ka1 - 2]);
        if (c == 1) *h = weno5_M_left_plus(&s[-ka1 - 2]);
        if (c >= 2) *h = weno5_M(&s[-ka1 - 2]);
        return 0;
    }
    if (br && r >= -kb1 - 1) {
        if (r == -1) *h = s[-ka1 - 1];
        if (r == -2) *h = s[-ka1] - 0.5 * minmod2(s[-ka1] - s[-ka1 - 1], s[-ka1] - s[-ka1 - 1]);
        if (r == -3) *h = weno5_M_right_plus(&s[-ka1 - 2]);
        if (r == -4) *h = weno5_M(&s[-ka1 - 2]);
        return 0;
    }
    return 1;
}

/* The flux derivative along one line, restating one launch of the
 * OCFD_d?1 (P) or OCFD_d?2 (M) launchers (src/OCFD_Schemes_Choose.cu:88-765)
 * with get_data_kernel (src/OCFD_Schemes.cu:159-290) and put_du_p/m_kernel
 * (src/OCFD_Schemes.cu:293-381).  f(q) for q in [-4, n+3] is at
 * f[(q+LAP)*fs]; du(i) at du[i*ds]; ajac(i) at ajac[(i+LAP)*as]; flags(j+1)
 * for face j+1/2 at flags[(j+1)*gs]. */
static void line_flux(int scheme, int side, int n, const double *f, long fs,
                      const double *ajac, long as, double h, int bl, int br,
                      int nonref, const int *flags, long gs, double *du, long ds,
                      double tvlim, double tvmax) {
    int ka1, kb1;
    scheme_extent(scheme, side, &ka1, &kb1);
    double *hf = (double *)malloc(sizeof(double) * (n + 1));
    for (int j = -1; j <= n - 1; j++) {
        double s[8];
        int center = (side == 0) ? j : j + 1; /* node at stencil[-ka1] */
        for (int k = 0; k <= kb1 - ka1; k++) {
            int q = center + ka1 + k;
            s[k] = f[(long)(q + LAP) * fs];
        }
        int c = j + 1, r = j - n;
        double hv = 0.0;
        int flag = side == 0 ? closure_P(bl, br, c, r, s, ka1, kb1, &hv)
                             : closure_M(bl, br, c, r, s, ka1, kb1, &hv);
        if (flag != 0) {
            int hy = flags ? flags[(long)(j + 1) * gs] : 0;
            hv = scheme_flux(scheme, side, hy, s, tvlim, tvmax);
        }
        hf[j + 1] = hv;
    }
    for (int i = 0; i < n; i++) {
        double tr = hf[i + 1], tl = hf[i];
        int zero = nonref && ((side == 0 && i == 0) || (side == 1 && i == n - 1));
        if (zero) du[(long)i * ds] += 0;
        else du[(long)i * ds] += -ajac[(long)(i + LAP) * as] * (tr - tl) / h;
    }
    free(hf);
}

/* ---- characteristic-space reconstruction (CHARTERIC = 1) ------------- */

/* Face state of the projection, get_para_charteric_p/m_kernel
 * (src/OCFD_flux_charteric.cu:321-520): means of u, v, w, c and the cofactor
 * row over the face's two nodes; n = A/|A|; l a unit vector normal to n
 * (no z component when |n3| <= |n2|, else no y component); m = n x l;
 * KK = (gamma-1)/c^2.  s0, s1 = u v w c a1 a2 a3 of the two nodes.
 * P = u v w c n1 n2 n3 l1 l2 l3 m1 m2 m3 KK. */
static void char_para(const double *s0, const double *s1, double gamma, double *P) {
    double u1 = (s0[0] + s1[0]) * 0.5, v1 = (s0[1] + s1[1]) * 0.5, w1 = (s0[2] + s1[2]) * 0.5;
    double c1 = (s0[3] + s1[3]) * 0.5;
    double a1 = (s0[4] + s1[4]) * 0.5, a2 = (s0[5] + s1[5]) * 0.5, a3 = (s0[6] + s1[6]) * 0.5;
    double ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
    double n1 = a1 / ss, n2 = a2 / ss, n3 = a3 / ss, l1, l2, l3;
    if (fabs(n3) <= fabs(n2)) {
        ss = sqrt(n1 * n1 + n2 * n2);
        l1 = -n2 / ss; l2 = n1 / ss; l3 = 0.0;
    } else {
        ss = sqrt(n1 * n1 + n3 * n3);
        l1 = -n3 / ss; l2 = 0.0; l3 = n1 / ss;
    }
    P[0] = u1; P[1] = v1; P[2] = w1; P[3] = c1;
    P[4] = n1; P[5] = n2; P[6] = n3; P[7] = l1; P[8] = l2; P[9] = l3;
    P[10] = n2 * l3 - n3 * l2; P[11] = n3 * l1 - n1 * l3; P[12] = n1 * l2 - n2 * l1;
    P[13] = (gamma - 1.0) / (c1 * c1);
}

/* Left (L, flux_charteric_ptoc_kernel, :617-685) and right (R,
 * flux_charteric_ctop_kernel, :687-752) eigenvector matrices, row-major. */
static void char_matrices(const double *P, double *L, double *R) {
    double u1 = P[0], v1 = P[1], w1 = P[2], c1 = P[3], n1 = P[4], n2 = P[5], n3 = P[6];
    double l1 = P[7], l2 = P[8], l3 = P[9], m1 = P[10], m2 = P[11], m3 = P[12], KK = P[13];
    double X1 = 1.0 / (2.0 * c1);
    double un = u1 * n1 + v1 * n2 + w1 * n3;
    double v2 = (u1 * u1 + v1 * v1 + w1 * w1) * 0.5;
    L[0] = 1.0 - KK * v2; L[1] = KK * u1; L[2] = KK * v1; L[3] = KK * w1; L[4] = -KK;
    L[5] = -(u1 * l1 + v1 * l2 + w1 * l3); L[6] = l1; L[7] = l2; L[8] = l3; L[9] = 0.0;
    L[10] = -(u1 * m1 + v1 * m2 + w1 * m3);
    L[11] = n2 * l3 - n3 * l2; L[12] = n3 * l1 - n1 * l3; L[13] = n1 * l2 - n2 * l1; L[14] = 0.0;
    L[15] = KK * v2 * 0.5 + X1 * un;
    L[16] = -X1 * n1 - KK * u1 * 0.5; L[17] = -X1 * n2 - KK * v1 * 0.5; L[18] = -X1 * n3 - KK * w1 * 0.5;
    L[19] = KK * 0.5;
    L[20] = L[15] - 2 * X1 * un;
    L[21] = 2 * X1 * n1 + L[16]; L[22] = 2 * X1 * n2 + L[17]; L[23] = 2 * X1 * n3 + L[18]; L[24] = L[19];
    double H = v2 + 1.0 / KK;
    R[0] = 1.0; R[1] = 0.0; R[2] = 0.0; R[3] = 1.0; R[4] = 1.0;
    R[5] = u1; R[6] = l1; R[7] = n2 * l3 - n3 * l2; R[8] = u1 - c1 * n1; R[9] = 2 * u1 - R[8];
    R[10] = v1; R[11] = l2; R[12] = n3 * l1 - n1 * l3; R[13] = v1 - c1 * n2; R[14] = 2 * v1 - R[13];
    R[15] = w1; R[16] = l3; R[17] = n1 * l2 - n2 * l1; R[18] = w1 - c1 * n3; R[19] = 2 * w1 - R[18];
    R[20] = v2; R[21] = u1 * l1 + v1 * l2 + w1 * l3; R[22] = u1 * m1 + v1 * m2 + w1 * m3;
    R[23] = H - c1 * (u1 * n1 + v1 * n2 + w1 * n3); R[24] = 2 * H - R[23];
}

static void matvec5(const double *M, const double *x, double *y) {
    for (int i = 0; i < 5; i++)
        y[i] = M[5 * i] * x[0] + M[5 * i + 1] * x[1] + M[5 * i + 2] * x[2] + M[5 * i + 3] * x[3] + M[5 * i + 4] * x[4];
}

/* Flux derivative of all five variables along one line in characteristic
 * space: OCFD_weno7_SYMBO_character_P/M_kernel and
 * OCFD_HybridAuto_character_P/M_Jameson_kernel (src/OCFD_flux_charteric.cu:
 * 754-908, 1000-1268).  Per face j+1/2 the eight split-flux vectors of the
 * stencil are projected with L of the face, each characteristic variable is
 * reconstructed (with the closures of the physical path), and R maps back;
 * a hybrid face with flag 1 stays in physical space (UP7).  du: the
 * reference's put_du_character_p/m_kernel (:16-119): with a closure on one
 * face only, the first (P, low closure) or last (high closure) node takes a
 * physical-space difference of the split flux (OCFD_bound_character_kernel_p
 * /m, :121-319) instead of the face difference, the non-reflecting zero
 * applies only then, and the low-side M special is never taken in y (its
 * test compares the y flag with the z value).  F[m][(q+LAP)*fs] = split flux
 * of variable m at node q; prim[0..6] = u v w c a1 a2 a3 at the same place. */
static void line_flux_char(int scheme, int side, int dir, int n, const double *const *F, long fs,
                           const double *const *prim, const double *ajac, double h, int bl, int br, int nonref,
                           const int *flags, long gs, double gamma, double *const *du, long ds, double tvlim,
                           double tvmax) {
    int ka1, kb1;
    scheme_extent(scheme, side, &ka1, &kb1);
    double *hf = (double *)malloc(sizeof(double) * 5 * (n + 1));
    for (int j = -1; j <= n - 1; j++) {
        double s[5][8], x[5], y[5], L[25], R[25], P[14];
        int center = (side == 0) ? j : j + 1;
        for (int m = 0; m < 5; m++)
            for (int k = 0; k <= kb1 - ka1; k++) s[m][k] = F[m][(long)(center + ka1 + k + LAP) * fs];
        int hy = flags ? flags[(long)(j + 1) * gs] : 0;
        int proj = !(scheme == OCFD_HYBRIDAUTO && hy == 1);
        if (proj) {
            double s0[7], s1[7];
            for (int t = 0; t < 7; t++) {
                s0[t] = prim[t][(long)(j + LAP) * fs];
                s1[t] = prim[t][(long)(j + 1 + LAP) * fs];
            }
            char_para(s0, s1, gamma, P);
            char_matrices(P, L, R);
            for (int k = 0; k <= kb1 - ka1; k++) {
                for (int m = 0; m < 5; m++) x[m] = s[m][k];
                matvec5(L, x, y);
                for (int m = 0; m < 5; m++) s[m][k] = y[m];
            }
        }
        int c = j + 1, r = j - n;
        for (int m = 0; m < 5; m++) {
            double hv = 0.0;
            int flag = side == 0 ? closure_P(bl, br, c, r, s[m], ka1, kb1, &hv)
                                 : closure_M(bl, br, c, r, s[m], ka1, kb1, &hv);
            if (flag != 0) hv = scheme_flux(scheme, side, hy, s[m], tvlim, tvmax);
            x[m] = hv;
        }
        if (proj) matvec5(R, x, y);
        else for (int m = 0; m < 5; m++) y[m] = x[m];
        for (int m = 0; m < 5; m++) hf[m * (n + 1) + j + 1] = y[m];
    }
    const int fy = bl && br ? 7 : bl ? 1 : br ? 4 : 0;
#define FQ(m, q) F[m][(long)((q) + LAP) * fs]
    for (int i = 0; i < n; i++) {
        double aj = ajac[(long)(i + LAP) * fs];
        for (int m = 0; m < 5; m++) {
            double t;
            int zero = 0;
            if (side == 0) {
                if (fy == 1 && i == 0) {
                    zero = nonref;
                    t = FQ(m, 1) - FQ(m, 0);
                } else if (fy == 4 && i == n - 1) {
                    double a = FQ(m, n - 1), b = FQ(m, n - 2), cc = FQ(m, n - 3);
                    t = (a + 0.5 * minmod2(a - b, a - b)) - (b + 0.5 * minmod2(a - b, b - cc));
                } else {
                    t = hf[m * (n + 1) + i + 1] - hf[m * (n + 1) + i];
                }
            } else {
                if (fy == 4 && i == n - 1) {
                    zero = nonref;
                    t = FQ(m, n - 1) - FQ(m, n - 2);
                } else if (fy == 1 && i == 0 && dir != 1) {
                    double a = FQ(m, 0), b = FQ(m, 1), cc = FQ(m, 2);
                    t = (b - 0.5 * minmod2(cc - b, b - a)) - (a - 0.5 * minmod2(b - a, b - a));
                } else {
                    t = hf[m * (n + 1) + i + 1] - hf[m * (n + 1) + i];
                }
            }
            if (zero) du[m][(long)i * ds] += 0;
            else du[m][(long)i * ds] += -aj * t / h;
        }
    }
#undef FQ
    free(hf);
}

/* Generic central derivative with D0 closures, OCFD_CD6_kernel/OCFD_CD8_kernel
 * (src/OCFD_Schemes.cu:96-156) + OCFD_D0bound_scheme_kernel
 * (src/OCFD_bound_Scheme.cu:79-186) + put_d0_kernel (src/OCFD_Schemes.cu:
 * 73-94).  f(q) at f[(q+LAP)*fs]; out(i) at out[i*os]. */
static void line_central(int svis, int n, const double *f, long fs, double h,
                         int bl, int br, double *out, long os) {
    int ka1 = (svis == OCFD_CD6) ? -3 : -4;
    for (int i = 0; i < n; i++) {
        double s[9];
        for (int k = 0; k <= -2 * ka1; k++) s[k] = f[(long)(i + ka1 + k + LAP) * fs];
        double tmp = 0.0;
        int done = 0;
        const double *c = &s[-ka1];
        if (bl) {
            if (i == 0) { tmp = (c[1] - c[0]); done = 1; }
            else if (i == 1) { tmp = (c[1] - c[-1]) * 0.5; done = 1; }
            else if (i == 2) { tmp = (c[-2] - 8.0 * c[-1] + 8.0 * c[1] - c[2]) / 12.0; done = 1; }
            else if (i == 3) { tmp = (c[3] - c[-3] - 9.0 * (c[2] - c[-2]) + 45.0 * (c[1] - c[-1])) / 60.0; done = 1; }
        }
        if (!done && br) {
            int t1 = i - n;
            if (t1 == -1) { tmp = (c[0] - c[-1]); done = 1; }
            else if (t1 == -2) { tmp = (c[1] - c[-1]) * 0.5; done = 1; }
            else if (t1 == -3) { tmp = (c[-2] - 8.0 * c[-1] + 8.0 * c[1] - c[2]) / 12.0; done = 1; }
            else if (t1 == -4) { tmp = (c[3] - c[-3] - 9.0 * (c[2] - c[-2]) + 45.0 * (c[1] - c[-1])) / 60.0; done = 1; }
        }
        if (!done) {
            if (svis == OCFD_CD6)
                tmp = (45.0 * (s[4] - s[2]) - 9.0 * (s[5] - s[1]) + (s[6] - s[0])) / 60.0;
            else
                tmp = (672.0 * (s[5] - s[3]) - 168.0 * (s[6] - s[2]) + 32.0 * (s[7] - s[1]) - 3.0 * (s[8] - s[0])) / 840.0;
        }
        out[(long)i * os] = tmp / h;
    }
}

/* ======================================================================
 * Public unit-level helpers
 * ====================================================================== */

double orc_face_flux(int scheme, int side, int hyflag, const double *s8, double tvlim, double tvmax) {
    return scheme_flux(scheme, side, hyflag, s8, tvlim, tvmax);
}

void orc_line_flux_derivative(int scheme, int side, int n, const double *f, const double *ajac, double h,
                              int bound_l, int bound_r, int nonref, const int *flags, double *du) {
    line_flux(scheme, side, n, f, 1, ajac, 1, h, bound_l, bound_r, nonref, flags, 1, du, 1, 5.0, 0.2);
}

void orc_line_central(int svis, int n, const double *f, double h, int bound_l, int bound_r, double *out) {
    line_central(svis, n, f, 1, h, bound_l, bound_r, out, 1);
}

double orc_jameson_dp(double pc, double pxm, double pxp, double pym, double pyp, double pzm, double pzp) {
    double dp0 = fabs(-pxm + 2 * pc - pxp) / (pxm + 2 * pc + pxp);
    dp0 += fabs(-pym + 2 * pc - pyp) / (pym + 2 * pc + pyp);
    dp0 += fabs(-pzm + 2 * pc - pzp) / (pzm + 2 * pc + pzp);
    return dp0;
}

/* Stager_Warming_ker, src/OCFD_split.cu:92-142 */
static void sw_point(double tmp0d, double u, double v, double w, double cc, double A1, double A2, double A3,
                     double gamma, double eps, double C1, double C3, double *fp, double *fm) {
    double ss, E1, E2, E3, E1P, E2P, E3P, E1M, E2M, E3M;
    double uc1, uc2, vc1, vc2, wc1, wc2, vvc1, vvc2, vv, W2;
    ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
    E1 = A1 * u + A2 * v + A3 * w;
    E2 = E1 - cc * ss;
    E3 = E1 + cc * ss;
    ss = 1.0 / ss;
    A1 *= ss; A2 *= ss; A3 *= ss;
    E1P = (E1 + sqrt(E1 * E1 + eps * eps)) * 0.50;
    E2P = (E2 + sqrt(E2 * E2 + eps * eps)) * 0.50;
    E3P = (E3 + sqrt(E3 * E3 + eps * eps)) * 0.50;
    E1M = E1 - E1P; E2M = E2 - E2P; E3M = E3 - E3P;
    uc1 = u - cc * A1; uc2 = u + cc * A1;
    vc1 = v - cc * A2; vc2 = v + cc * A2;
    wc1 = w - cc * A3; wc2 = w + cc * A3;
    vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.50;
    vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.50;
    vv = (gamma - 1.0) * (u * u + v * v + w * w);
    W2 = C3 * cc * cc;
    fp[0] = tmp0d * (C1 * E1P + E2P + E3P);
    fp[1] = tmp0d * (C1 * E1P * u + E2P * uc1 + E3P * uc2);
    fp[2] = tmp0d * (C1 * E1P * v + E2P * vc1 + E3P * vc2);
    fp[3] = tmp0d * (C1 * E1P * w + E2P * wc1 + E3P * wc2);
    fp[4] = tmp0d * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
    fm[0] = tmp0d * (C1 * E1M + E2M + E3M);
    fm[1] = tmp0d * (C1 * E1M * u + E2M * uc1 + E3M * uc2);
    fm[2] = tmp0d * (C1 * E1M * v + E2M * vc1 + E3M * vc2);
    fm[3] = tmp0d * (C1 * E1M * w + E2M * wc1 + E3M * wc2);
    fm[4] = tmp0d * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
}

void orc_split_point(double d, double u, double v, double w, double cc, double A1, double A2, double A3,
                     double gamma, double eps, double *out10) {
    double tmp0 = 1.0 / (2.0 * gamma);
    double C1 = 2.0 * (gamma - 1.0);
    double C3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    sw_point(d * tmp0, u, v, w, cc, A1, A2, A3, gamma, eps, C1, C3, out10, out10 + 5);
}

/* ======================================================================
 * Context
 * ====================================================================== */

static double *zalloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

orc *orc_create(const ocfd_config *cfg) {
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->c = *cfg;
    o->nx = cfg->nx; o->ny = cfg->ny; o->nz = cfg->nz;
    o->NX2 = o->nx + 2 * LAP; o->NY2 = o->ny + 2 * LAP; o->NZ2 = o->nz + 2 * LAP;
    o->nlap = (long)o->NX2 * o->NY2 * o->NZ2;
    o->nint = (long)o->nx * o->ny * o->nz;
    /* src/OCFD_init.cu:29-31 */
    o->hx = 1.0 / (o->nx - 1.0);
    o->hy = 1.0 / (o->ny - 1.0);
    o->hz = 1.0 / (o->nz - 1.0);
    double g = cfg->gamma, Ma = cfg->ama;
    double Tref = cfg->t_ref <= 0 ? 288.150 : cfg->t_ref; /* src/parameters.c:521 */
    /* opencfd_para_init, src/OCFD_init.cu:381-410 */
    o->Cv = 1.0 / (g * (g - 1.0) * Ma * Ma);
    o->Cp = 1.0 / ((g - 1.0) * Ma * Ma);
    o->Tsb = 110.4 / Tref;
    o->amu_C0 = 1.0 / cfg->re * (1.0 + 110.4 / Tref);
    o->tmp0 = 1.0 / (2.0 * g);
    o->split_C1 = 2.0 * (g - 1.0);
    o->split_C3 = (3.0 - g) / (2.0 * (g - 1.0));
    o->vis_c = o->Cp / cfg->pr;

    o->d = zalloc(o->nlap); o->u = zalloc(o->nlap); o->v = zalloc(o->nlap);
    o->w = zalloc(o->nlap); o->T = zalloc(o->nlap); o->P = zalloc(o->nlap);
    o->cc = zalloc(o->nlap);
    for (int k = 0; k < 13; k++) {
        o->A[k] = zalloc(o->nlap);
        for (long i = 0; i < o->nlap; i++) o->A[k][i] = 1.0; /* src/OCFD_Comput_Jacobian3d.cu:34-62 */
    }
    for (int k = 0; k < 3; k++) { o->fp[k] = zalloc(NV * o->nlap); o->fm[k] = zalloc(NV * o->nlap); }
    for (int k = 0; k < 4; k++) o->Ev[k] = zalloc(o->nlap);
    o->f_lap = zalloc(NV * o->nlap);
    o->f = zalloc(NV * o->nint); o->fn = zalloc(NV * o->nint); o->du = zalloc(NV * o->nint);
    o->Amu = zalloc(o->nint);
    for (int k = 0; k < 12; k++) o->grad[k] = zalloc(o->nint);
    for (int k = 0; k < 4; k++) o->visd[k] = zalloc(o->nint);
    o->flag[0] = (int *)calloc((size_t)(o->nx + 1) * o->ny * o->nz, sizeof(int));
    o->flag[1] = (int *)calloc((size_t)o->nx * (o->ny + 1) * o->nz, sizeof(int));
    o->flag[2] = (int *)calloc((size_t)o->nx * o->ny * (o->nz + 1), sizeof(int));
    o->ub1 = zalloc(4L * o->ny);
    o->pfx = zalloc(o->nx);
    o->pgz = zalloc(o->nz);
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    free(o->d); free(o->u); free(o->v); free(o->w); free(o->T); free(o->P); free(o->cc);
    for (int k = 0; k < 13; k++) free(o->A[k]);
    for (int k = 0; k < 3; k++) { free(o->fp[k]); free(o->fm[k]); free(o->flag[k]); }
    for (int k = 0; k < 4; k++) { free(o->Ev[k]); free(o->visd[k]); }
    for (int k = 0; k < 12; k++) free(o->grad[k]);
    free(o->f_lap); free(o->f); free(o->fn); free(o->du); free(o->Amu);
    free(o->ub1); free(o->pfx); free(o->pgz);
    free(o->in124); free(o->up124); free(o->coeff124); free(o->udist124);
    free(o->in124_user); free(o->up124_user);
    free(o);
}

/* ---------------------------------------------------------------------
 * Periodic face fill of one LAP field: exchange_boundary_{x,y,z}_packed_dev
 * at one rank (src/OCFD_mpi_dev.cu:186-310).  Only the 4-layer faces with the
 * other two indices in the interior are written; edges and corners are not.
 * --------------------------------------------------------------------- */
static void fill_x(orc *o, double *a) {
    for (int z = LAP; z < o->nz + LAP; z++)
        for (int y = LAP; y < o->ny + LAP; y++)
            for (int i = 0; i < LAP; i++) {
                a[IL(o, o->nx + LAP + i, y, z)] = a[IL(o, LAP + i, y, z)];
            }
    for (int z = LAP; z < o->nz + LAP; z++)
        for (int y = LAP; y < o->ny + LAP; y++)
            for (int i = 0; i < LAP; i++) a[IL(o, i, y, z)] = a[IL(o, o->nx + i, y, z)];
}
static void fill_y(orc *o, double *a) {
    for (int z = LAP; z < o->nz + LAP; z++)
        for (int i = 0; i < LAP; i++)
            for (int x = LAP; x < o->nx + LAP; x++) a[IL(o, x, o->ny + LAP + i, z)] = a[IL(o, x, LAP + i, z)];
    for (int z = LAP; z < o->nz + LAP; z++)
        for (int i = 0; i < LAP; i++)
            for (int x = LAP; x < o->nx + LAP; x++) a[IL(o, x, i, z)] = a[IL(o, x, o->ny + i, z)];
}
static void fill_z(orc *o, double *a) {
    for (int i = 0; i < LAP; i++)
        for (int y = LAP; y < o->ny + LAP; y++)
            for (int x = LAP; x < o->nx + LAP; x++) a[IL(o, x, y, o->nz + LAP + i)] = a[IL(o, x, y, LAP + i)];
    for (int i = 0; i < LAP; i++)
        for (int y = LAP; y < o->ny + LAP; y++)
            for (int x = LAP; x < o->nx + LAP; x++) a[IL(o, x, y, i)] = a[IL(o, x, y, o->nz + i)];
}
static void fill_dir(orc *o, double *a, int dir) {
    if (!o->c.iperiodic[dir]) return;
    if (dir == 0) fill_x(o, a); else if (dir == 1) fill_y(o, a); else fill_z(o, a);
}
/* exchange_boundary_xyz_packed_dev, src/OCFD_mpi_dev.cu:122-127 */
static void fill_xyz(orc *o, double *a) { fill_dir(o, a, 0); fill_dir(o, a, 1); fill_dir(o, a, 2); }

/* ---------------------------------------------------------------------
 * Metrics: comput_Jac3d (src/OCFD_Comput_Jacobian3d.cu:250-424)
 * --------------------------------------------------------------------- */
/* OCFD_d?0_CD6_kernel, src/OCFD_Schemes.cu:1236-1300, then the closures
 * OCFD_D?0_bound_kernel_m/p, src/OCFD_bound_Scheme.cu:189-349. */
static void metric_derivative(orc *o, const double *a, int dir, int bound, double *out) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int X = x + LAP, Y = y + LAP, Z = z + LAP;
                double r;
                if (dir == 0) {
                    double m3 = a[IL(o, X - 3, Y, Z)], m2 = a[IL(o, X - 2, Y, Z)], m1 = a[IL(o, X - 1, Y, Z)];
                    double p1 = a[IL(o, X + 1, Y, Z)], p2 = a[IL(o, X + 2, Y, Z)], p3 = a[IL(o, X + 3, Y, Z)];
                    r = (45.0 * (p1 - m1) - 9.0 * (p2 - m2) + (p3 - m3)) / (60.0 * o->hx);
                } else if (dir == 1) {
                    double m3 = a[IL(o, X, Y - 3, Z)], m2 = a[IL(o, X, Y - 2, Z)], m1 = a[IL(o, X, Y - 1, Z)];
                    double p1 = a[IL(o, X, Y + 1, Z)], p2 = a[IL(o, X, Y + 2, Z)], p3 = a[IL(o, X, Y + 3, Z)];
                    r = (45.0 * (p1 - m1) - 9.0 * (p2 - m2) + (p3 - m3)) / (60.0 * o->hy);
                } else {
                    double m3 = a[IL(o, X, Y, Z - 3)], m2 = a[IL(o, X, Y, Z - 2)], m1 = a[IL(o, X, Y, Z - 1)];
                    double p1 = a[IL(o, X, Y, Z + 1)], p2 = a[IL(o, X, Y, Z + 2)], p3 = a[IL(o, X, Y, Z + 3)];
                    r = (45.0 * (p1 - m1) - 9.0 * (p2 - m2) + (p3 - m3)) / (60.0 * o->hz);
                }
                out[I0(o, x, y, z)] = r;
            }
    if (!bound) return;
    int n = dir == 0 ? nx : dir == 1 ? ny : nz;
    if (n < LAP) return; /* CHECK_SIZE, src/OCFD_bound_Scheme.cu:27-41 aborts */
    double h = dir == 0 ? o->hx : dir == 1 ? o->hy : o->hz;
    int na = dir == 0 ? ny : nx, nb = dir == 2 ? ny : nz;
    for (int b = 0; b < nb; b++)
        for (int aa = 0; aa < na; aa++) {
            /* F(q) = field at interior index q along dir */
#define F(q) (dir == 0 ? a[IL(o, (q) + LAP, aa + LAP, b + LAP)] : dir == 1 ? a[IL(o, aa + LAP, (q) + LAP, b + LAP)] : a[IL(o, aa + LAP, b + LAP, (q) + LAP)])
#define O(q) out[dir == 0 ? I0(o, (q), aa, b) : dir == 1 ? I0(o, aa, (q), b) : I0(o, aa, b, (q))]
            O(0) = (F(1) - F(0)) / h;
            O(1) = (F(2) - F(0)) * 0.5 / h;
            O(2) = (F(0) - 8.0 * F(1) + 8.0 * F(3) - F(4)) / (12.0 * h);
            O(3) = (F(6) - F(0) - 9.0 * (F(5) - F(1)) + 45.0 * (F(4) - F(2))) / (60.0 * h);
            int t = n - 1;
            if (dir == 1) {
                /* OCFD_Dy0_bound_kernel_p, src/OCFD_bound_Scheme.cu:253-271 */
                O(n - 1) = -(F(t - 1) - F(t)) / h;
                O(n - 2) = -(F(t - 2) - F(t)) * 0.5 / h;
                O(n - 3) = -(F(t) - 8.0 * F(t - 1) + 8.0 * F(t - 3) - F(t - 4)) / (12.0 * h);
                O(n - 4) = -(F(t - 6) - F(t) - 9.0 * (F(t - 5) - F(t - 1)) + 45.0 * (F(t - 4) - F(t - 2))) / (60.0 * h);
            } else {
                /* OCFD_Dx0_bound_kernel_p / OCFD_Dz0_bound_kernel_p */
                O(n - 1) = (F(t) - F(t - 1)) / h;
                O(n - 2) = (F(t) - F(t - 2)) * 0.5 / h;
                O(n - 3) = (F(t - 4) - 8.0 * F(t - 3) + 8.0 * F(t - 1) - F(t)) / (12.0 * h);
                O(n - 4) = (F(t) - F(t - 6) - 9.0 * (F(t - 1) - F(t - 5)) + 45.0 * (F(t - 2) - F(t - 4))) / (60.0 * h);
            }
#undef F
#undef O
        }
}

static void mirror_y(orc *o, double *a, double sign);
static void ibc124_mirror_metrics(orc *o);

int orc_load_mesh(orc *o, const double *x, const double *y, const double *z) {
    const double *src[3] = {x, y, z};
    for (int k = 0; k < 3; k++) {
        for (int zz = 0; zz < o->nz; zz++)
            for (int yy = 0; yy < o->ny; yy++)
                for (int xx = 0; xx < o->nx; xx++)
                    o->A[k][IL(o, xx + LAP, yy + LAP, zz + LAP)] = src[k][I0(o, xx, yy, zz)];
        /* exchange_boundary_xyz on the host copy, src/OCFD_Comput_Jacobian3d.cu:139-141 */
        fill_xyz(o, o->A[k]);
    }
    /* boundary_Jac3d_Axx (src/OCFD_Comput_Jacobian3d.cu:592-617, IBC 124) */
    if (o->c.ibc == 124 && o->c.bc_npara[0] == 1) {
        mirror_y(o, o->A[AXX], 1.0); mirror_y(o, o->A[AYY], -1.0); mirror_y(o, o->A[AZZ], 1.0);
    }
    /* comput_Jac3d, src/OCFD_Comput_Jacobian3d.cu:393-424 */
    long N = o->nint;
    double *D = zalloc(9 * N);
    double *xi = D, *xj = D + N, *xk = D + 2 * N, *yi = D + 3 * N, *yj = D + 4 * N, *yk = D + 5 * N;
    double *zi = D + 6 * N, *zj = D + 7 * N, *zk = D + 8 * N;
    metric_derivative(o, o->A[AXX], 0, o->c.jac_bound[0], xi);
    metric_derivative(o, o->A[AYY], 0, o->c.jac_bound[0], yi);
    metric_derivative(o, o->A[AZZ], 0, o->c.jac_bound[0], zi);
    metric_derivative(o, o->A[AXX], 1, o->c.jac_bound[1], xj);
    metric_derivative(o, o->A[AYY], 1, o->c.jac_bound[1], yj);
    metric_derivative(o, o->A[AZZ], 1, o->c.jac_bound[1], zj);
    metric_derivative(o, o->A[AXX], 2, o->c.jac_bound[2], xk);
    metric_derivative(o, o->A[AYY], 2, o->c.jac_bound[2], yk);
    metric_derivative(o, o->A[AZZ], 2, o->c.jac_bound[2], zk);
    /* comput_Jac3d_kernal, src/OCFD_Comput_Jacobian3d.cu:250-391 */
    int nx = o->nx, ny = o->ny, nz = o->nz;
    for (int zz = 0; zz < nz; zz++)
        for (int yy = 0; yy < ny; yy++)
            for (int xx = 0; xx < nx; xx++) {
                long q = I0(o, xx, yy, zz);
                double xi1 = xi[q], xj1 = xj[q], xk1 = xk[q], yi1 = yi[q], yj1 = yj[q], yk1 = yk[q];
                double zi1 = zi[q], zj1 = zj[q], zk1 = zk[q];
                double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 - xi1 * zj1 * yk1);
                long L = IL(o, xx + LAP, yy + LAP, zz + LAP);
                o->A[AJAC][L] = Jac1;
                o->A[AKX][L] = (yj1 * zk1 - zj1 * yk1);
                o->A[AKY][L] = (zj1 * xk1 - xj1 * zk1);
                o->A[AKZ][L] = (xj1 * yk1 - yj1 * xk1);
                o->A[AIX][L] = (yk1 * zi1 - zk1 * yi1);
                o->A[AIY][L] = (zk1 * xi1 - xk1 * zi1);
                o->A[AIZ][L] = (xk1 * yi1 - yk1 * xi1);
                o->A[ASX][L] = (yi1 * zj1 - zi1 * yj1);
                o->A[ASY][L] = (zi1 * xj1 - xi1 * zj1);
                o->A[ASZ][L] = (xi1 * yj1 - yi1 * xj1);
            }
    /* zeroth-order ghost copies (faces), src/OCFD_Comput_Jacobian3d.cu:300-389 */
    for (int k = AKX; k <= AJAC; k++) {
        double *a = o->A[k];
        for (int zz = 0; zz < nz; zz++)
            for (int yy = 0; yy < ny; yy++) {
                for (int i = 0; i < LAP; i++) a[IL(o, i, yy + LAP, zz + LAP)] = a[IL(o, LAP, yy + LAP, zz + LAP)];
                for (int i = 1; i <= LAP; i++) a[IL(o, nx - 1 + LAP + i, yy + LAP, zz + LAP)] = a[IL(o, nx - 1 + LAP, yy + LAP, zz + LAP)];
            }
        for (int zz = 0; zz < nz; zz++)
            for (int xx = 0; xx < nx; xx++) {
                for (int i = 0; i < LAP; i++) a[IL(o, xx + LAP, i, zz + LAP)] = a[IL(o, xx + LAP, LAP, zz + LAP)];
                for (int i = 1; i <= LAP; i++) a[IL(o, xx + LAP, ny - 1 + LAP + i, zz + LAP)] = a[IL(o, xx + LAP, ny - 1 + LAP, zz + LAP)];
            }
        for (int yy = 0; yy < ny; yy++)
            for (int xx = 0; xx < nx; xx++) {
                for (int i = 0; i < LAP; i++) a[IL(o, xx + LAP, yy + LAP, i)] = a[IL(o, xx + LAP, yy + LAP, LAP)];
                for (int i = 1; i <= LAP; i++) a[IL(o, xx + LAP, yy + LAP, nz - 1 + LAP + i)] = a[IL(o, xx + LAP, yy + LAP, nz - 1 + LAP)];
            }
        /* src/OCFD_Comput_Jacobian3d.cu:225-234 */
        fill_xyz(o, a);
    }
    ibc124_mirror_metrics(o);
    free(D);
    return 0;
}

int orc_load_jacobi(orc *o, const double *a13) {
    /* src/OCFD_Comput_Jacobian3d.cu:148-215: interior from the file, ghosts keep
     * the initial 1.0 except the periodic faces. */
    for (int k = 0; k < 13; k++) {
        const double *s = a13 + (long)k * o->nint;
        for (int zz = 0; zz < o->nz; zz++)
            for (int yy = 0; yy < o->ny; yy++)
                for (int xx = 0; xx < o->nx; xx++)
                    o->A[k][IL(o, xx + LAP, yy + LAP, zz + LAP)] = s[I0(o, xx, yy, zz)];
        fill_xyz(o, o->A[k]);
    }
    ibc124_mirror_metrics(o); /* src/OCFD_init.cu:116-122 */
    return 0;
}

int orc_load_flow(orc *o, const double *d, const double *u, const double *v, const double *w, const double *T,
                  int istep, double tt) {
    const double *src[5] = {d, u, v, w, T};
    double *dst[5] = {o->d, o->u, o->v, o->w, o->T};
    for (int k = 0; k < 5; k++) {
        memset(dst[k], 0, sizeof(double) * o->nlap);
        for (int zz = 0; zz < o->nz; zz++)
            for (int yy = 0; yy < o->ny; yy++)
                for (int xx = 0; xx < o->nx; xx++)
                    dst[k][IL(o, xx + LAP, yy + LAP, zz + LAP)] = src[k][I0(o, xx, yy, zz)];
    }
    o->istep = istep;
    o->tt = tt;
    return 0;
}

int orc_set_inlet_profile(orc *o, const double *prof) {
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < o->ny; j++) o->ub1[k * o->ny + j] = prof[k * o->ny + j];
    o->have_profile = 1;
    return 0;
}

/* pri_to_cons_kernel over a job (interior indices), src/commen_kernel.cu:32-64 */
static void pri_to_cons(orc *o, int x0, int x1, int y0, int y1, int z0, int z1) {
    long N = o->nint;
    for (int z = z0; z < z1; z++)
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) {
                long L = IL(o, x + LAP, y + LAP, z + LAP), q = I0(o, x, y, z);
                double d = o->d[L], u = o->u[L], v = o->v[L], w = o->w[L], T = o->T[L];
                o->f[q] = d;
                o->f[q + N] = u * d;
                o->f[q + 2 * N] = v * d;
                o->f[q + 3 * N] = w * d;
                o->f[q + 4 * N] = d * ((u * u + v * v + w * w) * 0.5 + o->Cv * T);
            }
}

/* IBC 108 set-up, bc_user_Compression_conner_init, src/OCFD_boundary_init.c:
 * 376-516 (uses glibc rand() from its default seed as the reference does). */
static void ibc108_init(orc *o) {
    int MZMAX = o->c.bc_npara[0], MTMAX = o->c.bc_npara[1], INLET = o->c.bc_npara[2];
    double XDB = o->c.bc_rpara[2], XDE = o->c.bc_rpara[3], XUB = o->c.bc_rpara[6], SLZ = o->c.bc_rpara[7];
    int ny = o->ny, nx = o->nx, nz = o->nz;
    if (!(INLET == 1 && o->have_profile)) {
        if (INLET != 1)
            for (int j = 0; j < ny; j++) {
                o->ub1[j] = 1.; o->ub1[ny + j] = 1.; o->ub1[2 * ny + j] = 0.; o->ub1[3 * ny + j] = 1.;
            }
    }
    double *Axx = o->A[AXX], *Azz = o->A[AZZ];
    o->x_begin = 0;
    double *xa = zalloc(nx), *za = zalloc(nz);
    for (int i = 0; i < nx; i++) {
        if (Axx[IL(o, i + LAP, ny + LAP - 1, LAP)] <= XUB) o->x_begin = i;
        for (int k = 0; k < nz; k++) {
            xa[i] = Axx[IL(o, i + LAP, LAP, LAP)];
            za[k] = Azz[IL(o, LAP, LAP, k + LAP)];
        }
    }
    /* get_xs_blow_suction_multiwave, src/OCFD_boundary_init.c:446-498 */
    srand(1);
    int MZ1 = abs(MZMAX);
    double faiz[64], zl[64], ztmp = 0.;
    for (int k = 0; k < MZ1; k++) {
        faiz[k] = rand() / (double)RAND_MAX;
        if (k == 0) zl[k] = 1.; else zl[k] = zl[k - 1] / 1.25;
        ztmp = ztmp + zl[k];
    }
    for (int k = 0; k < MZ1; k++) zl[k] = zl[k] / ztmp;
    const double PI = 3.141592653589793;
    for (int i = 0; i < nx; i++) {
        if (xa[i] >= XDB && xa[i] <= XDE) {
            double seta = 10. * PI * (xa[i] - XDB) / (XDB - XDE);
            o->pfx[i] = 4. / sqrt(27.) * sin(seta) * (1. - cos(seta));
        } else o->pfx[i] = 0.;
    }
    for (int k = 0; k < nz; k++) {
        o->pgz[k] = 0.;
        double seta = za[k] / SLZ;
        if (MZMAX > 0) {
            for (int m = 0; m < MZMAX; m++) o->pgz[k] = o->pgz[k] + zl[m] * sin(2. * PI * (m + 1) * (seta + faiz[m]));
        } else if (MZMAX == 0) o->pgz[k] = 1.;
        else o->pgz[k] = sin(-2. * PI * MZMAX * seta);
    }
    /* get_fait_multifrequancy, src/OCFD_boundary_init.c:501-516 */
    double Ttmp = 0.;
    for (int k = 0; k < MTMAX; k++) {
        o->fait[k] = rand() / (double)RAND_MAX;
        o->TM[0] = 1.;
        if (k > 0) o->TM[k] = o->TM[k - 1] / 1.25;
        Ttmp = Ttmp + o->TM[k];
    }
    for (int k = 0; k < MTMAX; k++) o->TM[k] = o->TM[k] / Ttmp;
    free(xa); free(za);
}

/* bc_user_Compression_conner, src/OCFD_boundary_compression_conner.cu:27-179 */
static void ibc108(orc *o) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    double *d = o->d, *u = o->u, *v = o->v, *w = o->w, *T = o->T;
    const double *ub1 = o->ub1;
    /* do_ub1_inlet_kernel: i = 0..LAP at every interior (y,z) */
    for (int z = LAP; z < nz + LAP; z++)
        for (int y = LAP; y < ny + LAP; y++) {
            int yl = y - LAP;
            for (int i = 0; i <= LAP; i++) {
                long L = IL(o, i, y, z);
                d[L] = ub1[yl]; u[L] = ub1[yl + ny]; v[L] = ub1[yl + 2 * ny]; w[L] = 0.; T[L] = ub1[yl + 3 * ny];
            }
        }
    /* do_ub1_top_kernel */
    for (int z = LAP; z < nz + LAP; z++)
        for (int x = o->x_begin + LAP; x < nx + LAP; x++) {
            long L = IL(o, x, ny + LAP - 1, z);
            d[L] = ub1[ny - 1]; u[L] = ub1[ny + ny - 1]; v[L] = ub1[2 * ny + ny - 1]; w[L] = 0.; T[L] = ub1[3 * ny + ny - 1];
        }
    double ht = 0.;
    double BETA = o->c.bc_rpara[4];
    int MTMAX = o->c.bc_npara[1];
    if (BETA > 0.) {
        for (int m = 0; m < MTMAX; m++) ht = ht + o->TM[m] * sin((m + 1) * BETA * o->tt);
    } else ht = 1.;
    double EPSL = o->c.bc_rpara[1], TW = o->c.bc_rpara[0];
    /* do_wall_dist_kernel */
    for (int z = LAP; z < nz + LAP; z++)
        for (int x = LAP; x < nx + LAP; x++) {
            long L = IL(o, x, LAP, z);
            u[L] = 0.;
            v[L] = EPSL * o->pfx[x - LAP] * o->pgz[z - LAP] * ht;
            w[L] = 0.;
        }
    if (o->c.bc_npara[3] == 0) {
        if (TW > 0) {
            /* do_wall_Tp1_kernel */
            for (int z = LAP; z < nz + LAP; z++)
                for (int x = LAP; x < nx + LAP; x++) {
                    double pw = (18. * d[IL(o, x, LAP + 1, z)] * T[IL(o, x, LAP + 1, z)]
                                 - 9. * d[IL(o, x, LAP + 2, z)] * T[IL(o, x, LAP + 2, z)]
                                 + 2. * d[IL(o, x, LAP + 3, z)] * T[IL(o, x, LAP + 3, z)]) / 11.;
                    T[IL(o, x, LAP, z)] = TW;
                    d[IL(o, x, LAP, z)] = pw / T[IL(o, x, LAP, z)];
                    for (int i = 0; i < LAP; i++) {
                        d[IL(o, x, i, z)] = d[IL(o, x, 2 * LAP - i, z)];
                        u[IL(o, x, i, z)] = -u[IL(o, x, 2 * LAP - i, z)];
                        v[IL(o, x, i, z)] = -v[IL(o, x, 2 * LAP - i, z)];
                        w[IL(o, x, i, z)] = -w[IL(o, x, 2 * LAP - i, z)];
                        T[IL(o, x, i, z)] = T[IL(o, x, 2 * LAP - i, z)];
                    }
                }
        } else {
            /* do_wall_Tp2_kernel */
            for (int z = LAP; z < nz + LAP; z++)
                for (int x = LAP; x < nx + LAP; x++) {
                    T[IL(o, x, LAP, z)] = (18. * T[IL(o, x, LAP + 1, z)] - 9. * T[IL(o, x, LAP + 2, z)] + 2. * T[IL(o, x, LAP + 3, z)]) / 11.;
                    double pw = (18. * d[IL(o, x, LAP + 1, z)] * T[IL(o, x, LAP + 1, z)]
                                 - 9. * d[IL(o, x, LAP + 2, z)] * T[IL(o, x, LAP + 2, z)]
                                 + 2. * d[IL(o, x, LAP + 3, z)] * T[IL(o, x, LAP + 3, z)]) / 11.;
                    d[IL(o, x, LAP, z)] = pw / T[IL(o, x, LAP, z)];
                }
        }
    }
}

/* y-ghost mirror about j = 0 and j = ny-1 over the interior x and z,
 * boundary_Jac3d_kernal_y_l/r (src/OCFD_Comput_Jacobian3d.cu:429-448) and
 * do_symmetry_kernel_m/p (src/OCFD_boundary_Liftbody3D.cu:84-115). */
static void mirror_y(orc *o, double *a, double sign) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    for (int z = LAP; z < nz + LAP; z++)
        for (int x = LAP; x < nx + LAP; x++) {
            for (int y = 0; y < LAP; y++) a[IL(o, x, y, z)] = sign * a[IL(o, x, 2 * LAP - y, z)];
            for (int y = ny + LAP; y < ny + 2 * LAP; y++) a[IL(o, x, y, z)] = sign * a[IL(o, x, 2 * (ny + LAP - 1) - y, z)];
        }
}

/* boundary_Jac3d_Liftbody_Ajac (src/OCFD_Comput_Jacobian3d.cu:619-662) */
static void ibc124_mirror_metrics(orc *o) {
    if (o->c.ibc != 124 || o->c.bc_npara[0] != 1) return;
    static const double sign[10] = {1, -1, 1, -1, 1, -1, 1, -1, 1, 1};
    for (int k = AKX; k <= AJAC; k++) mirror_y(o, o->A[k], sign[k - AKX]);
}

int orc_set_bc_tables(orc *o, const double *inlet5, const double *upper5) {
    long nyz = (long)o->ny * o->nz, nxy = (long)o->nx * o->ny;
    if (inlet5) {
        free(o->in124_user);
        o->in124_user = (double *)malloc(sizeof(double) * 5 * nyz);
        memcpy(o->in124_user, inlet5, sizeof(double) * 5 * nyz);
    }
    if (upper5) {
        free(o->up124_user);
        o->up124_user = (double *)malloc(sizeof(double) * 5 * nxy);
        memcpy(o->up124_user, upper5, sizeof(double) * 5 * nxy);
    }
    return 0;
}

/* bc_user_Liftbody3d_init (src/OCFD_boundary_init.c:95-318), parameters of
 * bc_parameter (:50-71); rand() from glibc's default seed (the reference
 * seeds it with the clock). */
static int ibc124_init(orc *o) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long nyz = (long)ny * nz, nxy = (long)nx * ny;
    int withleading = o->c.bc_npara[1], upper_mode = o->c.bc_npara[2];
    int MZMAX = o->c.bc_npara[3], MTMAX = o->c.bc_npara[4];
    double EPSL_WALL = o->c.bc_rpara[2], EPSL_UPPER = o->c.bc_rpara[3];
    double DB = o->c.bc_rpara[5], DE = o->c.bc_rpara[6];
    const double PI = 3.141592653589793;
    if (!o->in124) {
        o->in124 = zalloc(5 * nyz);
        o->up124 = zalloc(5 * nxy);
        o->coeff124 = zalloc(3 * nxy);
        o->udist124 = zalloc(nxy);
    }
    double *d[5] = {o->d, o->u, o->v, o->w, o->T};
    if (withleading == 1) {
        if (!o->in124_user) return -1;
        memcpy(o->in124, o->in124_user, sizeof(double) * 5 * nyz);
    } else {
        for (int m = 0; m < 5; m++)
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++) o->in124[m * nyz + (long)k * ny + j] = d[m][IL(o, LAP, j + LAP, k + LAP)];
    }
    if (upper_mode == 1) {
        if (!o->up124_user) return -1;
        memcpy(o->up124, o->up124_user, sizeof(double) * 5 * nxy);
    }
    const double *X = o->A[AXX], *Y = o->A[AYY], *Z = o->A[AZZ];
    srand(1);
    o->v_dist_need = 0;
    if (MZMAX != 0) {
        /* get_xy_blow_suction_multiwave (:320-372) */
        o->v_dist_need = 1;
        int MZ1 = abs(MZMAX);
        double *faiz = zalloc(MZ1 + 1), *zl = zalloc(MZ1 + 1), *fx = zalloc(nx), *gz = zalloc(ny);
        double ztmp = 0.;
        for (int k = 0; k < MZ1; k++) {
            faiz[k] = rand() / (double)RAND_MAX;
            if (k == 0) zl[k] = 1.; else zl[k] = zl[k - 1] / 1.25;
            ztmp = ztmp + zl[k];
        }
        for (int k = 0; k < MZ1; k++) zl[k] = zl[k] / ztmp;
        for (int i = 0; i < nx; i++) {
            double xa = X[IL(o, i + LAP, LAP, LAP)];
            if (xa >= DB && xa <= DE) {
                double seta = 10. * PI * (xa - DB) / (DB - DE);
                fx[i] = 4. / sqrt(27.) * sin(seta) * (1. - cos(seta));
            } else fx[i] = 0.;
        }
        for (int j = 0; j < ny; j++) {
            gz[j] = 0.;
            double seta = ((double)j) / ny;
            if (MZMAX > 0) {
                for (int m = 0; m < MZMAX; m++) gz[j] = gz[j] + zl[m] * sin(2. * PI * (m + 1) * (seta + faiz[m]));
            } else {
                gz[j] = sin(-2. * PI * MZMAX * seta);
            }
        }
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double vdw = fx[i] * gz[j];
                long q = (long)j * nx + i;
                o->coeff124[q] = 0;
                o->coeff124[nxy + q] = EPSL_WALL * vdw * sin(2 * PI / ny * j);
                o->coeff124[2 * nxy + q] = EPSL_WALL * vdw * cos(2 * PI / ny * j);
            }
        free(faiz); free(zl); free(fx); free(gz);
    } else {
        double *tv = zalloc(nxy), *tu = zalloc(nxy);
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double r = (rand() / (double)RAND_MAX - 0.5) * 2.0;
                tv[(long)j * nx + i] = EPSL_WALL * r;
                r = (rand() / (double)RAND_MAX - 0.5) * 2.0;
                tu[(long)j * nx + i] = EPSL_UPPER * r;
            }
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long q = (long)j * nx + i;
                long L0 = IL(o, i + LAP, j + LAP, LAP), L1 = IL(o, i + LAP, j + LAP, LAP + 1);
                double vdw = 0.0;
                if (X[L0] >= DB && X[L0] <= DE) {
                    o->v_dist_need = 1;
                    vdw = tv[q];
                }
                o->udist124[q] = tu[q];
                double xn = X[L1] - X[L0], yn = Y[L1] - Y[L0], zn = Z[L1] - Z[L0];
                double sn = sqrt(xn * xn + yn * yn + zn * zn);
                o->coeff124[q] = vdw * xn / sn;
                o->coeff124[nxy + q] = vdw * yn / sn;
                o->coeff124[2 * nxy + q] = vdw * zn / sn;
            }
        free(tv); free(tu);
    }
    /* get_fait_multifrequancy (:501-517) */
    double Ttmp = 0.;
    for (int k = 0; k < MTMAX && k < 64; k++) {
        o->fait[k] = rand() / (double)RAND_MAX;
        o->TM[0] = 1.;
        if (k > 0) o->TM[k] = o->TM[k - 1] / 1.25;
        Ttmp = Ttmp + o->TM[k];
    }
    for (int k = 0; k < MTMAX && k < 64; k++) o->TM[k] = o->TM[k] / Ttmp;
    return 0;
}

/* bc_user_Liftbody3d (src/OCFD_boundary_Liftbody3D.cu:185-294) */
static void ibc124(orc *o) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long nyz = (long)ny * nz, nxy = (long)nx * ny;
    double *d = o->d, *u = o->u, *v = o->v, *w = o->w, *T = o->T;
    /* do_u2d_inlet_kernel: i = 0..LAP */
    for (int z = LAP; z < nz + LAP; z++)
        for (int y = LAP; y < ny + LAP; y++) {
            long q = (long)(z - LAP) * ny + (y - LAP);
            for (int i = 0; i <= LAP; i++) {
                long L = IL(o, i, y, z);
                d[L] = o->in124[q]; u[L] = o->in124[nyz + q]; v[L] = o->in124[2 * nyz + q];
                w[L] = o->in124[3 * nyz + q]; T[L] = o->in124[4 * nyz + q];
            }
        }
    /* upper face k = nz-1 */
    double aoa = o->c.bc_rpara[0] * 3.141592653589793 / 180;
    double sin_aoa = sin(aoa), cos_aoa = cos(aoa);
    for (int y = LAP; y < ny + LAP; y++)
        for (int x = LAP; x < nx + LAP; x++) {
            long q = (long)(y - LAP) * nx + (x - LAP), L = IL(o, x, y, nz + LAP - 1);
            double dist = o->udist124[q];
            if (o->c.bc_npara[2] == 1) { /* do_u2d_upper_kernel */
                d[L] = o->up124[q]; u[L] = o->up124[nxy + q] + dist; v[L] = o->up124[2 * nxy + q];
                w[L] = o->up124[3 * nxy + q]; T[L] = o->up124[4 * nxy + q];
            } else { /* do_u_dist_upper_kernel */
                d[L] = 1.0; u[L] = cos_aoa + dist; v[L] = 0.0; w[L] = sin_aoa; T[L] = 1.0;
            }
        }
    /* wall k = 0 */
    double ht = 0.;
    double BETA = o->c.bc_rpara[4];
    int MTMAX = o->c.bc_npara[4];
    if (BETA > 0.) {
        for (int m = 0; m < MTMAX; m++) ht = ht + o->TM[m] * sin((m + 1) * BETA * o->tt);
    } else ht = 1.;
    double TW = o->c.bc_rpara[1];
    for (int y = LAP; y < ny + LAP; y++)
        for (int x = LAP; x < nx + LAP; x++) {
            long q = (long)(y - LAP) * nx + (x - LAP);
            long L0 = IL(o, x, y, LAP), L1 = IL(o, x, y, LAP + 1), L2 = IL(o, x, y, LAP + 2);
            if (o->v_dist_need) {
                u[L0] = o->coeff124[q] * ht;
                v[L0] = o->coeff124[nxy + q] * ht;
                w[L0] = o->coeff124[2 * nxy + q] * ht;
            } else {
                u[L0] = 0.0; v[L0] = 0.0; w[L0] = 0.0;
            }
            if (TW > 0.0) {
                T[L0] = TW;
                d[L0] = (4.0 * d[L1] * T[L1] - d[L2] * T[L2]) / (3.0 * TW);
            } else {
                T[L0] = (4.0 * T[L1] - T[L2]) / 3.0;
                d[L0] = (4.0 * d[L1] - d[L2]) / 3.0;
            }
        }
    /* symmetry planes */
    if (o->c.bc_npara[0] == 1) {
        mirror_y(o, d, 1.0); mirror_y(o, u, 1.0); mirror_y(o, v, -1.0); mirror_y(o, w, 1.0); mirror_y(o, T, 1.0);
    }
}

/* OCFD_bc, src/OCFD_boundary.cu:18-92 */
static void ocfd_bc(orc *o) {
    if (o->c.ibc == 108) ibc108(o);
    if (o->c.ibc == 124) ibc124(o);
    int nx = o->nx, ny = o->ny, nz = o->nz;
    pri_to_cons(o, 0, 1, 0, ny, 0, nz);
    pri_to_cons(o, nx - 1, nx, 0, ny, 0, nz);
    pri_to_cons(o, 0, nx, 0, 1, 0, nz);
    pri_to_cons(o, 0, nx, ny - 1, ny, 0, nz);
    pri_to_cons(o, 0, nx, 0, ny, 0, 1);
    pri_to_cons(o, 0, nx, 0, ny, nz - 1, nz);
}

/* get_Amu, src/commen_kernel.cu:122-140 */
static void get_amu(orc *o) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < o->nz; z++)
        for (int y = 0; y < o->ny; y++)
            for (int x = 0; x < o->nx; x++) {
                double tmp = o->T[IL(o, x + LAP, y + LAP, z + LAP)];
                o->Amu[I0(o, x, y, z)] = o->amu_C0 * sqrt(tmp * tmp * tmp) / (o->Tsb + tmp);
            }
}

/* get_duvwT / cons_to_pri_kernel, src/commen_kernel.cu:77-118 */
static void cons_to_pri(orc *o) {
    long N = o->nint;
    double g = o->c.gamma, Ma = o->c.ama;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < o->nz; z++)
        for (int y = 0; y < o->ny; y++)
            for (int x = 0; x < o->nx; x++) {
                long q = I0(o, x, y, z), L = IL(o, x + LAP, y + LAP, z + LAP);
                double d1, u1, v1, w1, T1, d2, T2;
                d1 = o->f[q];
                o->d[L] = d2 = d1;
                u1 = o->f[q + N]; u1 = u1 / d1; o->u[L] = u1;
                v1 = o->f[q + 2 * N]; v1 = v1 / d1; o->v[L] = v1;
                w1 = o->f[q + 3 * N]; w1 = w1 / d1; o->w[L] = w1;
                T1 = o->f[q + 4 * N];
                o->T[L] = T2 = (T1 - 0.5 * d1 * (u1 * u1 + v1 * v1 + w1 * w1)) / (d1 * o->Cv);
                o->P[L] = T2 * d2 / (g * Ma * Ma);
            }
}

int orc_init(orc *o) {
    /* pri_to_cons over the interior, src/OCFD_init.cu:99-110 */
    pri_to_cons(o, 0, o->nx, 0, o->ny, 0, o->nz);
    if (o->c.ibc == 108) ibc108_init(o);
    if (o->c.ibc == 124 && ibc124_init(o) != 0) return -1;
    /* NS_solver_real prologue, src/OCFD_NS_Solver.cu:46-54 */
    fill_xyz(o, o->d); fill_xyz(o, o->u); fill_xyz(o, o->v); fill_xyz(o, o->w); fill_xyz(o, o->T);
    ocfd_bc(o);
    get_amu(o);
    return 0;
}

/* Set_Scheme_HybridAuto style 2: Comput_P + Comput_Scheme_point_Jameson,
 * src/OCFD_Schemes_hybrid_auto.cu:21-68, 611-703 */
static void hybrid_sensor(orc *o) {
    double p00 = 1.0 / (o->c.gamma * o->c.ama * o->c.ama);
    for (long i = 0; i < o->nlap; i++) o->P[i] = p00 * o->d[i] * o->T[i];
    int nx = o->nx, ny = o->ny, nz = o->nz;
    double i1 = o->c.hy_dp_intv[0], i2 = o->c.hy_dp_intv[1];
    const double *P = o->P;
    for (int dir = 0; dir < 3; dir++) {
        int ex = nx + (dir == 0), ey = ny + (dir == 1), ez = nz + (dir == 2);
#pragma omp parallel for schedule(static)
        for (int z = 0; z < ez; z++)
            for (int y = 0; y < ey; y++)
                for (int x = 0; x < ex; x++) {
                    int X = x + LAP - (dir == 0), Y = y + LAP - (dir == 1), Z = z + LAP - (dir == 2);
                    double pc = P[IL(o, X, Y, Z)];
                    double dp0 = fabs(-P[IL(o, X - 1, Y, Z)] + 2 * pc - P[IL(o, X + 1, Y, Z)]) /
                                 (P[IL(o, X - 1, Y, Z)] + 2 * pc + P[IL(o, X + 1, Y, Z)]);
                    dp0 += fabs(-P[IL(o, X, Y - 1, Z)] + 2 * pc - P[IL(o, X, Y + 1, Z)]) /
                           (P[IL(o, X, Y - 1, Z)] + 2 * pc + P[IL(o, X, Y + 1, Z)]);
                    dp0 += fabs(-P[IL(o, X, Y, Z - 1)] + 2 * pc - P[IL(o, X, Y, Z + 1)]) /
                           (P[IL(o, X, Y, Z - 1)] + 2 * pc + P[IL(o, X, Y, Z + 1)]);
                    int kp = 1;
                    if (dp0 > i1) kp += 1;
                    if (dp0 > i2) kp += 1;
                    o->flag[dir][(long)x + (long)ex * ((long)y + (long)ey * z)] = kp;
                }
    }
}

/* du_invis_Jacobian3d (src/OCFD_Stream.cu:289-305): ghost fill, sound speed
 * (sound_speed_kernel, src/commen_kernel.cu:144-153, over the LAP box),
 * Stager_Warming over the LAP box (src/OCFD_split.cu:145-198), then for each
 * direction the P launcher (5 variables) and the M launcher (5 variables). */
static void du_invis(orc *o) {
    fill_xyz(o, o->d); fill_xyz(o, o->u); fill_xyz(o, o->v); fill_xyz(o, o->w); fill_xyz(o, o->T);
    double Ma = o->c.ama;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < o->nlap; i++) o->cc[i] = sqrt(o->T[i]) / Ma;
    const int ai[3][3] = {{AKX, AKY, AKZ}, {AIX, AIY, AIZ}, {ASX, ASY, ASZ}};
#pragma omp parallel for schedule(static)
    for (long i = 0; i < o->nlap; i++) {
        double tmp0d = o->d[i] * o->tmp0;
        for (int dir = 0; dir < 3; dir++) {
            double fp[5], fm[5];
            sw_point(tmp0d, o->u[i], o->v[i], o->w[i], o->cc[i], o->A[ai[dir][0]][i], o->A[ai[dir][1]][i],
                     o->A[ai[dir][2]][i], o->c.gamma, o->c.epsl_sw, o->split_C1, o->split_C3, fp, fm);
            for (int m = 0; m < 5; m++) {
                o->fp[dir][i + m * o->nlap] = fp[m];
                o->fm[dir][i + m * o->nlap] = fm[m];
            }
        }
    }
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long N = o->nint;
    for (int dir = 0; dir < 3; dir++) {
        int n = dir == 0 ? nx : dir == 1 ? ny : nz;
        long fs = dir == 0 ? 1 : dir == 1 ? o->NX2 : (long)o->NX2 * o->NY2;
        long ds = dir == 0 ? 1 : dir == 1 ? nx : (long)nx * ny;
        double h = dir == 0 ? o->hx : dir == 1 ? o->hy : o->hz;
        int bl = o->c.dif_bound[2 * dir], br = o->c.dif_bound[2 * dir + 1];
        int na = dir == 0 ? ny : nx, nb = dir == 2 ? ny : nz;
        int hyb = o->c.scheme_invis == OCFD_HYBRIDAUTO;
        int ex = nx + (dir == 0), ey = ny + (dir == 1);
        long gs = dir == 0 ? 1 : dir == 1 ? ex : (long)ex * ey;
        if (o->c.charteric) {
            for (int side = 0; side < 2; side++) {
                const double *F = side == 0 ? o->fp[dir] : o->fm[dir];
                int nonref = o->c.non_ref[2 * dir + side];
#pragma omp parallel for collapse(2) schedule(static)
                for (int b = 0; b < nb; b++)
                    for (int a = 0; a < na; a++) {
                        int X = dir == 0 ? 0 : a, Y = dir == 0 ? a : dir == 1 ? 0 : b, Z = dir == 2 ? 0 : b;
                        long base = IL(o, dir == 0 ? 0 : X + LAP, dir == 1 ? 0 : Y + LAP, dir == 2 ? 0 : Z + LAP);
                        const double *fl[5], *pr[7];
                        double *dul[5];
                        for (int m = 0; m < 5; m++) {
                            fl[m] = F + m * o->nlap + base;
                            dul[m] = o->du + m * N + I0(o, X, Y, Z);
                        }
                        pr[0] = o->u + base; pr[1] = o->v + base; pr[2] = o->w + base; pr[3] = o->cc + base;
                        for (int t = 0; t < 3; t++) pr[4 + t] = o->A[ai[dir][t]] + base;
                        const int *fl_h = NULL;
                        if (hyb) fl_h = o->flag[dir] + (long)X + (long)ex * ((long)Y + (long)ey * Z);
                        line_flux_char(o->c.scheme_invis, side, dir, n, fl, fs, pr, o->A[AJAC] + base, h, bl, br,
                                       nonref, fl_h, gs, o->c.gamma, dul, ds, 5.0, 0.2);
                    }
            }
            continue;
        }
        for (int side = 0; side < 2; side++) {
            const double *F = side == 0 ? o->fp[dir] : o->fm[dir];
            int nonref = o->c.non_ref[2 * dir + side];
            for (int m = 0; m < 5; m++)
#pragma omp parallel for collapse(2) schedule(static)
                for (int b = 0; b < nb; b++)
                    for (int a = 0; a < na; a++) {
                        int X = dir == 0 ? 0 : a, Y = dir == 0 ? a : dir == 1 ? 0 : b, Z = dir == 2 ? 0 : b;
                        const double *fl = F + m * o->nlap + IL(o, dir == 0 ? 0 : X + LAP, dir == 1 ? 0 : Y + LAP, dir == 2 ? 0 : Z + LAP);
                        const double *aj = o->A[AJAC] + IL(o, dir == 0 ? 0 : X + LAP, dir == 1 ? 0 : Y + LAP, dir == 2 ? 0 : Z + LAP);
                        double *dul = o->du + m * N + I0(o, X, Y, Z);
                        const int *fl_h = NULL;
                        if (hyb) fl_h = o->flag[dir] + (long)X + (long)ex * ((long)Y + (long)ey * Z);
                        line_flux(o->c.scheme_invis, side, n, fl, fs, aj, fs, h, bl, br, nonref, fl_h, gs, dul, ds,
                                  5.0, 0.2);
                    }
        }
    }
}

/* du_vis_Jacobian3d (src/OCFD_Stream.cu:307-335) with du_viscous_Jacobian3d_*
 * (src/OCFD_NS_Jacobian3d.cu:58-440) and YF_Pe_XF (src/commen_kernel.cu:158-169) */
/* Derivatives along xi/eta/zeta of the first nf of u,v,w,T with the viscous
 * scheme and its D0 closures: grad[3*f + dir] (OCFD_dx0/dy0/dz0,
 * src/OCFD_Schemes_Choose.cu:19-84). */
static void central_grads(orc *o, int nf) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const double *src4[4] = {o->u, o->v, o->w, o->T};
    int sv = o->c.scheme_vis;
    for (int fI = 0; fI < nf; fI++)
        for (int dir = 0; dir < 3; dir++) {
            int n = dir == 0 ? nx : dir == 1 ? ny : nz;
            long fs = dir == 0 ? 1 : dir == 1 ? o->NX2 : (long)o->NX2 * o->NY2;
            long os = dir == 0 ? 1 : dir == 1 ? nx : (long)nx * ny;
            double h = dir == 0 ? o->hx : dir == 1 ? o->hy : o->hz;
            int bl = o->c.dif_bound[2 * dir], br = o->c.dif_bound[2 * dir + 1];
            int na = dir == 0 ? ny : nx, nb = dir == 2 ? ny : nz;
#pragma omp parallel for collapse(2) schedule(static)
            for (int b = 0; b < nb; b++)
                for (int a = 0; a < na; a++) {
                    int X = dir == 0 ? 0 : a, Y = dir == 0 ? a : dir == 1 ? 0 : b, Z = dir == 2 ? 0 : b;
                    const double *fl = src4[fI] + IL(o, dir == 0 ? 0 : X + LAP, dir == 1 ? 0 : Y + LAP, dir == 2 ? 0 : Z + LAP);
                    line_central(sv, n, fl, fs, h, bl, br, o->grad[3 * fI + dir] + I0(o, X, Y, Z), os);
                }
        }
}

static void du_vis(orc *o) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long N = o->nint;
    int sv = o->c.scheme_vis;
    /* 12 gradients: grad[3*f + dir] for f in u,v,w,T */
    central_grads(o, 4);
    const double *uk = o->grad[0], *ui = o->grad[1], *us = o->grad[2];
    const double *vk = o->grad[3], *vi = o->grad[4], *vs = o->grad[5];
    const double *wk = o->grad[6], *wi = o->grad[7], *ws = o->grad[8];
    const double *Tk = o->grad[9], *Ti = o->grad[10], *Ts = o->grad[11];
    const int ai[3][3] = {{AKX, AKY, AKZ}, {AIX, AIY, AIZ}, {ASX, ASY, ASZ}};
    for (int dir = 0; dir < 3; dir++) {
        /* vis_flux_ker, src/OCFD_NS_Jacobian3d.cu:239-281 (interior only) */
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    long q = I0(o, x, y, z), L = IL(o, x + LAP, y + LAP, z + LAP);
                    double Akx = o->A[AKX][L], Aix = o->A[AIX][L], Asx = o->A[ASX][L];
                    double Aky = o->A[AKY][L], Aiy = o->A[AIY][L], Asy = o->A[ASY][L];
                    double Akz = o->A[AKZ][L], Aiz = o->A[AIZ][L], Asz = o->A[ASZ][L];
                    double Amu = o->Amu[q];
                    double ux, vx, wx, uy, vy, wy, uz, vz, wz, div;
                    double s11, s12, s13, s22, s23, s33, E1, E2, E3;
                    ux = uk[q] * Akx + ui[q] * Aix + us[q] * Asx;
                    vx = vk[q] * Akx + vi[q] * Aix + vs[q] * Asx;
                    wx = wk[q] * Akx + wi[q] * Aix + ws[q] * Asx;
                    uy = uk[q] * Aky + ui[q] * Aiy + us[q] * Asy;
                    vy = vk[q] * Aky + vi[q] * Aiy + vs[q] * Asy;
                    wy = wk[q] * Aky + wi[q] * Aiy + ws[q] * Asy;
                    uz = uk[q] * Akz + ui[q] * Aiz + us[q] * Asz;
                    vz = vk[q] * Akz + vi[q] * Aiz + vs[q] * Asz;
                    wz = wk[q] * Akz + wi[q] * Aiz + ws[q] * Asz;
                    div = ux + vy + wz;
                    s11 = (2.0 * ux - 2.0 / 3.0 * div) * Amu;
                    s22 = (2.0 * vy - 2.0 / 3.0 * div) * Amu;
                    s33 = (2.0 * wz - 2.0 / 3.0 * div) * Amu;
                    s12 = (uy + vx) * Amu;
                    s13 = (uz + wx) * Amu;
                    s23 = (vz + wy) * Amu;
                    double Amuk = Amu * o->vis_c;
                    double Tx = Tk[q] * Akx + Ti[q] * Aix + Ts[q] * Asx;
                    double Ty = Tk[q] * Aky + Ti[q] * Aiy + Ts[q] * Asy;
                    double Tz = Tk[q] * Akz + Ti[q] * Aiz + Ts[q] * Asz;
                    double uu = o->u[L], vv = o->v[L], ww = o->w[L];
                    E1 = uu * s11 + vv * s12 + ww * s13 + Amuk * Tx;
                    E2 = uu * s12 + vv * s22 + ww * s23 + Amuk * Ty;
                    E3 = uu * s13 + vv * s23 + ww * s33 + Amuk * Tz;
                    double Aj1 = o->A[AJAC][L];
                    double akx = o->A[ai[dir][0]][L] * Aj1;
                    double aky = o->A[ai[dir][1]][L] * Aj1;
                    double akz = o->A[ai[dir][2]][L] * Aj1;
                    o->Ev[0][L] = (akx * s11 + aky * s12 + akz * s13);
                    o->Ev[1][L] = (akx * s12 + aky * s22 + akz * s23);
                    o->Ev[2][L] = (akx * s13 + aky * s23 + akz * s33);
                    o->Ev[3][L] = (akx * E1 + aky * E2 + akz * E3);
                }
        for (int k = 0; k < 4; k++) fill_dir(o, o->Ev[k], dir);
        /* central derivative of Ev along dir, then du[1..4] += Ajac*vis */
        int n = dir == 0 ? nx : dir == 1 ? ny : nz;
        long fs = dir == 0 ? 1 : dir == 1 ? o->NX2 : (long)o->NX2 * o->NY2;
        long os = dir == 0 ? 1 : dir == 1 ? nx : (long)nx * ny;
        double h = dir == 0 ? o->hx : dir == 1 ? o->hy : o->hz;
        int bl = o->c.dif_bound[2 * dir], br = o->c.dif_bound[2 * dir + 1];
        int na = dir == 0 ? ny : nx, nb = dir == 2 ? ny : nz;
        for (int k = 0; k < 4; k++)
#pragma omp parallel for collapse(2) schedule(static)
            for (int b = 0; b < nb; b++)
                for (int a = 0; a < na; a++) {
                    int X = dir == 0 ? 0 : a, Y = dir == 0 ? a : dir == 1 ? 0 : b, Z = dir == 2 ? 0 : b;
                    const double *fl = o->Ev[k] + IL(o, dir == 0 ? 0 : X + LAP, dir == 1 ? 0 : Y + LAP, dir == 2 ? 0 : Z + LAP);
                    line_central(sv, n, fl, fs, h, bl, br, o->visd[k] + I0(o, X, Y, Z), os);
                }
        for (int k = 0; k < 4; k++)
#pragma omp parallel for schedule(static)
            for (int z = 0; z < nz; z++)
                for (int y = 0; y < ny; y++)
                    for (int x = 0; x < nx; x++) {
                        long q = I0(o, x, y, z);
                        double ajac = o->A[AJAC][IL(o, x + LAP, y + LAP, z + LAP)];
                        o->du[q + (k + 1) * N] += ajac * o->visd[k][q];
                    }
    }
}

/* du_comput, src/OCFD_Stream.cu:28-54 (STREAM=0) */
int orc_rhs(orc *o, int krk) {
    if (o->c.scheme_invis == OCFD_HYBRIDAUTO && krk == 1) hybrid_sensor(o);
    memset(o->du, 0, sizeof(double) * NV * o->nint);
    du_invis(o);
    du_vis(o);
    return 0;
}

/* OCFD_time_advance, src/OCFD_time.cu:44-123 */
static void time_advance(orc *o, int krk) {
    long N = o->nint;
    double dt = o->c.dt;
    for (int m = 0; m < 5; m++)
#pragma omp parallel for schedule(static)
        for (int z = 0; z < o->nz; z++)
            for (int y = 0; y < o->ny; y++)
                for (int x = 0; x < o->nx; x++) {
                    long q = I0(o, x, y, z) + m * N;
                    if (krk == 1) o->f[q] = o->fn[q] + dt * o->du[q];
                    else if (krk == 2) {
                        double t1 = 3.0 / 4.0, t2 = 1.0 / 4.0;
                        o->f[q] = t1 * o->fn[q] + t2 * (o->f[q] + dt * o->du[q]);
                    } else {
                        double t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
                        o->f[q] = t1 * o->fn[q] + t2 * (o->f[q] + dt * o->du[q]);
                        o->f_lap[m * o->nlap + IL(o, x + LAP, y + LAP, z + LAP)] = o->f[q];
                    }
                }
}

/* One SSP-RK3 stage on flat arrays (same arithmetic as time_advance). */
void orc_rk_stage(int n, int krk, double dt, double *f, const double *fn, const double *du) {
    for (int q = 0; q < n; q++) {
        if (krk == 1) f[q] = fn[q] + dt * du[q];
        else if (krk == 2) {
            double t1 = 3.0 / 4.0, t2 = 1.0 / 4.0;
            f[q] = t1 * fn[q] + t2 * (f[q] + dt * du[q]);
        } else {
            double t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
            f[q] = t1 * fn[q] + t2 * (f[q] + dt * du[q]);
        }
    }
}

/* filter_{x,y}3d_shock_kernel restated for one output point along dir with
 * the sigma values from the pressure, src/OCFD_filtering.cu:366-416,451-495 */
static double filt_sigma(orc *o, int dir, int X, int Y, int Z, double rth) {
    long st = dir == 0 ? 1 : dir == 1 ? o->NX2 : (long)o->NX2 * o->NY2;
    long L = IL(o, X, Y, Z);
    const double *P = o->P;
    double p_m2 = P[L - 2 * st], p_m1 = P[L - st], p_00 = P[L], p_p1 = P[L + st], p_p2 = P[L + 2 * st];
    double dp0 = 0.25 * (-p_p1 + 2.0 * p_00 - p_m1);
    double dp1 = 0.25 * (-p_p2 + 2.0 * p_p1 - p_00);
    double dp2 = 0.25 * (-p_00 + 2.0 * p_m1 - p_m2);
    double ri = 0.5 * ((dp0 - dp1) * (dp0 - dp1) + (dp0 - dp2) * (dp0 - dp2)) / (p_00 * p_00) + 1e-16;
    ri = 1.0 - rth / ri;
    return 0.5 * (ri + fabs(ri));
}

static void filter_point(orc *o, int dir, int X, int Y, int Z, double s0, double rth) {
    const double c1 = -0.2103830, c2 = 0.0396170;
    long st = dir == 0 ? 1 : dir == 1 ? o->NX2 : (long)o->NX2 * o->NY2;
    int sx = dir == 0, sy = dir == 1, sz = dir == 2;
    double hm = filt_sigma(o, dir, X - sx, Y - sy, Z - sz, rth);
    double h0 = filt_sigma(o, dir, X, Y, Z, rth);
    double hp = filt_sigma(o, dir, X + sx, Y + sy, Z + sz, rth);
    double Sc1 = 0.5 * (h0 + hp) * s0;
    double Sc2 = 0.5 * (h0 + hm) * s0;
    long L = IL(o, X, Y, Z);
    long q = I0(o, X - LAP, Y - LAP, Z - LAP);
    for (int m = 0; m < NV; m++) {
        const double *fl = o->f_lap + m * o->nlap;
        o->f[q + m * o->nint] = fl[L] - (Sc1 * (c1 * (fl[L + st] - fl[L]) + c2 * (fl[L + 2 * st] - fl[L - st]))
                                        - Sc2 * (c1 * (fl[L] - fl[L - st]) + c2 * (fl[L + st] - fl[L - 2 * st])));
    }
}

/* filtering, src/OCFD_filtering.cu:20-183 and the launch geometry of
 * filter_{x,y,z}3d_shock (src/OCFD_filtering.cu:418-606).  The x and y
 * launches filter the points ib <= i < ie-1 (in the filtered direction; the
 * block overlap of the kernels skips the last index).  filter_z3d_shock
 * launches the y kernel with a z-extended block (src/OCFD_filtering.cu:604):
 * it filters in y over ceil((je-jb)/8) blocks of 6 points. */
static void orc_filtering(orc *o) {
    int any = 0;
    for (int m = 0; m < o->c.nfilter; m++)
        if (o->istep % o->c.filter_npara[m][0] == 0) any = 1;
    if (!any) return;
    int nx = o->nx, ny = o->ny, nz = o->nz;
    /* set_para_filtering clamps, src/OCFD_filtering.cu:20-73 */
    int jx = 0, jy = 0, jz = 0;
    int para[OCFD_MAX_FILTERS][11];
    for (int k = 0; k < o->c.nfilter; k++) {
        for (int q = 0; q < 11; q++) para[k][q] = o->c.filter_npara[k][q];
        if (para[k][1] == 1) jx = 1;
        if (para[k][2] == 1) jy = 1;
        if (para[k][3] == 1) jz = 1;
        int ib = para[k][4], ie = para[k][5], jb = para[k][6], je = para[k][7], kb = para[k][8], ke = para[k][9];
        if (!o->c.iperiodic[0]) { if (ib < 6) ib = 6; if (ie > nx - 5) ie = nx - 5; }
        if (!o->c.iperiodic[1]) { if (jb < 6) jb = 6; if (je > ny - 5) je = ny - 5; }
        if (!o->c.iperiodic[2]) { if (kb < 6) kb = 6; if (ke > nz - 5) ke = nz - 5; }
        para[k][4] = ib; para[k][5] = ie; para[k][6] = jb; para[k][7] = je; para[k][8] = kb; para[k][9] = ke;
    }
    if (jx) { fill_dir(o, o->P, 0); for (int m = 0; m < 5; m++) fill_dir(o, o->f_lap + m * o->nlap, 0); }
    if (jy) { fill_dir(o, o->P, 1); for (int m = 0; m < 5; m++) fill_dir(o, o->f_lap + m * o->nlap, 1); }
    if (jz) { fill_dir(o, o->P, 2); for (int m = 0; m < 5; m++) fill_dir(o, o->f_lap + m * o->nlap, 2); }
    for (int m = 0; m < o->c.nfilter; m++) {
        if (!(o->tt <= o->c.filter_rpara[m][2])) continue;
        if (o->istep % para[m][0] != 0) continue;
        if (para[m][10] != 2) continue; /* only the shock filter exists (Fo9p aborts) */
        int ib = para[m][4] + LAP, ie = para[m][5] + LAP, jb = para[m][6] + LAP, je = para[m][7] + LAP;
        int kb = para[m][8] + LAP, ke = para[m][9] + LAP;
        double s0 = o->c.filter_rpara[m][0], rth = o->c.filter_rpara[m][1];
        if (para[m][1] == 1)
#pragma omp parallel for schedule(static)
            for (int Z = kb; Z < ke; Z++)
                for (int Y = jb; Y < je; Y++)
                    for (int X = ib; X < ie - 1; X++) filter_point(o, 0, X, Y, Z, s0, rth);
        if (para[m][2] == 1)
#pragma omp parallel for schedule(static)
            for (int Z = kb; Z < ke; Z++)
                for (int Y = jb; Y < je - 1; Y++)
                    for (int X = ib; X < ie; X++) filter_point(o, 1, X, Y, Z, s0, rth);
        if (para[m][3] == 1) {
            int nby = (je - jb + 7) / 8;
            int yend = jb + 6 * nby;
            if (yend > je - 1) yend = je - 1;
#pragma omp parallel for schedule(static)
            for (int Z = kb; Z < ke; Z++)
                for (int Y = jb; Y < yend; Y++)
                    for (int X = ib; X < ie; X++) filter_point(o, 1, X, Y, Z, s0, rth);
        }
    }
}

/* NS_solver_real loop body, src/OCFD_NS_Solver.cu:63-101 */
int orc_step(orc *o, int nsteps) {
    for (int s = 0; s < nsteps; s++) {
        double *t = o->f; o->f = o->fn; o->fn = t;
        for (int krk = 1; krk <= 3; krk++) {
            orc_rhs(o, krk);
            time_advance(o, krk);
            cons_to_pri(o);
            ocfd_bc(o);
            get_amu(o);
        }
        o->istep++;
        o->tt += o->c.dt;
        orc_filtering(o);
    }
    return 0;
}

/* ---- getters ------------------------------------------------------------ */
static void get_interior(orc *o, const double *a, double *out) {
    for (int z = 0; z < o->nz; z++)
        for (int y = 0; y < o->ny; y++)
            for (int x = 0; x < o->nx; x++) out[I0(o, x, y, z)] = a[IL(o, x + LAP, y + LAP, z + LAP)];
}
int orc_get_flow(orc *o, double *d, double *u, double *v, double *w, double *T, int *istep, double *tt) {
    if (d) get_interior(o, o->d, d);
    if (u) get_interior(o, o->u, u);
    if (v) get_interior(o, o->v, v);
    if (w) get_interior(o, o->w, w);
    if (T) get_interior(o, o->T, T);
    if (istep) *istep = o->istep;
    if (tt) *tt = o->tt;
    return 0;
}
int orc_num_threads(void) { return omp_get_max_threads(); }

int orc_get_state(orc *o, double *f5) { memcpy(f5, o->f, sizeof(double) * NV * o->nint); return 0; }
int orc_get_du(orc *o, double *du5) { memcpy(du5, o->du, sizeof(double) * NV * o->nint); return 0; }
int orc_get_pressure(orc *o, double *p) { get_interior(o, o->P, p); return 0; }
int orc_get_pressure_lap(orc *o, double *p) { memcpy(p, o->P, sizeof(double) * o->nlap); return 0; }
/* primitive k (0..4 = d,u,v,w,T) over the whole LAP box, ghosts included */
int orc_get_prim_lap(orc *o, int k, double *a) {
    const double *src[5] = {o->d, o->u, o->v, o->w, o->T};
    if (k < 0 || k > 4) return -1;
    memcpy(a, src[k], sizeof(double) * o->nlap);
    return 0;
}
/* The filter alone on the current state, for the pins: f_lap <- f (as the
 * third RK stage leaves it), p and the primitives from f (cons_to_pri), then
 * filtering with every filter due. */
int orc_filter_now(orc *o) {
    long N = o->nint;
    for (int m = 0; m < NV; m++)
        for (int z = 0; z < o->nz; z++)
            for (int y = 0; y < o->ny; y++)
                for (int x = 0; x < o->nx; x++)
                    o->f_lap[m * o->nlap + IL(o, x + LAP, y + LAP, z + LAP)] = o->f[I0(o, x, y, z) + m * N];
    cons_to_pri(o);
    int save = o->istep;
    o->istep = 0;
    orc_filtering(o);
    o->istep = save;
    return 0;
}
int orc_get_metric_lap(orc *o, int k, double *a) { memcpy(a, o->A[k], sizeof(double) * o->nlap); return 0; }
int orc_get_amu(orc *o, double *amu) { memcpy(amu, o->Amu, sizeof(double) * o->nint); return 0; }
int orc_get_metrics(orc *o, double *a13) {
    for (int k = 0; k < 13; k++) get_interior(o, o->A[k], a13 + (long)k * o->nint);
    return 0;
}
int orc_get_flags(orc *o, signed char *fx, signed char *fy, signed char *fz) {
    long n0 = (long)(o->nx + 1) * o->ny * o->nz, n1 = (long)o->nx * (o->ny + 1) * o->nz, n2 = (long)o->nx * o->ny * (o->nz + 1);
    if (fx) for (long i = 0; i < n0; i++) fx[i] = (signed char)o->flag[0][i];
    if (fy) for (long i = 0; i < n1; i++) fy[i] = (signed char)o->flag[1][i];
    if (fz) for (long i = 0; i < n2; i++) fz[i] = (signed char)o->flag[2][i];
    return 0;
}

/* Q criterion (get_Q / get_Q_kernal, src/OCFD_ana.cu:299-458) at the interior
 * points from the current primitives (their ghost layers as they stand). */
int orc_get_q(orc *o, double *q) {
    central_grads(o, 3);
    const double *uk = o->grad[0], *ui = o->grad[1], *us = o->grad[2];
    const double *vk = o->grad[3], *vi = o->grad[4], *vs = o->grad[5];
    const double *wk = o->grad[6], *wi = o->grad[7], *ws = o->grad[8];
    for (int z = 0; z < o->nz; z++)
        for (int y = 0; y < o->ny; y++)
            for (int x = 0; x < o->nx; x++) {
                long p = I0(o, x, y, z), L = IL(o, x + LAP, y + LAP, z + LAP);
                double Akx = o->A[AKX][L], Aky = o->A[AKY][L], Akz = o->A[AKZ][L];
                double Aix = o->A[AIX][L], Aiy = o->A[AIY][L], Aiz = o->A[AIZ][L];
                double Asx = o->A[ASX][L], Asy = o->A[ASY][L], Asz = o->A[ASZ][L];
                double ux = uk[p] * Akx + ui[p] * Aix + us[p] * Asx;
                double vx = vk[p] * Akx + vi[p] * Aix + vs[p] * Asx;
                double wx = wk[p] * Akx + wi[p] * Aix + ws[p] * Asx;
                double uy = uk[p] * Aky + ui[p] * Aiy + us[p] * Asy;
                double vy = vk[p] * Aky + vi[p] * Aiy + vs[p] * Asy;
                double wy = wk[p] * Aky + wi[p] * Aiy + ws[p] * Asy;
                double uz = uk[p] * Akz + ui[p] * Aiz + us[p] * Asz;
                double vz = vk[p] * Akz + vi[p] * Aiz + vs[p] * Asz;
                double wz = wk[p] * Akz + wi[p] * Aiz + ws[p] * Asz;
                double aj = o->A[AJAC][L];
                q[p] = (ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy) * aj * aj;
            }
    return 0;
}

/* Characteristic matrices of a face (pins of the projection). */
void orc_char_matrices(const double *s0, const double *s1, double gamma, double *L, double *R) {
    double P[14];
    char_para(s0, s1, gamma, P);
    char_matrices(P, L, R);
}

/* One line of the characteristic flux derivative with contiguous arrays:
 * F (5 x (n+8)), prim (7 x (n+8)) = u v w c a1 a2 a3, ajac (n+8), flags
 * (n+1, face j+1/2 at j+1) or NULL; du (5 x n) accumulated. */
void orc_line_flux_char(int scheme, int side, int dir, int n, const double *F, const double *prim,
                        const double *ajac, double h, int bl, int br, int nonref, const int *flags, double gamma,
                        double *du) {
    const double *f5[5], *p7[7];
    double *d5[5];
    for (int m = 0; m < 5; m++) {
        f5[m] = F + (long)m * (n + 8);
        d5[m] = du + (long)m * n;
    }
    for (int t = 0; t < 7; t++) p7[t] = prim + (long)t * (n + 8);
    line_flux_char(scheme, side, dir, n, f5, 1, p7, ajac, h, bl, br, nonref, flags, 1, gamma, d5, 1, 5.0, 0.2);
}
