/*
 * ocfd_oracle.h - CPU restatement of the OpenCFD-SCU right-hand-side path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never linked into libocfd_hip.so and never used as a
 * fallback.
 *
 * The oracle keeps the reference's data layout (LAP=4 ghost layers around
 * every field, interior-only SoA for f/fn/du) and its launch semantics
 * (which points each kernel writes, in which order terms are added to du), so
 * that a difference from the HIP path points at an arithmetic difference and
 * not at a layout difference.  See ocfd_oracle.c for the per-function
 * reference citations.
 */
#ifndef OCFD_ORACLE_H
#define OCFD_ORACLE_H

#include "../include/ocfd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc *orc_create(const ocfd_config *cfg);
void orc_destroy(orc *o);
int orc_load_mesh(orc *o, const double *x, const double *y, const double *z);
int orc_load_jacobi(orc *o, const double *a13);
int orc_load_flow(orc *o, const double *d, const double *u, const double *v,
                  const double *w, const double *T, int istep, double tt);
int orc_set_inlet_profile(orc *o, const double *prof4xny);
/* IBC 124 tables: inlet5 (5 x nz*ny), upper5 (5 x ny*nx); either NULL. */
int orc_set_bc_tables(orc *o, const double *inlet5, const double *upper5);
int orc_init(orc *o);
int orc_rhs(orc *o, int krk);
int orc_step(orc *o, int nsteps);
/* OpenMP threads the point loops use. */
int orc_num_threads(void);
int orc_get_flow(orc *o, double *d, double *u, double *v, double *w,
                 double *T, int *istep, double *tt);
int orc_get_state(orc *o, double *f5);
int orc_get_du(orc *o, double *du5);
int orc_get_pressure(orc *o, double *p);
/* The pressure on the whole LAP box ((nz+8)(ny+8)(nx+8), x fastest). */
int orc_get_pressure_lap(orc *o, double *p);
int orc_get_prim_lap(orc *o, int k, double *a);
int orc_filter_now(orc *o);
/* Metric array k (0..12: Axx Ayy Azz Akx ... Ajac) over the LAP box. */
int orc_get_metric_lap(orc *o, int k, double *a);
int orc_get_amu(orc *o, double *amu);
int orc_get_metrics(orc *o, double *a13);
int orc_get_flags(orc *o, signed char *fx, signed char *fy, signed char *fz);
/* Characteristic projection of a face from the u v w c a1 a2 a3 of its two
 * nodes: left and right eigenvector matrices (5x5, row-major). */
void orc_char_matrices(const double *s0, const double *s1, double gamma, double *L, double *R);
/* One line of the CHARTERIC = 1 flux derivative (contiguous arrays). */
void orc_line_flux_char(int scheme, int side, int dir, int n, const double *F, const double *prim,
                        const double *ajac, double h, int bl, int br, int nonref, const int *flags, double gamma,
                        double *du);
/* Q criterion at the interior points (nz*ny*nx). */
int orc_get_q(orc *o, double *q);

/* Unit-level entry points used by the pinning tests. */
/* Half-node flux of one scheme on a stencil s[0..7] laid out as the
 * reference's kernels gather it (P: nodes j-3..j+4 for face j+1/2;
 * M: nodes j-3..j+4 as well, i.e. x-4..x+3 around x=j+1). `scheme` is a
 * scheme id; for OCFD_HYBRIDAUTO `hyflag` selects 1=UP7, 2=WENO7, 3=WENO5. */
double orc_face_flux(int scheme, int side, int hyflag, const double *s8,
                     double tv_limiter, double tv_max);
/* Steger-Warming split of one point: out[0..4]=f+, out[5..9]=f-. */
void orc_split_point(double d, double u, double v, double w, double cc,
                     double A1, double A2, double A3, double gamma,
                     double epsl_sw, double *out10);
/* 1-D inviscid derivative of one variable along a line of n interior points
 * with 4 ghosts each side: du[i] += -ajac[i]*(h(i+1/2)-h(i-1/2))/h.  fpm has
 * n+8 values (node -4..n+3).  flags: hybrid flags per face (n+1) or NULL. */
void orc_line_flux_derivative(int scheme, int side, int n, const double *f,
                              const double *ajac, double h, int bound_l,
                              int bound_r, int nonref, const int *flags,
                              double *du);
/* 1-D central derivative (CD6/CD8 with D0 closures) of n+8 values. */
void orc_line_central(int scheme_vis, int n, const double *f, double h,
                      int bound_l, int bound_r, double *out);
/* One SSP-RK3 stage (OCFD_time_advance_ker1/2/3) on n flat values. */
void orc_rk_stage(int n, int krk, double dt, double *f, const double *fn,
                  const double *du);
/* Jameson pressure sensor value at one node from its 6 neighbours. */
double orc_jameson_dp(double pc, double pxm, double pxp, double pym,
                      double pyp, double pzm, double pzp);

#ifdef __cplusplus
}
#endif
#endif
