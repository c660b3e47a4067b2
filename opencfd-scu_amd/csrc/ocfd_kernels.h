// ocfd_kernels.h - launchers of the gfx950 kernels (host-callable).
#pragma once

#include "ocfd_dev.h"
#include "../../include/ocfd.h"

// Inviscid sweep along one direction.
struct SweepArgs {
    Geo g;
    Consts c;
    const double *d, *u, *v, *w, *T;     // primitives (LAP layout)
    const double *m1, *m2, *m3;          // the direction's cofactors
    const double *jac;                   // Ajac
    const int8_t *flag;                  // hybrid flags of this direction (face j+1/2 at node j) or null
    double *du;                          // 5 fields, stride fs
    long long fs;
    double h;                            // hx / hy / hz
    int bl, br;                          // DIF_BOUND on the low / high face
    int nrp, nrm;                        // NON_REFLETION for the P / M flux
};
int launch_sweep(int scheme, int dir, const SweepArgs& a, bool first, hipStream_t st);

// Viscous terms.
struct ViscArgs {
    Geo g;
    Consts c;
    int svis;                                // 203 / 204
    int bnd[6];                              // DIF_BOUND
    const double *u, *v, *w, *T, *amu;
    const double *M[M_COUNT];
    double *ev;                              // 12 fields: Ev1..4 for x, y, z
    double *du;                              // 5 fields
    long long fs;
};
// Gradients of u,v,w,T (kept in registers) and the 12 viscous fluxes.
void launch_visc_gf(const ViscArgs& a, hipStream_t st);
void launch_visc_div(const ViscArgs& a, hipStream_t st);

// Periodic face fill (exchange_boundary_?_packed_dev at one rank): copy the
// 4-layer faces of `nf` fields (stride fs) in direction dir.
void launch_fill(const Geo& g, double* base, int nf, long long fs, int dir, hipStream_t st);

// RK stage + cons->prim + viscosity.
struct StageArgs {
    Geo g;
    Consts c;
    int krk;
    double dt;
    double *f;
    const double *fn, *du;
    double *flap;                            // 5 fields or null (written in stage 3 when non-null)
    double *d, *u, *v, *w, *T, *P, *amu;
    long long fs;
};
void launch_stage(const StageArgs& a, hipStream_t st);
void launch_c2p_all(const StageArgs& a, hipStream_t st);

// IBC 108 and the face refresh.
struct BCArgs {
    Geo g;
    Consts c;
    double *d, *u, *v, *w, *T, *amu, *f;
    long long fs;
    const double *ub1, *fx, *gz;           // inlet profile [4][ny], blowing tables
    int x_begin;
    double ht, epsl, tw;
    int wall_not_normal;
};
void launch_bc108(const BCArgs& a, hipStream_t st);
void launch_face_refresh(const BCArgs& a, hipStream_t st);
void launch_p2c_all(const BCArgs& a, hipStream_t st);
void launch_amu_all(const BCArgs& a, hipStream_t st);

// hybrid_auto style 2 sensor: flags of the three directions.
void launch_sensor(const Geo& g, const Consts& c, const double* d, const double* T, int8_t* fx, int8_t* fy,
                   int8_t* fz, hipStream_t st);

// Shock-capturing filter in direction dir over [lo,hi) boxes (interior indices).
struct FilterArgs {
    Geo g;
    const double *P, *flap;
    double *f;
    long long fs;
    double s0, rth;
    int dir;
    int lo[3], hi[3];
};
void launch_filter(const FilterArgs& a, hipStream_t st);

// Metrics from mesh coordinates (init).
void launch_metric_deriv(const Geo& g, const double* coord, int dir, int bound, double* out, hipStream_t st);
void launch_cofactors(const Geo& g, const double* D9, long long fs, double* const* M, hipStream_t st);
void launch_set(const Geo& g, double* base, long long n, double value, hipStream_t st);
