// ocfd_kernels.h - launchers of the gfx950 kernels (host-callable).
//
// S is the storage type of the primitives, the viscosity, the right-hand side
// and the viscous fluxes: double on the plain path, float on the mixed path
// (which keeps the conserved state f/fn, the metrics and the filter input in
// double).  Every launcher is instantiated for both.
#pragma once

#include "ocfd_dev.h"
#include "../../include/ocfd.h"

// Inviscid sweep along one direction.
template <typename S>
struct SweepArgsT {
    Geo g;
    Consts c;
    const S *d, *u, *v, *w, *T;          // primitives (LAP layout)
    const double *m1, *m2, *m3;          // the direction's cofactors
    const double *jac;                   // Ajac
    const int8_t *flag;                  // hybrid flags of this direction (face j+1/2 at node j) or null
    S *du;                               // 5 fields, stride fs
    long long fs;
    double h;                            // hx / hy / hz
    int bl, br;                          // DIF_BOUND on the low / high face
    int nrp, nrm;                        // NON_REFLETION for the P / M flux
    // The learned hybrid sensor, run by the x sweep of the first RK stage
    // (hybrid_auto only): when P is set the x sweep computes the flags of its
    // rows' nodes (i >= -1) itself, writes them to fx, fy, fz and the
    // probabilities to prob, and uses them for its own faces.
    struct {
        const double* M[M_COUNT];        // cofactors and Ajac (the features' contravariant velocities)
        const float* P;                  // packed network weights (launch_mlp_pack), or null
        float lo, hi;                    // confidence band that falls back to the plain sensor
        int8_t *fx, *fy, *fz;
        float* prob;
    } mlp;
};
template <typename S>
int launch_sweep(int scheme, int dir, const SweepArgsT<S>& a, bool first, hipStream_t st);
// The same sweep with the reconstruction in characteristic space
// (CHARTERIC = 1): WENO7_SYMBO, WENO7_SYMBO_LIM and SCHEME_HYBRIDAUTO only
// (returns -1 for the other schemes, which the reference cannot run so).
template <typename S>
int launch_sweep_char(int scheme, int dir, const SweepArgsT<S>& a, bool first, hipStream_t st);

// Viscous terms.
template <typename S>
struct ViscArgsT {
    Geo g;
    Consts c;
    int svis;                                // 203 / 204
    int bnd[6];                              // DIF_BOUND
    const S *u, *v, *w, *T, *amu;
    const double *M[M_COUNT];
    S *ev;                                   // 12 fields: Ev1..4 for x, y, z
    S *du;                                   // 5 fields
    S *dv;                                   // 4 fields: the viscous part of du[1..4], or null (added to du)
    long long fs;
};
// Gradients of u,v,w,T (kept in registers) and the 12 viscous fluxes.
template <typename S> void launch_visc_gf(const ViscArgsT<S>& a, hipStream_t st);
// The divergence in launches of visc_div_components<S>() flux components
// each, starting at component c0 (0..3): written to dv when it is set (the
// stage adds du + dv, so the divergence may run beside the inviscid sweeps),
// else added to du after the sweeps.
template <typename S> int visc_div_components();
template <typename S> void launch_visc_div(const ViscArgsT<S>& a, int c0, hipStream_t st);
// Q criterion of u, v, w at every interior point into q (LAP layout, FP64).
template <typename S> void launch_qcrit(const ViscArgsT<S>& a, double* q, hipStream_t st);

// Periodic face fill (exchange_boundary_?_packed_dev at one rank): copy the
// 4-layer faces of `nf` fields (stride fs) in direction dir.
template <typename T>
void launch_fill(const Geo& g, T* base, int nf, long long fs, int dir, hipStream_t st);

// RK stage + cons->prim + viscosity.
template <typename S>
struct StageArgsT {
    Geo g;
    Consts c;
    int krk;
    double dt;
    double *f;
    const double *fn;
    const S *du;                             // right-hand side without the Jacobian factor
    const S *dv;                             // 4 fields, the viscous part of du[1..4], or null
    const double *jac;                       // Ajac: the stage advances with Ajac * (du + dv)
    double *flap;                            // 5 fields or null (written in stage 3 when non-null)
    S *d, *u, *v, *w, *T;
    S *amu;                                  // viscosity, or null (the stage then skips it)
    S *P;                                    // pressure, or null (the stage then skips it)
    long long fs;
};
template <typename S> void launch_stage(const StageArgsT<S>& a, hipStream_t st);

// SSP-RK3 stage of one point (OCFD_time_advance_ker1/2/3, src/OCFD_time.cu:
// 44-94) with cons->prim + p (cons_to_pri_kernel, src/commen_kernel.cu:
// 77-109) and Sutherland's law (get_Amu_kernal, src/commen_kernel.cu:122-131);
// du[m] is the point's right-hand side (with the Jacobian factor applied).
template <typename S>
__device__ __forceinline__ void stage_point(const StageArgsT<S>& A, long long q, const double* du) {
    const long long fs = A.fs;
    const double dt = A.dt;
    double fv[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const long long o = m * fs + q;
        // the conserved state is streamed (non-temporal loads and stores):
        // nothing reads it again before the next stage, while the
        // primitives rho, u, v, w, T written below are read at once by the
        // next right-hand side and stay ordinary stores (measured: the stage 174 -> 147 us
        // on the 512x128x64 ramp; streaming the primitives too, 167 us)
        double r;
        const double fno = __builtin_nontemporal_load(&A.fn[o]);
        const double fo = A.krk == 1 ? 0.0 : __builtin_nontemporal_load(&A.f[o]);
        if (A.krk == 1) {
            r = fno + dt * du[m];
        } else if (A.krk == 2) {
            const double t1 = 3.0 / 4.0, t2 = 1.0 / 4.0;
            r = t1 * fno + t2 * (fo + dt * du[m]);
        } else {
            const double t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
            r = t1 * fno + t2 * (fo + dt * du[m]);
        }
        __builtin_nontemporal_store(r, &A.f[o]);
        if (A.flap) __builtin_nontemporal_store(r, &A.flap[o]);
        fv[m] = r;
    }
    const double d1 = fv[0];
    const double u1 = fv[1] / d1, v1 = fv[2] / d1, w1 = fv[3] / d1;
    const double T2 = (fv[4] - 0.5 * d1 * (u1 * u1 + v1 * v1 + w1 * w1)) / (d1 * A.c.cv);
    A.d[q] = S(d1); A.u[q] = S(u1); A.v[q] = S(v1); A.w[q] = S(w1); A.T[q] = S(T2);
    // the pressure (read after the step) is streamed; the viscosity only
    // when amu is set (the time step leaves it null: the gradient kernel forms
    // the viscosity from T itself)
    if (A.P) __builtin_nontemporal_store(S(T2 * d1 / (A.c.gamma * A.c.ama * A.c.ama)), &A.P[q]);
    if (A.amu) __builtin_nontemporal_store(S(A.c.amu_c0 * sqrt(T2 * T2 * T2) / (A.c.tsb + T2)), &A.amu[q]);
}

template <typename S> void launch_c2p_all(const StageArgsT<S>& a, hipStream_t st);

// IBC 108 and the face refresh.
template <typename S>
struct BCArgsT {
    Geo g;
    Consts c;
    S *d, *u, *v, *w, *T, *amu;
    double *f;
    long long fs;
    const double *ub1, *fx, *gz;           // inlet profile [4][ny], blowing tables
    int x_begin;
    double ht, epsl, tw;
    int wall_not_normal;
    int mu_faces;                          // viscosity refreshed on faces: bit 2*dir (low), 2*dir+1 (high)
    // IBC 124 (lifting body): inlet [5][nz][ny], upper [5][ny][nx] (or null:
    // free stream), wall disturbance [3][ny][nx] (or null: none), upper
    // disturbance [ny][nx]
    const double *in124, *up124, *coeff124, *udist124;
    double sin_aoa, cos_aoa;
};
template <typename S> void launch_bc108(const BCArgsT<S>& a, hipStream_t st);
// IBC 108 followed by the face refresh, in one launch (the same values as
// launch_bc108 then launch_face_refresh).
template <typename S> void launch_bc108_faces(const BCArgsT<S>& a, hipStream_t st);
// IBC 124 (bc_user_Liftbody3d, src/OCFD_boundary_Liftbody3D.cu:185-294):
// inlet, upper face, wall, then the symmetry planes when sym != 0.
template <typename S> void launch_bc124(const BCArgsT<S>& a, int sym, hipStream_t st);
// Mirror of one LAP field into the y ghost layers about j = 0 and j = ny-1,
// times sign, over the interior x and z (boundary_Jac3d_kernal_y_l/r,
// src/OCFD_Comput_Jacobian3d.cu:429-448; do_symmetry_kernel_m/p).
template <typename T> void launch_mirror_y(const Geo& g, T* f, double sign, hipStream_t st);
template <typename S> void launch_face_refresh(const BCArgsT<S>& a, hipStream_t st);
template <typename S> void launch_p2c_all(const BCArgsT<S>& a, hipStream_t st);
template <typename S> void launch_amu_all(const BCArgsT<S>& a, hipStream_t st);

// hybrid_auto style 2 sensor: flags of the three directions.
template <typename S>
void launch_sensor(const Geo& g, const Consts& c, const S* d, const S* T, int8_t* fx, int8_t* fy, int8_t* fz,
                   hipStream_t st);

// Shock-capturing filter in direction dir over [lo,hi) boxes (interior indices).
template <typename S>
struct FilterArgsT {
    Geo g;
    const S *P;
    const double *flap;
    double *f;
    long long fs;
    double s0, rth;
    int dir;
    int lo[3], hi[3];
};
template <typename S> void launch_filter(const FilterArgsT<S>& a, hipStream_t st);

// Metrics from mesh coordinates (init).
void launch_metric_deriv(const Geo& g, const double* coord, int dir, int bound, double* out, hipStream_t st);
void launch_cofactors(const Geo& g, const double* D9, long long fs, double* const* M, hipStream_t st);
void launch_set(const Geo& g, double* base, long long n, double value, hipStream_t st);
// Step diagnostics: per-block partial sums of E = f[4] and T over the interior
// (diag_sum_blocks(g) partials each) and NaN / negative counts of T over the
// LAP box (2*DIAG_SCAN_BLOCKS ints).  T_base is the T field (index 0 = first
// LAP-box element).
constexpr int DIAG_SCAN_BLOCKS = 1024;
int diag_sum_blocks(const Geo& g);
template <typename S>
void launch_diag(const Geo& g, const double* E, const S* T_base, double* partE, double* partT, int* scan_part,
                 hipStream_t st);
// The partials above reduced on the device by one block: out4 = {sum E,
// sum T, NaN count, negative count} (the counts stored as long long).
void launch_diag_final(const double* partE, const double* partT, long long nb, const int* scan_part, double* out4,
                       hipStream_t st);
// Running time average of the primitives: avg (5 FP64 fields, stride fs)
// becomes (n avg + x)/(n + 1) at every interior point.
template <typename S>
void launch_time_avg(const Geo& g, const S* d, const S* u, const S* v, const S* w, const S* T, double* avg,
                     long long fs, int n, hipStream_t st);

// Learned sensor (k_mlp.hip): parameters W1 [32][24], b1, W2 [32][32], b2,
// W3, b3 as 1889 floats.
constexpr int MLP_NPARAM = 1889;
// The weights as the inference kernels read them (launch_mlp_pack): the
// BF16 MFMA fragments and biases of each of a wave's 64 lanes, lane fastest.
constexpr int MLP_PACKED_FLOATS = 2564;
void launch_mlp_pack(const float* params, float* packed, hipStream_t st);
template <typename S>
struct MlpSensorArgsT {
    Geo g;
    Consts c;
    const S *d, *u, *v, *w, *T;
    const double *M[M_COUNT];
    const float *params;          // packed weights (launch_mlp_pack)
    float lo, hi;                 // confidence band that falls back to the plain sensor
    int8_t *fx, *fy, *fz;
    float *prob;                  // probability per node (LAP layout) or null
};
// plane 0: the whole pass; 1 / 2: only the ghost plane j = -1 / k = -1 (the
// rest is computed by the x sweep of the first RK stage)
template <typename S> void launch_mlp_sensor(const MlpSensorArgsT<S>& a, int plane, hipStream_t st);
// Features and plain labels of the interior nodes (i, j, k) with i, j, k
// multiples of stride, in lattice order: mlp_collect_count(g, stride) rows.
template <typename S> void launch_mlp_collect(const MlpSensorArgsT<S>& a, int stride, float* X, float* y, hipStream_t st);
long long mlp_collect_count(const Geo& g, int stride);
// One training step on the nb samples idx[0..nb) of (X, y) (idx null: rows
// 0..nb): per-sample vectors into per (nb * mlp_per_sample_floats()), the
// batch-mean gradient into grad (may be null), the mean loss into *loss, and
// with adam_t > 0 the Adam update of params (moments m1, m2).
void launch_mlp_step(float* params, const float* X, const float* y, const int* idx, int nb, float wpos, float* per,
                     float* grad, float* m1, float* m2, int adam_t, float lr, float* loss, hipStream_t st);
size_t mlp_per_sample_floats();

// dst[i] = (To)src[i] for n elements (precision conversion of whole fields).
template <typename To, typename From>
void launch_convert(To* dst, const From* src, long long n, hipStream_t st);
