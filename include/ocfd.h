/*
 * ocfd.h - C ABI of the MI355X (gfx950) right-hand-side path of OpenCFD-SCU.
 *
 * Two layers are exported by libocfd_hip.so:
 *
 *  1. Reference-named entry points with the same signatures and meaning as the
 *     reference's extern "C" functions.  They act on one process-wide context,
 *     so the reference's own control flow (opencfd.c) can call them unchanged.
 *
 *  2. An explicit context API (ocfd_*) underneath them: the context owns every
 *     device buffer, the caller owns every host array, each call returns an int
 *     status (0 = ok), one HIP stream per context, no hidden host threads, not
 *     thread-safe per context.
 *
 * Host arrays at the boundary are FP64 in the reference's file order: x fastest,
 * then y, then z, interior points only (no ghost layers).  Internally the
 * library keeps the reference's LAP=4 ghost layout on the GPU.
 *
 * Citations are to the reference tree (OpenCFD-SCU); paths are relative to it.
 */
#ifndef OCFD_H
#define OCFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCFD_LAP 4          /* head/config_parameters.h:15 */
#define OCFD_NVARS 5        /* head/config_parameters.h:17 */
#define OCFD_MAX_FILTERS 10 /* src/parameters.c:102 (NF_max) */
#define OCFD_MAX_ANA 16     /* ANA_EVENT& keys kept (the reference mallocs N_ana) */
#define OCFD_ANA_NPARA 100  /* ANA_NPARA&/ANA_RPARA& values, src/parameters.c:346-358 */
/* Wall-disturbance time modes kept: MTMAX (BC_NPARA[1] for IBC 108,
 * BC_NPARA[4] for IBC 124) above this is refused by ocfd_create
 * (src/OCFD_boundary_init.c:501-517 sizes its tables to MTMAX). */
#define OCFD_MTMAX 64

/* Scheme identifiers, src/parameters.c:567-637. */
enum {
    OCFD_CD6 = 203, OCFD_CD8 = 204,
    OCFD_UP7 = 301, OCFD_WENO5 = 302, OCFD_WENO7 = 303, OCFD_WENO7_SYMBO = 304,
    OCFD_WENO7_SYMBO_LIM = 305, OCFD_NND2 = 306, OCFD_OMP6_HR = 307,
    OCFD_OMP6_LD = 308, OCFD_OMP6_CD8 = 309, OCFD_HYBRIDAUTO = 310
};

/* Arithmetic of the bulk kernels. */
enum { OCFD_PREC_FP64 = 0, OCFD_PREC_MIXED = 1 };

/* Status codes. */
enum {
    OCFD_OK = 0, OCFD_ERR_ARG = 1, OCFD_ERR_CONFIG = 2, OCFD_ERR_IO = 3,
    OCFD_ERR_HIP = 4, OCFD_ERR_UNSUPPORTED = 5, OCFD_ERR_STATE = 6,
    OCFD_ERR_NUMERIC = 7 /* NaN or too many negative T (the reference aborts) */
};

/*
 * Contents of opencfd-scu.in (src/parameters.c:166-564).  Field names follow
 * the input keys; comments give the key.
 */
typedef struct ocfd_config {
    int nx, ny, nz;              /* GRID_3D                                  */
    int npx0, npy0, npz0;        /* PARALLEL_3D (must be 1 1 1 here)         */
    int stream_mode;             /* STREAM (0 supported)                     */
    int charteric;               /* CHARTERIC (1: SYMBO / hybrid_auto only)  */
    int test;                    /* TEST                                     */
    int iperiodic[3];            /* IPERIODIC                                */
    int jac_bound[3];            /* JAC_BOUND                                */
    int dif_bound[6];            /* DIF_BOUND  (x-, x+, y-, y+, z-, z+)      */
    int non_ref[6];              /* NON_REFLETION (same order)               */
    int scheme_invis;            /* SCHEME_INVIS -> 301..310                 */
    int scheme_vis;              /* SCHEME_VIS   -> 203/204                  */
    int hy_style;                /* HY_STYLE (2 supported)                   */
    int hy_smooth_dp;            /* HY_SMOOTH_DP                             */
    int hy_patch_zone;           /* HY_PATCH_ZONE                            */
    double hy_dp_intv[2];        /* HY_DP_INTV                               */
    double re, ama, gamma, pr;   /* RE AMA GAMMA PR                          */
    double t_ref;                /* T_REF (<=0 -> 288.15)                    */
    double epsl_sw;              /* EPSL_SW                                  */
    double dt, end_time;         /* DT END_TIME                              */
    int kstep_show, kstep_save;  /* KSTEP_SHOW KSTEP_SAVE                    */
    int init_stat;               /* INIT_STAT                                */
    int ibc;                     /* IBC                                      */
    int n_bc_npara, n_bc_rpara;
    int bc_npara[32];            /* BC_NPARA                                 */
    double bc_rpara[32];         /* BC_RPARA                                 */
    int nfilter;                 /* number of FILTER_NPARA& keys             */
    int filter_npara[OCFD_MAX_FILTERS][11]; /* FILTER_NPARA&                */
    double filter_rpara[OCFD_MAX_FILTERS][3]; /* FILTER_RPARA&              */
    int precision;               /* OCFD_PREC_* (not an input key)           */
    int nana;                    /* number of ANA_EVENT& keys                */
    int ana_event[OCFD_MAX_ANA][2];       /* ANA_EVENT&: style, every Kstep  */
    int n_ana_npara[OCFD_MAX_ANA];        /* values given in ANA_NPARA&      */
    int n_ana_rpara[OCFD_MAX_ANA];        /* values given in ANA_RPARA&      */
    int ana_npara[OCFD_MAX_ANA][OCFD_ANA_NPARA];    /* ANA_NPARA&            */
    double ana_rpara[OCFD_MAX_ANA][OCFD_ANA_NPARA]; /* ANA_RPARA&            */
} ocfd_config;

typedef struct ocfd_ctx ocfd_ctx;

/* ---- configuration / formats (host side) ------------------------------- */

/* Parse an opencfd-scu.in file with the reference's key rules
 * (src/parameters.c:197-564, 667-792): spaces in keys are ignored, the last
 * occurrence of a key wins, '#' lines never match, unknown keys are ignored,
 * numbered FILTER_NPARA&/FILTER_RPARA& families are discovered by stripping
 * digits, and so are ANA_EVENT&/ANA_NPARA&/ANA_RPARA& (src/parameters.c:
 * 328-358).  Unset keys are 0.  Returns OCFD_ERR_CONFIG on a bad scheme name
 * (the reference exits, src/parameters.c:661-664). */
int ocfd_read_config(const char *path, ocfd_config *cfg);

/* Flow file (opencfd.dat / OCFD%08d.dat): header record {int32 Istep, f64 tt}
 * then d,u,v,w,T each as nz Fortran records of nx*ny f64
 * (src/OCFD_IO.c:24-228).  Arrays are nx*ny*nz, x fastest. */
int ocfd_read_flow_file(const char *path, int nx, int ny, int nz, double *d,
                        double *u, double *v, double *w, double *T,
                        int *istep, double *tt);
int ocfd_write_flow_file(const char *path, int nx, int ny, int nz,
                         const double *d, const double *u, const double *v,
                         const double *w, const double *T, int istep,
                         double tt);
/* nfield 3-D arrays written/read back to back as Fortran records
 * (OCFD3d-Mesh.dat: 3 fields; OCFD3d-Jacobi.dat: 13 fields,
 * src/OCFD_Comput_Jacobian3d.cu:122-215). */
int ocfd_read_records(const char *path, int nx, int ny, int nz, int nfield,
                      double *out);
int ocfd_write_records(const char *path, int nx, int ny, int nz, int nfield,
                       const double *in);

/* ---- explicit context API ---------------------------------------------- */

/* Allocate a context on HIP device `device`.  Replaces opencfd_mem_init_all
 * (src/OCFD_init.cu:134-235, 414-488) and opencfd_para_init
 * (src/OCFD_init.cu:381-410). */
int ocfd_create(const ocfd_config *cfg, int device, ocfd_ctx **out);
int ocfd_destroy(ocfd_ctx *ctx);
const char *ocfd_last_error(void);

/* Grid: mesh coordinates x,y,z (interior, nx*ny*nz each).  Metrics are
 * computed on the GPU as Init_Jacobian3d/comput_Jac3d do
 * (src/OCFD_Comput_Jacobian3d.cu:30-424). */
int ocfd_load_mesh(ocfd_ctx *ctx, const double *x, const double *y,
                   const double *z);
/* Grid: the 13 arrays of OCFD3d-Jacobi.dat in file order Axx Ayy Azz Akx Aky
 * Akz Aix Aiy Aiz Asx Asy Asz Ajac, back to back (13*nx*ny*nz). */
int ocfd_load_jacobi(ocfd_ctx *ctx, const double *arrays13);

/* Flow: primitive variables (interior).  Replaces read_file+memcpy_inner
 * (src/OCFD_IO.c:24-91, src/OCFD_init.cu:83-97). */
int ocfd_load_flow(ocfd_ctx *ctx, const double *d, const double *u,
                   const double *v, const double *w, const double *T,
                   int istep, double tt);
/* Inlet profile for IBC 108: ny rows of (d,u,v,T) as 4*ny doubles laid out
 * [var][j] (flow1d-inlet.dat, src/OCFD_boundary_init.c:376-395).  Required
 * when BC_NPARA[2] (INLET_BOUNDARY) = 1: ocfd_init then fails with
 * OCFD_ERR_STATE without it; otherwise optional, and the uniform (1,1,0,1)
 * profile is used. */
int ocfd_set_inlet_profile(ocfd_ctx *ctx, const double *prof4xny);

/* IBC 124 boundary tables (bc_user_Liftbody3d_init,
 * src/OCFD_boundary_init.c:118-201): inlet5 = d,u,v,w,T at x = 0 as five
 * arrays of nz*ny (y fastest; the reference's flow-inlet-section.dat, used
 * when IF_WITHLEADING = 1), upper5 = five arrays of ny*nx (x fastest; its
 * flow-outboundary.dat, used when IFLAG_UPPERBOUNDARY = 1).  Either may be
 * NULL.  Call before ocfd_init. */
int ocfd_set_bc_tables(ocfd_ctx *ctx, const double *inlet5, const double *upper5);
/* Finish initialisation: prim->cons, boundary-condition set-up, the initial
 * ghost fill, BC and viscosity of NS_solver_real (src/OCFD_init.cu:99-131,
 * src/OCFD_NS_Solver.cu:46-54). */
int ocfd_init(ocfd_ctx *ctx);

/* One right-hand side evaluation into the context's du (du_comput,
 * src/OCFD_Stream.cu:28-54). */
int ocfd_rhs(ocfd_ctx *ctx, int krk);
/* nsteps full time steps (body of NS_solver_real, src/OCFD_NS_Solver.cu:63-167,
 * without show/save/analysis). */
int ocfd_step(ocfd_ctx *ctx, int nsteps);
int ocfd_sync(ocfd_ctx *ctx);

/* Copies back to the host (interior, x fastest). */
int ocfd_get_flow(ocfd_ctx *ctx, double *d, double *u, double *v, double *w,
                  double *T, int *istep, double *tt);
int ocfd_get_state(ocfd_ctx *ctx, double *f5);   /* conserved, [5][nz][ny][nx] */
int ocfd_get_du(ocfd_ctx *ctx, double *du5);     /* last RHS, same layout */
int ocfd_get_pressure(ocfd_ctx *ctx, double *p); /* interior p */
int ocfd_get_amu(ocfd_ctx *ctx, double *amu);
/* 13 metric arrays in OCFD3d-Jacobi.dat order, interior: the mesh
 * coordinates as loaded, then the nine cofactors and Ajac. */
int ocfd_get_metrics(ocfd_ctx *ctx, double *arrays13);
/* Step diagnostics printed every KSTEP_SHOW steps by NS_solver_real
 * (src/OCFD_NS_Solver.cu:124-152, ana_residual and ana_NAN_and_NT in
 * src/OCFD_ana.cu:50-222): the means of the total energy f[4] and of T over
 * the interior, and the numbers of NaN and of negative values of T over the
 * field including its ghost layers.  Any pointer may be null. */
int ocfd_diagnostics(ocfd_ctx *ctx, double *mean_energy, double *mean_T,
                     long long *n_nan_T, long long *n_neg_T);
/* ocfd_diagnostics without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned host
 * memory (code under HIP graph capture elsewhere in the process, serving
 * loops, harnesses that forbid hipMalloc).  The partial sums and counts live
 * in one file-scope device array of OCFD_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends; its address is fetched by ocfd_create);
 * the results return through a 32-byte copy.  Grids whose partials do not
 * fit, (ceil(nx/64) ceil(ny/4) nz) > 523,774 (e.g. above 1024x384x256 by a
 * third), are refused with OCFD_ERR_UNSUPPORTED.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two contexts on
 * different host threads) race.  Same values as ocfd_diagnostics (the counts
 * exactly, the means up to the order of the final sum). */
#define OCFD_NOALLOC_SCRATCH_BYTES (8LL << 20)
int ocfd_diagnostics_noalloc(ocfd_ctx *ctx, double *mean_energy, double *mean_T,
                             long long *n_nan_T, long long *n_neg_T);
/* Write OCFD3d-Jacobi.dat (the 13 arrays above) so that a later run reads the
 * metrics instead of deriving them from OCFD3d-Mesh.dat.  The reference only
 * reads this file (src/OCFD_Comput_Jacobian3d.cu:148-215) and has no writer. */
int ocfd_write_jacobi(ocfd_ctx *ctx, const char *path);
/* hybrid_auto flags: x (nx+1)*ny*nz, y nx*(ny+1)*nz, z nx*ny*(nz+1). */
int ocfd_get_hybrid_flags(ocfd_ctx *ctx, int8_t *fx, int8_t *fy, int8_t *fz);

/* ---- analysis events (OCFD_ana, src/OCFD_ana.cu:588-619) ---------------
 * ocfd_analysis runs the event of `style` with the parameters of ANA_NPARA&
 * number `id` of the configuration (the id-th ANA_EVENT& key), writing its
 * files into the working directory as the reference does:
 *   100  NaN / negative-T check of T over the LAP box (ana_NAN_and_NT,
 *        src/OCFD_ana.cu:105-222): prints each location, returns
 *        OCFD_ERR_NUMERIC on a NaN or on more than 10 negative T.
 *   101  running time average of d,u,v,w,T (ana_time_average, 224-298): the
 *        first call reads ./opencfd.average when present (src/OCFD_IO.c:
 *        94-151), and OCFD%08d.average is written when Istep % KSTEP_SAVE == 0.
 *   102  Scheme_{x,y,z}%08d.dat, the flags of the plane z = NZ/2
 *        (HybridAuto_scheme_IO, src/OCFD_Schemes_hybrid_auto.cu:706-744).
 *   103  Q criterion to Q.dat (records) and Q000000.dat (text) (get_Q,
 *        src/OCFD_ana.cu:299-458; the reference then exits, which is the
 *        caller's business here).
 *   104  Savedata-XY%03d.dat, 105 Savedata-YZ%03d.dat, 106 Savedata-XZ%03d.dat:
 *        planes of d,u,v,w,T appended (ana_saveplaneXY/YZ/XZ, 459-586).
 *   107  fractions of the three hybrid scheme types per direction, printed
 *        (HybridAuto_scheme_Proportion, src/OCFD_Schemes_hybrid_auto.cu:
 *        747-830); nothing when the scheme is not hybrid_auto.
 * Other styles are ignored, as in the reference's switch. */
int ocfd_analysis(ocfd_ctx *ctx, int style, int id);
/* Q criterion at the interior points (nz*ny*nx). */
int ocfd_get_q(ocfd_ctx *ctx, double *q);
/* The time averages of style 101 (interior, nz*ny*nx each), the number of
 * samples in them and their accumulated time; OCFD_ERR_STATE before the
 * first style-101 event. */
int ocfd_get_time_average(ocfd_ctx *ctx, double *d, double *u, double *v,
                          double *w, double *T, int *nsamples, double *tt_avg);
/* The nine numbers style 107 prints: the fractions of type 1, 2, 3 flags in
 * x, then y, then z, with the reference's normalisation (count over the
 * direction's face array / (the two other global sizes) / (N + 1)). */
int ocfd_hybrid_proportions(ocfd_ctx *ctx, double frac9[9]);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on the
 * context's stream and accumulated per kernel name. */
/* ---- learned hybrid sensor (no counterpart in the reference) ----------
 * A 24 -> 32 -> 32 -> 1 MLP (GELU, sigmoid) over 8 stencil features per
 * direction (normalised p and rho jumps, the largest outer p jump, the
 * contravariant velocity jump over c) gives the probability that a node needs
 * WENO.  Parameters: 1889 floats, W1[32][24] b1[32] W2[32][32] b2[32] W3[32]
 * b3.  With ocfd_set_sensor the hybrid_auto flags become: prob < lo -> UP7,
 * prob > hi -> max(plain kp, WENO7), otherwise the plain style-2 kp.
 * params = NULL switches back to the plain sensor. */
int ocfd_set_sensor(ocfd_ctx *ctx, const float *params, float lo, float hi);
/* Where the learned sensor runs (no counterpart in the reference), after the
 * first stage's periodic fill: on != 0 (the default) inside the x sweep of
 * the first stage (the BF16-MFMA forward of the 64 nodes of each x-sweep
 * wave, with the ghost planes j = -1 and k = -1 in a pass of their own);
 * on = 0 in its own pass over all nodes.  The same flags and probabilities
 * either way.  CHARTERIC = 1 always uses the separate pass. */
int ocfd_set_sensor_fused(ocfd_ctx *ctx, int on);
/* Interior probabilities of the last learned-sensor pass (nz*ny*nx floats). */
int ocfd_get_sensor_prob(ocfd_ctx *ctx, float *prob);

/* Trainer: samples (24 features + plain label), parameters and Adam moments
 * on one device.  Training is FP32 and deterministic for a given seed. */
typedef struct ocfd_mlp ocfd_mlp;
int ocfd_mlp_create(int device, long long capacity, ocfd_mlp **out);
int ocfd_mlp_destroy(ocfd_mlp *m);
int ocfd_mlp_set_params(ocfd_mlp *m, const float *params); /* resets Adam */
int ocfd_mlp_get_params(ocfd_mlp *m, float *params);
long long ocfd_mlp_sample_count(ocfd_mlp *m);
int ocfd_mlp_clear(ocfd_mlp *m);
/* Append the interior nodes whose indices are all multiples of stride, from
 * the context's current state, with the plain sensor's label (kp > 1). */
int ocfd_mlp_collect(ocfd_mlp *m, ocfd_ctx *ctx, int stride);
int ocfd_mlp_put_samples(ocfd_mlp *m, const float *X, const float *y, long long n);
int ocfd_mlp_get_samples(ocfd_mlp *m, long long first, long long n, float *X, float *y);
/* Batch-mean weighted BCE and its gradient (1889 floats) on samples
 * [first, first+n); positives weighted by wpos. */
int ocfd_mlp_grad(ocfd_mlp *m, long long first, int n, float wpos, float *grads, float *loss);
/* FP32 forward probabilities of samples [first, first+n). */
int ocfd_mlp_predict(ocfd_mlp *m, long long first, int n, float *prob);
/* Adam (lr; beta 0.9, 0.999) over shuffled mini-batches; epoch_loss (may be
 * NULL) receives the mean loss of each epoch. */
int ocfd_mlp_train(ocfd_mlp *m, int epochs, int batch, float lr, float wpos, unsigned long long seed,
                   float *epoch_loss);
/* ocfd_set_sensor with the trainer's current parameters. */
int ocfd_set_sensor_from(ocfd_ctx *ctx, ocfd_mlp *m, float lo, float hi);

int ocfd_profile_enable(ocfd_ctx *ctx, int on);
int ocfd_profile_count(ocfd_ctx *ctx);
/* i-th entry: name (NUL-terminated, <=63 chars), launches, total ms. */
int ocfd_profile_entry(ocfd_ctx *ctx, int i, char *name, int *launches,
                       double *total_ms);
int ocfd_profile_reset(ocfd_ctx *ctx);
/* Wall time of the last ocfd_step call measured with HIP events, ms. */
double ocfd_last_step_ms(ocfd_ctx *ctx);

/* ---- reference-named entry points (process-wide context) -------------- */

/* read_parameters (src/parameters.c:197): reads ./opencfd-scu.in. */
void read_parameters(void);
/* opencfd_mem_init_all (src/OCFD_init.cu:134). */
void opencfd_mem_init_all(void);
/* init (src/OCFD_init.cu:26): metrics from ./OCFD3d-Jacobi.dat or
 * ./OCFD3d-Mesh.dat, flow from Opencfd.msg/opencfd.dat, IBC set-up. */
void init(void);
/* NS_solver_real (src/OCFD_NS_Solver.cu:39): the time loop to END_TIME with
 * the ANA_EVENT analysis events, the KSTEP_SHOW diagnostics and the
 * KSTEP_SAVE output files, in the reference's order. */
void NS_solver_real(void);
/* du_comput (src/OCFD_Stream.cu:28). */
void du_comput(int KRK);
/* OCFD_time_advance (src/OCFD_time.cu:96). */
void OCFD_time_advance(int KRK);
/* get_duvwT (src/commen_kernel.cu:111). */
void get_duvwT(void);
/* get_Amu (src/commen_kernel.cu:133). */
void get_Amu(void);
/* OCFD_bc (src/OCFD_boundary.cu:18). */
void OCFD_bc(void);
/* filtering (src/OCFD_filtering.cu:75); the pointers are the reference's host
 * mirrors and are not dereferenced. */
void filtering(double *pf, double *pf0, double *pp);
/* OCFD_ana (src/OCFD_ana.cu:588): ocfd_analysis on the process-wide context;
 * a failed style 100 aborts the process as MPI_Abort does. */
void OCFD_ana(int style, int ID);
/* opencfd_mem_finalize_all (src/OCFD_init.cu:238). */
void opencfd_mem_finalize_all(void);
/* Access to the process-wide context used by the functions above. */
ocfd_ctx *ocfd_global_ctx(void);

#ifdef __cplusplus
}
#endif
#endif /* OCFD_H */
