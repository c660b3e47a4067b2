// k_point.hip - pointwise and boundary kernels (gfx950): RK stage fused with
// cons->prim and viscosity, periodic face fills, IBC 108, the face refresh,
// the hybrid_auto sensor, the shock-capturing filter and the metric kernels.
//
// S (double or float) is the storage type of the primitives and the
// viscosity; the conserved state, the metrics and all arithmetic here are
// double.
#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

// ---------------------------------------------------------------------------
// Periodic face fill, exchange_boundary_{x,y,z}_packed_dev at one rank
// (src/OCFD_mpi_dev.cu:186-310): ghost[n..n+3] <- interior[0..3],
// ghost[-4..-1] <- interior[n-4..n-1], other two indices interior only.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(Geo g, T* base, int nf, long long fs, int dir) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // position in the 8 x a x b slab
    const int na = dir == 0 ? g.ny : g.nx;
    const int nb = dir == 2 ? g.ny : g.nz;
    const long long tot = 8LL * na * nb;
    if (t >= tot) return;
    int l, a, b;
    if (dir == 0) { l = t % 8; a = (t / 8) % na; b = t / (8 * na); }
    else { a = t % na; l = (t / na) % 8; b = t / (8 * na); }
    const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
    int dst, src;
    if (l < 4) { dst = n + l; src = l; } else { dst = l - 8; src = n + l - 8; }
    long long qd, qs;
    if (dir == 0) { qd = gidx(g, dst, a, b); qs = gidx(g, src, a, b); }
    else if (dir == 1) { qd = gidx(g, a, dst, b); qs = gidx(g, a, src, b); }
    else { qd = gidx(g, a, b, dst); qs = gidx(g, a, b, src); }
    for (int m = 0; m < nf; m++) base[m * fs + qd] = base[m * fs + qs];
}

// ---------------------------------------------------------------------------
// SSP-RK3 stage (OCFD_time_advance_ker1/2/3, src/OCFD_time.cu:44-94) fused
// with cons->prim + p (cons_to_pri_kernel, src/commen_kernel.cu:77-109) and
// Sutherland's law (get_Amu_kernal, src/commen_kernel.cu:122-131).
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_stage(StageArgsT<S> A) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= A.g.nx || j >= A.g.ny) return;
    const long long q = gidx(A.g, i, j, k);
    // du, plus its viscous part dv (components 1..4) when that is kept apart
    // du and dv are read once, as streaming (non-temporal) loads
    double du[5];
#pragma unroll
    for (int m = 0; m < 5; m++) du[m] = double(__builtin_nontemporal_load(&A.du[m * A.fs + q]));
    if (A.dv) {
#pragma unroll
        for (int m = 1; m < 5; m++) du[m] += double(__builtin_nontemporal_load(&A.dv[(m - 1) * A.fs + q]));
    }
    // the sweeps and the divergence leave out the node's Ajac, by which the
    // reference multiplies each of their terms: it is applied to the sum here
    const double ajac = __builtin_nontemporal_load(&A.jac[q]);
#pragma unroll
    for (int m = 0; m < 5; m++) du[m] *= ajac;
    stage_point(A, q, du);
}

// ---------------------------------------------------------------------------
// prim->cons of one point (pri_to_cons_kernel, src/commen_kernel.cu:32-64),
// used by the face refresh of OCFD_bc (src/OCFD_boundary.cu:40-91) and at
// initialisation.
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void p2c(const BCArgsT<S>& A, long long q) {
    const double d = A.d[q], u = A.u[q], v = A.v[q], w = A.w[q], T = A.T[q];
    A.f[q] = d;
    A.f[A.fs + q] = u * d;
    A.f[2 * A.fs + q] = v * d;
    A.f[3 * A.fs + q] = w * d;
    A.f[4 * A.fs + q] = d * ((u * u + v * v + w * w) * 0.5 + A.c.cv * T);
}
// ---------------------------------------------------------------------------
// IBC 108, bc_user_Compression_conner (src/OCFD_boundary_compression_conner.cu
// 27-179: inlet, top, wall velocity, wall temperature/density with mirrored
// ghosts), get_Amu on the faces it changes, and the face refresh.  By
// default one launch does all of it (k_bc108_faces below); the three
// launches here keep the reference's order where points are shared, and
// are the path the merged kernel is checked against: (A) inlet and top (they
// write the same values where they overlap), (B) per wall column the wall
// velocity then the wall temperature/density and mirrored ghosts (which read
// the inlet column at i = 0, so B follows A), (C) per face point the viscosity
// (faces x=0, y=0, y=ny-1, when IBC 108 ran) and prim->cons on all six faces.
template <typename S>
__global__ void k_bc_inlet_top(BCArgsT<S> A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    const int ny = A.g.ny;
    if (t < ny) {
        const int j = t;
        for (int i = -4; i <= 0; i++) {
            const long long q = gidx(A.g, i, j, k);
            A.d[q] = S(A.ub1[j]); A.u[q] = S(A.ub1[j + ny]); A.v[q] = S(A.ub1[j + 2 * ny]); A.w[q] = S(0);
            A.T[q] = S(A.ub1[j + 3 * ny]);
        }
    } else {
        const int i = A.x_begin + (t - ny);
        if (i >= A.g.nx) return;
        const long long q = gidx(A.g, i, ny - 1, k);
        A.d[q] = S(A.ub1[ny - 1]); A.u[q] = S(A.ub1[ny + ny - 1]); A.v[q] = S(A.ub1[2 * ny + ny - 1]); A.w[q] = S(0);
        A.T[q] = S(A.ub1[3 * ny + ny - 1]);
    }
}
template <typename S>
__global__ void k_bc_wall(BCArgsT<S> A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= A.g.nx) return;
    const Geo& g = A.g;
    {  // do_wall_dist_kernel
        const long long q = gidx(g, i, 0, k);
        A.u[q] = S(0);
        A.v[q] = S(A.epsl * A.fx[i] * A.gz[k] * A.ht);
        A.w[q] = S(0);
    }
    if (A.wall_not_normal != 0) return;
    const long long q0 = gidx(g, i, 0, k), q1 = gidx(g, i, 1, k), q2 = gidx(g, i, 2, k), q3 = gidx(g, i, 3, k);
    auto D = [&](long long q) { return double(A.d[q]); };
    auto T = [&](long long q) { return double(A.T[q]); };
    if (A.tw > 0) {
        const double pw = (18. * D(q1) * T(q1) - 9. * D(q2) * T(q2) + 2. * D(q3) * T(q3)) / 11.;
        A.T[q0] = S(A.tw);
        A.d[q0] = S(pw / T(q0));
        for (int l = 0; l < 4; l++) {  // ghost j = l-4 mirrors interior j = 4-l
            const long long qg = gidx(g, i, l - 4, k), qm = gidx(g, i, 4 - l, k);
            A.d[qg] = A.d[qm];
            A.u[qg] = -A.u[qm];
            A.v[qg] = -A.v[qm];
            A.w[qg] = -A.w[qm];
            A.T[qg] = A.T[qm];
        }
    } else {
        A.T[q0] = S((18. * T(q1) - 9. * T(q2) + 2. * T(q3)) / 11.);
        const double pw = (18. * D(q1) * T(q1) - 9. * D(q2) * T(q2) + 2. * D(q3) * T(q3)) / 11.;
        A.d[q0] = S(pw / T(q0));
    }
}
template <typename S>
__global__ void k_faces(BCArgsT<S> A) {
    const Geo& g = A.g;
    const long long nyz = (long long)g.ny * g.nz, nxz = (long long)g.nx * g.nz, nxy = (long long)g.nx * g.ny;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    bool mu = false;
    if (t < 2 * nyz) {
        const int side = (int)(t / nyz);
        const long long r = t % nyz;
        i = side ? g.nx - 1 : 0; j = (int)(r % g.ny); k = (int)(r / g.ny);
        mu = (A.mu_faces >> side) & 1;
    } else if ((t -= 2 * nyz) < 2 * nxz) {
        const int side = (int)(t / nxz);
        const long long r = t % nxz;
        i = (int)(r % g.nx); j = side ? g.ny - 1 : 0; k = (int)(r / g.nx);
        mu = (A.mu_faces >> (2 + side)) & 1;
    } else if ((t -= 2 * nxz) < 2 * nxy) {
        const int side = (int)(t / nxy);
        const long long r = t % nxy;
        i = (int)(r % g.nx); j = (int)(r / g.nx); k = side ? g.nz - 1 : 0;
        mu = (A.mu_faces >> (4 + side)) & 1;
    } else {
        return;
    }
    const long long q = gidx(g, i, j, k);
    if (mu && A.amu) {
        const double T = double(A.T[q]);
        A.amu[q] = S(A.c.amu_c0 * sqrt(T * T * T) / (A.c.tsb + T));
    }
    p2c(A, q);
}

// ---------------------------------------------------------------------------
// IBC 108 and the face refresh in one launch.  A thread of a face point
// derives the point's primitives after the boundary condition from the
// state before it, by the rules and in the order of (A) and (B) above (inlet
// columns i = -4..0, then the top row from x_begin, then the wall row j = 0
// with rows 1..3 as (A) left them), writes them, and refreshes the point as
// (C) does.  The ghost layers the condition writes (inlet i = -4..-1, wall
// j = -4..-1) have threads of their own.  A point on two faces is done by
// two threads with the same values; a thread reads from memory only points
// the condition leaves unchanged.
// ---------------------------------------------------------------------------
template <typename S> struct Prim5 { S d, u, v, w, T; };

// primitives of (i, j, k) after (A)
template <typename S>
__device__ __forceinline__ Prim5<S> bc108_after_a(const BCArgsT<S>& A, int i, int j, int k) {
    const int ny = A.g.ny;
    Prim5<S> p;
    if (i <= 0 && i >= -4 && j >= 0 && j < ny) {
        p.d = S(A.ub1[j]); p.u = S(A.ub1[j + ny]); p.v = S(A.ub1[j + 2 * ny]); p.w = S(0); p.T = S(A.ub1[j + 3 * ny]);
    } else if (j == ny - 1 && i >= A.x_begin && i < A.g.nx) {
        p.d = S(A.ub1[ny - 1]); p.u = S(A.ub1[ny + ny - 1]); p.v = S(A.ub1[2 * ny + ny - 1]); p.w = S(0);
        p.T = S(A.ub1[3 * ny + ny - 1]);
    } else {
        const long long q = gidx(A.g, i, j, k);
        p.d = A.d[q]; p.u = A.u[q]; p.v = A.v[q]; p.w = A.w[q]; p.T = A.T[q];
    }
    return p;
}
// primitives of (i, j, k) after (A) and (B)
template <typename S>
__device__ __forceinline__ Prim5<S> bc108_final(const BCArgsT<S>& A, int i, int j, int k) {
    Prim5<S> p = bc108_after_a(A, i, j, k);
    if (j != 0 || i < 0 || i >= A.g.nx) return p;
    p.u = S(0);
    p.v = S(A.epsl * A.fx[i] * A.gz[k] * A.ht);
    p.w = S(0);
    if (A.wall_not_normal != 0) return p;
    const Prim5<S> p1 = bc108_after_a(A, i, 1, k), p2 = bc108_after_a(A, i, 2, k), p3 = bc108_after_a(A, i, 3, k);
    const double D1 = p1.d, D2 = p2.d, D3 = p3.d, T1 = p1.T, T2 = p2.T, T3 = p3.T;
    if (A.tw > 0) {
        const double pw = (18. * D1 * T1 - 9. * D2 * T2 + 2. * D3 * T3) / 11.;
        p.T = S(A.tw);
        p.d = S(pw / double(p.T));
    } else {
        p.T = S((18. * T1 - 9. * T2 + 2. * T3) / 11.);
        const double pw = (18. * D1 * T1 - 9. * D2 * T2 + 2. * D3 * T3) / 11.;
        p.d = S(pw / double(p.T));
    }
    return p;
}
template <typename S>
__device__ __forceinline__ void put_prim(const BCArgsT<S>& A, long long q, const Prim5<S>& p) {
    A.d[q] = p.d; A.u[q] = p.u; A.v[q] = p.v; A.w[q] = p.w; A.T[q] = p.T;
}
template <typename S>
__global__ void k_bc108_faces(BCArgsT<S> A) {
    const Geo& g = A.g;
    const long long nyz = (long long)g.ny * g.nz, nxz = (long long)g.nx * g.nz, nxy = (long long)g.nx * g.ny;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // inlet ghosts i = -4..-1: the four of a row on consecutive threads, 32
    // contiguous bytes per field (measured: the launch 209 -> 197 us on the
    // 1024x384x256 swept ramp, against one row per thread at a fixed i)
    if (t < 4 * nyz) {
        const int i = -4 + (int)(t & 3);
        const long long r = t >> 2;
        const int j = (int)(r % g.ny), k = (int)(r / g.ny);
        put_prim(A, gidx(g, i, j, k), bc108_after_a(A, i, j, k));
        return;
    }
    t -= 4 * nyz;
    const bool ghosts = A.tw > 0 && A.wall_not_normal == 0;
    if (t < 4 * nxz) {  // wall ghosts j = -4..-1 mirror rows 4..1 (rho, T even; velocity odd)
        if (!ghosts) return;
        const int l = (int)(t / nxz);
        const long long r = t % nxz;
        const int i = (int)(r % g.nx), k = (int)(r / g.nx);
        const Prim5<S> m = bc108_after_a(A, i, 4 - l, k);
        Prim5<S> p;
        p.d = m.d; p.u = -m.u; p.v = -m.v; p.w = -m.w; p.T = m.T;
        put_prim(A, gidx(g, i, l - 4, k), p);
        return;
    }
    t -= 4 * nxz;
    int i, j, k;
    bool mu = false;
    if (t < 2 * nyz) {
        const int side = (int)(t / nyz);
        const long long r = t % nyz;
        i = side ? g.nx - 1 : 0; j = (int)(r % g.ny); k = (int)(r / g.ny);
        mu = (A.mu_faces >> side) & 1;
    } else if ((t -= 2 * nyz) < 2 * nxz) {
        const int side = (int)(t / nxz);
        const long long r = t % nxz;
        i = (int)(r % g.nx); j = side ? g.ny - 1 : 0; k = (int)(r / g.nx);
        mu = (A.mu_faces >> (2 + side)) & 1;
    } else if ((t -= 2 * nxz) < 2 * nxy) {
        const int side = (int)(t / nxy);
        const long long r = t % nxy;
        i = (int)(r % g.nx); j = (int)(r / g.nx); k = side ? g.nz - 1 : 0;
        mu = (A.mu_faces >> (4 + side)) & 1;
    } else {
        return;
    }
    const long long q = gidx(g, i, j, k);
    const Prim5<S> p = bc108_final(A, i, j, k);
    put_prim(A, q, p);
    const double T = p.T;
    if (mu && A.amu) A.amu[q] = S(A.c.amu_c0 * sqrt(T * T * T) / (A.c.tsb + T));
    const double d = p.d, u = p.u, v = p.v, w = p.w;
    A.f[q] = d;
    A.f[A.fs + q] = u * d;
    A.f[2 * A.fs + q] = v * d;
    A.f[3 * A.fs + q] = w * d;
    A.f[4 * A.fs + q] = d * ((u * u + v * v + w * w) * 0.5 + A.c.cv * T);
}

// ---------------------------------------------------------------------------
// IBC 124, bc_user_Liftbody3d (src/OCFD_boundary_Liftbody3D.cu:32-294):
// inlet (x ghosts and i = 0 from the inlet table), upper face (k = nz-1: free
// stream at the angle of attack, or the upper table, plus the upper
// disturbance on u), wall (k = 0: blowing/suction times ht when a
// disturbance table exists, else no slip; T = Tw with the density from
// 2nd-order extrapolated pressure, or T and rho extrapolated when Tw <= 0),
// then the symmetry planes (y ghosts mirrored about j = 0 and j = ny-1, v
// odd).  One launch per step of the reference, in its order.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void k_bc124_inlet(BCArgsT<S> A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    const Geo& g = A.g;
    if (j >= g.ny) return;
    const long long t = (long long)g.nz * g.ny, o = (long long)k * g.ny + j;
    for (int i = -4; i <= 0; i++) {
        const long long q = gidx(g, i, j, k);
        A.d[q] = S(A.in124[o]);
        A.u[q] = S(A.in124[t + o]);
        A.v[q] = S(A.in124[2 * t + o]);
        A.w[q] = S(A.in124[3 * t + o]);
        A.T[q] = S(A.in124[4 * t + o]);
    }
}
template <typename S>
__global__ void k_bc124_upper(BCArgsT<S> A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Geo& g = A.g;
    if (i >= g.nx) return;
    const long long q = gidx(g, i, j, g.nz - 1), t = (long long)g.ny * g.nx, o = (long long)j * g.nx + i;
    const double dist = A.udist124[o];
    if (A.up124) {  // do_u2d_upper_kernel
        A.d[q] = S(A.up124[o]);
        A.u[q] = S(A.up124[t + o] + dist);
        A.v[q] = S(A.up124[2 * t + o]);
        A.w[q] = S(A.up124[3 * t + o]);
        A.T[q] = S(A.up124[4 * t + o]);
    } else {  // do_u_dist_upper_kernel
        A.d[q] = S(1.0);
        A.u[q] = S(A.cos_aoa + dist);
        A.v[q] = S(0.0);
        A.w[q] = S(A.sin_aoa);
        A.T[q] = S(1.0);
    }
}
template <typename S>
__global__ void k_bc124_wall(BCArgsT<S> A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Geo& g = A.g;
    if (i >= g.nx) return;
    const long long q0 = gidx(g, i, j, 0), q1 = gidx(g, i, j, 1), q2 = gidx(g, i, j, 2);
    const long long t = (long long)g.ny * g.nx, o = (long long)j * g.nx + i;
    if (A.coeff124) {  // do_wall_kernel_*_V
        A.u[q0] = S(A.coeff124[o] * A.ht);
        A.v[q0] = S(A.coeff124[t + o] * A.ht);
        A.w[q0] = S(A.coeff124[2 * t + o] * A.ht);
    } else {  // do_wall_kernel_*_NV
        A.u[q0] = S(0.0);
        A.v[q0] = S(0.0);
        A.w[q0] = S(0.0);
    }
    if (A.tw > 0.0) {
        A.T[q0] = S(A.tw);
        A.d[q0] = S((4.0 * double(A.d[q1]) * double(A.T[q1]) - double(A.d[q2]) * double(A.T[q2])) / (3.0 * A.tw));
    } else {
        A.T[q0] = S((4.0 * double(A.T[q1]) - double(A.T[q2])) / 3.0);
        A.d[q0] = S((4.0 * double(A.d[q1]) - double(A.d[q2])) / 3.0);
    }
}
template <typename T>
__global__ void k_mirror_y(Geo g, T* f, double sign) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y, k = blockIdx.z;
    if (i >= g.nx) return;
    // l = 0..3: ghost j = -1-l from interior 1+l; l = 4..7: ghost j = ny+(l-4)
    // from interior ny-2-(l-4)
    const int jg = l < 4 ? -1 - l : g.ny + (l - 4);
    const int js = l < 4 ? 1 + l : g.ny - 2 - (l - 4);
    f[gidx(g, i, jg, k)] = T(sign * double(f[gidx(g, i, js, k)]));
}

// ---------------------------------------------------------------------------
// hybrid_auto style 2 (Comput_P + Comput_Scheme_point_{x,y,z}_Jameson_kernel,
// src/OCFD_Schemes_hybrid_auto.cu:48-68, 611-703): p = p00*rho*T on the fly,
// dp = sum over directions |p- - 2p + p+| / (p- + 2p + p+), kp = 1 + (dp>i1) +
// (dp>i2).  Node (i,j,k) with i,j,k >= -1 serves the x flags when j,k >= 0,
// the y flags when i,k >= 0 and the z flags when i,j >= 0.
// ---------------------------------------------------------------------------
#ifndef OCFD_SENSOR_KZ
#define OCFD_SENSOR_KZ 16  // nodes along z per block of the sensor
#endif
constexpr int SENSOR_KZ = OCFD_SENSOR_KZ;
// A block owns a 64 x 4 tile of (i, j) nodes and marches through SENSOR_KZ
// planes: p of the tile's plane k and its one-node rim is staged in LDS (x
// and y neighbours), each column keeps p at k-1, k, k+1 in registers (z
// neighbours), so every p is formed once per block instead of seven times.
// Same arithmetic as p00*rho*T and the dp sum in x, y, z order.
template <typename S>
__global__ void __launch_bounds__(256) k_sensor(Geo g, Consts c, const S* __restrict__ d,
                                                const S* __restrict__ T, int8_t* fx, int8_t* fy, int8_t* fz) {
    constexpr int W = 66;                      // tile row with its rim
    __shared__ double pl[6 * W];               // rows j0-1 .. j0+4, columns i0-1 .. i0+64
    const Blk B = xcd_block();
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int i0 = B.x * 64 - 1, j0 = B.y * 4 - 1;
    const int i = i0 + tx, j = j0 + ty;
    const int kb = B.z * SENSOR_KZ - 1, ke = min(kb + SENSOR_KZ, g.nz);
    const bool own = i < g.nx && j < g.ny;
    auto P = [&](long long o) { return c.p00 * double(d[o]) * double(T[o]); };
    // the lane's own column: its real position inside the LAP box (an own
    // lane's neighbours are at most nx, ny), else any valid column (unused)
    const long long col = (i <= g.nx + 3 && j <= g.ny + 3) ? gidx(g, i, j, 0) : gidx(g, 0, 0, 0);
    // rim slots: lanes 0..131 the rows j0-1 and j0+4 (66 columns each),
    // lanes 132..139 the columns i0-1 and i0+64 of the 4 tile rows
    int hs = -1;
    long long ho = -1;
    if (tid < 2 * W) {
        const int r = tid / W, xx = tid % W, gi = i0 - 1 + xx, gj = r ? j0 + 4 : j0 - 1;
        hs = (r ? 5 : 0) * W + xx;
        if (gi <= g.nx + 3 && gj <= g.ny + 3) ho = gidx(g, gi, gj, 0);
    } else if (tid < 2 * W + 8) {
        const int r = tid - 2 * W, side = r >> 2, row = r & 3, gi = side ? i0 + 64 : i0 - 1, gj = j0 + row;
        hs = (row + 1) * W + (side ? W - 1 : 0);
        if (gi <= g.nx + 3 && gj <= g.ny + 3) ho = gidx(g, gi, gj, 0);
    }
    if (kb >= ke) return;
    double pm = P(col + (long long)(kb - 1) * g.sz), pc = P(col + (long long)kb * g.sz);
    double pp = P(col + (long long)(kb + 1) * g.sz);
    double hv = ho >= 0 ? P(ho + (long long)kb * g.sz) : 0.0;
    for (int k = kb; k < ke; k++) {
        pl[(ty + 1) * W + tx + 1] = pc;
        if (hs >= 0) pl[hs] = hv;
        __syncthreads();
        // the next plane's values are in flight while this one is decided
        double pn = 0.0, hn = 0.0;
        if (k + 1 < ke) {
            pn = P(col + (long long)(k + 2) * g.sz);
            hn = ho >= 0 ? P(ho + (long long)(k + 1) * g.sz) : 0.0;
        }
        if (own) {
            const long long q = gidx(g, i, j, k);
            double am = pl[(ty + 1) * W + tx], ap = pl[(ty + 1) * W + tx + 2];
            double dp0 = fabs(-am + 2 * pc - ap) / (am + 2 * pc + ap);
            am = pl[ty * W + tx + 1]; ap = pl[(ty + 2) * W + tx + 1];
            dp0 += fabs(-am + 2 * pc - ap) / (am + 2 * pc + ap);
            dp0 += fabs(-pm + 2 * pc - pp) / (pm + 2 * pc + pp);
            int kp = 1;
            if (dp0 > c.hy_i1) kp += 1;
            if (dp0 > c.hy_i2) kp += 1;
            if (j >= 0 && k >= 0) fx[q] = (int8_t)kp;
            if (i >= 0 && k >= 0) fy[q] = (int8_t)kp;
            if (i >= 0 && j >= 0) fz[q] = (int8_t)kp;
        }
        __syncthreads();
        pm = pc; pc = pp; pp = pn; hv = hn;
    }
}

// ---------------------------------------------------------------------------
// Shock-capturing filter (filter_{x,y}3d_shock_kernel, src/OCFD_filtering.cu:
// 366-416, 451-495) for one direction over a box of output points.
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ double filt_sigma(const S* __restrict__ P, long long L, long long st, double rth) {
    const double p_m2 = P[L - 2 * st], p_m1 = P[L - st], p_00 = P[L], p_p1 = P[L + st], p_p2 = P[L + 2 * st];
    const double dp0 = 0.25 * (-p_p1 + 2.0 * p_00 - p_m1);
    const double dp1 = 0.25 * (-p_p2 + 2.0 * p_p1 - p_00);
    const double dp2 = 0.25 * (-p_00 + 2.0 * p_m1 - p_m2);
    double ri = 0.5 * ((dp0 - dp1) * (dp0 - dp1) + (dp0 - dp2) * (dp0 - dp2)) / (p_00 * p_00) + 1e-16;
    ri = 1.0 - rth / ri;
    return 0.5 * (ri + fabs(ri));
}
template <typename S>
__global__ void __launch_bounds__(256) k_filter(FilterArgsT<S> A) {
    const Blk B = xcd_block();
    const int i = A.lo[0] + B.x * 64 + threadIdx.x, j = A.lo[1] + B.y * 4 + threadIdx.y;
    const int k = A.lo[2] + B.z;
    if (i >= A.hi[0] || j >= A.hi[1] || k >= A.hi[2]) return;
    const Geo& g = A.g;
    const long long L = gidx(g, i, j, k);
    const long long st = A.dir == 0 ? 1 : A.dir == 1 ? (long long)g.ldx : g.sz;
    const double hm = filt_sigma(A.P, L - st, st, A.rth);
    const double h0 = filt_sigma(A.P, L, st, A.rth);
    const double hp = filt_sigma(A.P, L + st, st, A.rth);
    const double Sc1 = 0.5 * (h0 + hp) * A.s0;
    const double Sc2 = 0.5 * (h0 + hm) * A.s0;
    const double c1 = -0.2103830, c2 = 0.0396170;
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const double* F = A.flap + (long long)m * A.fs;
        const double r = F[L] - (Sc1 * (c1 * (F[L + st] - F[L]) + c2 * (F[L + 2 * st] - F[L - st])) -
                                 Sc2 * (c1 * (F[L] - F[L - st]) + c2 * (F[L + st] - F[L - 2 * st])));
        __builtin_nontemporal_store(r, &A.f[(long long)m * A.fs + L]);  // streamed (125 -> 96 us)
    }
}

// ---------------------------------------------------------------------------
// Metrics (init): OCFD_d?0_CD6_kernel (src/OCFD_Schemes.cu:1236-1300) with the
// OCFD_D?0_bound_kernel_m/p closures (src/OCFD_bound_Scheme.cu:189-349), then
// comput_Jac3d_kernal (src/OCFD_Comput_Jacobian3d.cu:250-391).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_metric_deriv(Geo g, const double* __restrict__ a, int dir, int bound,
                                                      double* out) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const long long q = gidx(g, i, j, k);
    const long long st = dir == 0 ? 1 : dir == 1 ? (long long)g.ldx : g.sz;
    const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
    const int p = dir == 0 ? i : dir == 1 ? j : k;
    const double h = dir == 0 ? g.hx : dir == 1 ? g.hy : g.hz;
    auto F = [&](int o) { return a[q + (long long)o * st]; };
    double r = (45.0 * (F(1) - F(-1)) - 9.0 * (F(2) - F(-2)) + (F(3) - F(-3))) / (60.0 * h);
    if (bound && n >= 4) {
        // left closures written relative to node 0 (F(o) is node p+o)
        if (p == 0) r = (F(1) - F(0)) / h;
        else if (p == 1) r = (F(1) - F(-1)) * 0.5 / h;
        else if (p == 2) r = (F(-2) - 8.0 * F(-1) + 8.0 * F(1) - F(2)) / (12.0 * h);
        else if (p == 3) r = (F(3) - F(-3) - 9.0 * (F(2) - F(-2)) + 45.0 * (F(1) - F(-1))) / (60.0 * h);
        // right closures (run after the left ones in the reference)
        if (p == n - 1) r = dir == 1 ? -(F(-1) - F(0)) / h : (F(0) - F(-1)) / h;
        else if (p == n - 2) r = dir == 1 ? -(F(-1) - F(1)) * 0.5 / h : (F(1) - F(-1)) * 0.5 / h;
        else if (p == n - 3)
            r = dir == 1 ? -(F(2) - 8.0 * F(1) + 8.0 * F(-1) - F(-2)) / (12.0 * h)
                         : (F(-2) - 8.0 * F(-1) + 8.0 * F(1) - F(2)) / (12.0 * h);
        else if (p == n - 4)
            r = dir == 1 ? -(F(-3) - F(3) - 9.0 * (F(-2) - F(2)) + 45.0 * (F(-1) - F(1))) / (60.0 * h)
                         : (F(3) - F(-3) - 9.0 * (F(2) - F(-2)) + 45.0 * (F(1) - F(-1))) / (60.0 * h);
    }
    out[q] = r;
}

struct MetricOut { double* M[M_COUNT]; };

__global__ void __launch_bounds__(256) k_cofactors(Geo g, const double* __restrict__ D, long long fs,
                                                   MetricOut O) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const long long q = gidx(g, i, j, k);
    // D: xi xj xk yi yj yk zi zj zk
    const double xi1 = D[q], xj1 = D[fs + q], xk1 = D[2 * fs + q];
    const double yi1 = D[3 * fs + q], yj1 = D[4 * fs + q], yk1 = D[5 * fs + q];
    const double zi1 = D[6 * fs + q], zj1 = D[7 * fs + q], zk1 = D[8 * fs + q];
    double v[M_COUNT];
    v[M_JAC] = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                      xi1 * zj1 * yk1);
    v[M_KX] = (yj1 * zk1 - zj1 * yk1);
    v[M_KY] = (zj1 * xk1 - xj1 * zk1);
    v[M_KZ] = (xj1 * yk1 - yj1 * xk1);
    v[M_IX] = (yk1 * zi1 - zk1 * yi1);
    v[M_IY] = (zk1 * xi1 - xk1 * zi1);
    v[M_IZ] = (xk1 * yi1 - yk1 * xi1);
    v[M_SX] = (yi1 * zj1 - zi1 * yj1);
    v[M_SY] = (zi1 * xj1 - xi1 * zj1);
    v[M_SZ] = (xi1 * yj1 - yi1 * xj1);
#pragma unroll
    for (int m = 0; m < M_COUNT; m++) {
        double* a = O.M[m];
        a[q] = v[m];
        // zeroth-order ghost copies on the six faces
        if (i == 0) for (int l = 1; l <= 4; l++) a[gidx(g, -l, j, k)] = v[m];
        if (i == g.nx - 1) for (int l = 1; l <= 4; l++) a[gidx(g, i + l, j, k)] = v[m];
        if (j == 0) for (int l = 1; l <= 4; l++) a[gidx(g, i, -l, k)] = v[m];
        if (j == g.ny - 1) for (int l = 1; l <= 4; l++) a[gidx(g, i, j + l, k)] = v[m];
        if (k == 0) for (int l = 1; l <= 4; l++) a[gidx(g, i, j, -l)] = v[m];
        if (k == g.nz - 1) for (int l = 1; l <= 4; l++) a[gidx(g, i, j, k + l)] = v[m];
    }
}

// Whole-interior prim->cons (src/OCFD_init.cu:99-110) and viscosity
// (get_Amu, src/commen_kernel.cu:133-140), used at initialisation.
template <typename S>
__global__ void __launch_bounds__(256) k_p2c_all(BCArgsT<S> A) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= A.g.nx || j >= A.g.ny) return;
    p2c(A, gidx(A.g, i, j, k));
}
// cons->prim + p at every interior point (get_duvwT, src/commen_kernel.cu:111-118).
template <typename S>
__global__ void __launch_bounds__(256) k_c2p_all(StageArgsT<S> A) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= A.g.nx || j >= A.g.ny) return;
    const long long q = gidx(A.g, i, j, k), fs = A.fs;
    const double d1 = A.f[q];
    const double u1 = A.f[fs + q] / d1, v1 = A.f[2 * fs + q] / d1, w1 = A.f[3 * fs + q] / d1;
    const double T2 = (A.f[4 * fs + q] - 0.5 * d1 * (u1 * u1 + v1 * v1 + w1 * w1)) / (d1 * A.c.cv);
    A.d[q] = S(d1); A.u[q] = S(u1); A.v[q] = S(v1); A.w[q] = S(w1); A.T[q] = S(T2);
    A.P[q] = S(T2 * d1 / (A.c.gamma * A.c.ama * A.c.ama));
}
template <typename S>
__global__ void __launch_bounds__(256) k_amu_all(BCArgsT<S> A) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= A.g.nx || j >= A.g.ny) return;
    const long long q = gidx(A.g, i, j, k);
    const double T = double(A.T[q]);
    A.amu[q] = S(A.c.amu_c0 * sqrt(T * T * T) / (A.c.tsb + T));
}

// Running time average of d, u, v, w, T over the interior
// (ana_time_average_kernel, src/OCFD_ana.cu:240-254): m = (n m + x)/(n + 1)
// with n the number of samples already in m.  avg holds 5 FP64 fields.
template <typename S>
__global__ void __launch_bounds__(256) k_time_avg(Geo g, const S* __restrict__ d, const S* __restrict__ u,
                                                  const S* __restrict__ v, const S* __restrict__ w,
                                                  const S* __restrict__ T, double* __restrict__ avg, long long fs,
                                                  int n) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const long long q = gidx(g, i, j, k);
    const S* src[5] = {d, u, v, w, T};
    const double dn = double(n);
#pragma unroll
    for (int m = 0; m < 5; m++) {
        double* a = avg + m * fs + q;
        *a = (dn * *a + double(src[m][q])) / (dn + 1.0);
    }
}

__global__ void k_set(double* p, long long n, double v) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// ---------------------------------------------------------------------------
// Step diagnostics of NS_solver_real (src/OCFD_NS_Solver.cu:124-152):
// ana_residual sums of the total energy f[4] and of T over the interior
// (src/OCFD_ana.cu:50-104), and the NaN / negative-T scan of
// ana_NAN_and_NT over the LAP box (src/OCFD_ana.cu:105-222).  One partial
// per block, reduced in a fixed order (LDS tree here, the block index order
// on the host): the sums do not depend on scheduling.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_diag_sums(Geo g, const double* __restrict__ E, const S* __restrict__ T,
                                                   double* partE, double* partT) {
    __shared__ double se[256], st[256];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    double e = 0.0, t = 0.0;
    if (i < g.nx && j < g.ny) {
        const long long q = gidx(g, i, j, k);
        e = E[q];
        t = double(T[q]);
    }
    se[tid] = e;
    st[tid] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            se[tid] += se[tid + w];
            st[tid] += st[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        partE[b] = se[0];
        partT[b] = st[0];
    }
}

template <typename S>
__global__ void __launch_bounds__(256) k_diag_scan(const S* __restrict__ T, long long n, int* part) {
    __shared__ int sn[256], sg[256];
    const int tid = threadIdx.x;
    int nn = 0, ng = 0;
    for (long long q = (long long)blockIdx.x * 256 + tid; q < n; q += (long long)gridDim.x * 256) {
        const S v = T[q];
        nn += (v != v);
        ng += (v < S(0));
    }
    sn[tid] = nn;
    sg[tid] = ng;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sn[tid] += sn[tid + w];
            sg[tid] += sg[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[2 * blockIdx.x] = sn[0];
        part[2 * blockIdx.x + 1] = sg[0];
    }
}

// Final reduction of the diagnostics partials on the device (one block):
// thread t sums partials t, t+256, ... in order, then an LDS tree; out =
// {sum E, sum T, NaN count, negative count} (the counts as long long bits).
__global__ void __launch_bounds__(256) k_diag_final(const double* __restrict__ partE,
                                                    const double* __restrict__ partT, long long nb,
                                                    const int* __restrict__ scan_part, double* out) {
    __shared__ double se[256], st[256];
    __shared__ long long sn[256], sg[256];
    const int tid = threadIdx.x;
    double e = 0.0, t = 0.0;
    for (long long b = tid; b < nb; b += 256) {
        e += partE[b];
        t += partT[b];
    }
    long long nn = 0, ng = 0;
    for (int b = tid; b < DIAG_SCAN_BLOCKS; b += 256) {
        nn += scan_part[2 * b];
        ng += scan_part[2 * b + 1];
    }
    se[tid] = e;
    st[tid] = t;
    sn[tid] = nn;
    sg[tid] = ng;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            se[tid] += se[tid + w];
            st[tid] += st[tid + w];
            sn[tid] += sn[tid + w];
            sg[tid] += sg[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = se[0];
        out[1] = st[0];
        reinterpret_cast<long long*>(out)[2] = sn[0];
        reinterpret_cast<long long*>(out)[3] = sg[0];
    }
}

template <typename To, typename From>
__global__ void k_convert(To* dst, const From* src, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = To(src[t]);
}

dim3 pgrid(const Geo& g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }

}  // namespace

template <typename T>
void launch_fill(const Geo& g, T* base, int nf, long long fs, int dir, hipStream_t st) {
    const int na = dir == 0 ? g.ny : g.nx, nb = dir == 2 ? g.ny : g.nz;
    const long long tot = 8LL * na * nb;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, g, base, nf, fs, dir);
}

template <typename S>
void launch_stage(const StageArgsT<S>& a, hipStream_t st) {
    hipLaunchKernelGGL(k_stage<S>, pgrid(a.g), dim3(64, 4), 0, st, a);
}

template <typename S>
void launch_bc108(const BCArgsT<S>& a, hipStream_t st) {
    const Geo& g = a.g;
    const int nt = g.ny + (g.nx - a.x_begin > 0 ? g.nx - a.x_begin : 0);
    hipLaunchKernelGGL(k_bc_inlet_top<S>, dim3((nt + 63) / 64, g.nz), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_bc_wall<S>, dim3((g.nx + 63) / 64, g.nz), dim3(64), 0, st, a);
}

template <typename S>
void launch_bc108_faces(const BCArgsT<S>& a, hipStream_t st) {
    const Geo& g = a.g;
    const long long tot = 4LL * g.ny * g.nz + 4LL * g.nx * g.nz + 2LL * g.ny * g.nz + 2LL * g.nx * g.nz +
                          2LL * g.nx * g.ny;
    hipLaunchKernelGGL(k_bc108_faces<S>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a);
}

template <typename T>
void launch_mirror_y(const Geo& g, T* f, double sign, hipStream_t st) {
    hipLaunchKernelGGL(k_mirror_y<T>, dim3((g.nx + 63) / 64, 8, g.nz), dim3(64), 0, st, g, f, sign);
}

template <typename S>
void launch_bc124(const BCArgsT<S>& a, int sym, hipStream_t st) {
    const Geo& g = a.g;
    hipLaunchKernelGGL(k_bc124_inlet<S>, dim3((g.ny + 63) / 64, g.nz), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_bc124_upper<S>, dim3((g.nx + 63) / 64, g.ny), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_bc124_wall<S>, dim3((g.nx + 63) / 64, g.ny), dim3(64), 0, st, a);
    if (sym) {
        S* f[5] = {a.d, a.u, a.v, a.w, a.T};
        for (int m = 0; m < 5; m++) launch_mirror_y(g, f[m], m == 2 ? -1.0 : 1.0, st);
    }
}

template <typename S>
void launch_face_refresh(const BCArgsT<S>& a, hipStream_t st) {
    const Geo& g = a.g;
    const long long tot = 2LL * g.ny * g.nz + 2LL * g.nx * g.nz + 2LL * g.nx * g.ny;
    hipLaunchKernelGGL(k_faces<S>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a);
}

template <typename S>
void launch_sensor(const Geo& g, const Consts& c, const S* d, const S* T, int8_t* fx, int8_t* fy, int8_t* fz,
                   hipStream_t st) {
    // nodes i, j, k = -1 .. n-1
    dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, (g.nz + 1 + SENSOR_KZ - 1) / SENSOR_KZ);
    hipLaunchKernelGGL(k_sensor<S>, grid, dim3(64, 4), 0, st, g, c, d, T, fx, fy, fz);
}

template <typename S>
void launch_filter(const FilterArgsT<S>& a, hipStream_t st) {
    const int ex = a.hi[0] - a.lo[0], ey = a.hi[1] - a.lo[1], ez = a.hi[2] - a.lo[2];
    if (ex <= 0 || ey <= 0 || ez <= 0) return;
    hipLaunchKernelGGL(k_filter<S>, dim3((ex + 63) / 64, (ey + 3) / 4, ez), dim3(64, 4), 0, st, a);
}

void launch_metric_deriv(const Geo& g, const double* coord, int dir, int bound, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_metric_deriv, pgrid(g), dim3(64, 4), 0, st, g, coord, dir, bound, out);
}

void launch_cofactors(const Geo& g, const double* D9, long long fs, double* const* M, hipStream_t st) {
    MetricOut o;
    for (int m = 0; m < M_COUNT; m++) o.M[m] = M[m];
    hipLaunchKernelGGL(k_cofactors, pgrid(g), dim3(64, 4), 0, st, g, D9, fs, o);
}

template <typename S>
void launch_p2c_all(const BCArgsT<S>& a, hipStream_t st) {
    hipLaunchKernelGGL(k_p2c_all<S>, pgrid(a.g), dim3(64, 4), 0, st, a);
}

template <typename S>
void launch_c2p_all(const StageArgsT<S>& a, hipStream_t st) {
    hipLaunchKernelGGL(k_c2p_all<S>, pgrid(a.g), dim3(64, 4), 0, st, a);
}

template <typename S>
void launch_amu_all(const BCArgsT<S>& a, hipStream_t st) {
    hipLaunchKernelGGL(k_amu_all<S>, pgrid(a.g), dim3(64, 4), 0, st, a);
}

void launch_set(const Geo& g, double* base, long long n, double value, hipStream_t st) {
    (void)g;
    hipLaunchKernelGGL(k_set, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, base, n, value);
}

template <typename To, typename From>
void launch_convert(To* dst, const From* src, long long n, hipStream_t st) {
    hipLaunchKernelGGL((k_convert<To, From>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n);
}

template <typename S>
void launch_time_avg(const Geo& g, const S* d, const S* u, const S* v, const S* w, const S* T, double* avg,
                     long long fs, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_time_avg<S>, pgrid(g), dim3(64, 4), 0, st, g, d, u, v, w, T, avg, fs, n);
}

int diag_sum_blocks(const Geo& g) { return ((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz; }

template <typename S>
void launch_diag(const Geo& g, const double* E, const S* T_base, double* partE, double* partT, int* scan_part,
                 hipStream_t st) {
    hipLaunchKernelGGL(k_diag_sums<S>, pgrid(g), dim3(64, 4), 0, st, g, E, T_base, partE, partT);
    // the LAP box is the first sz*(nz+8) elements of the field
    hipLaunchKernelGGL(k_diag_scan<S>, dim3(DIAG_SCAN_BLOCKS), dim3(256), 0, st, T_base, g.sz * (g.nz + 8), scan_part);
}

void launch_diag_final(const double* partE, const double* partT, long long nb, const int* scan_part, double* out4,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, partE, partT, nb, scan_part, out4);
}

#define OCFD_INSTANTIATE(S)                                                                                     \
    template void launch_fill<S>(const Geo&, S*, int, long long, int, hipStream_t);                          \
    template void launch_stage<S>(const StageArgsT<S>&, hipStream_t);                                        \
    template void launch_bc108<S>(const BCArgsT<S>&, hipStream_t);                                           \
    template void launch_bc108_faces<S>(const BCArgsT<S>&, hipStream_t);                                     \
    template void launch_bc124<S>(const BCArgsT<S>&, int, hipStream_t);                                      \
    template void launch_mirror_y<S>(const Geo&, S*, double, hipStream_t);                                   \
    template void launch_face_refresh<S>(const BCArgsT<S>&, hipStream_t);                                    \
    template void launch_sensor<S>(const Geo&, const Consts&, const S*, const S*, int8_t*, int8_t*, int8_t*, \
                                   hipStream_t);                                                              \
    template void launch_filter<S>(const FilterArgsT<S>&, hipStream_t);                                      \
    template void launch_p2c_all<S>(const BCArgsT<S>&, hipStream_t);                                         \
    template void launch_c2p_all<S>(const StageArgsT<S>&, hipStream_t);                                      \
    template void launch_amu_all<S>(const BCArgsT<S>&, hipStream_t);                                        \
    template void launch_diag<S>(const Geo&, const double*, const S*, double*, double*, int*, hipStream_t);  \
    template void launch_time_avg<S>(const Geo&, const S*, const S*, const S*, const S*, const S*, double*, \
                                     long long, int, hipStream_t);
OCFD_INSTANTIATE(double)
OCFD_INSTANTIATE(float)
#undef OCFD_INSTANTIATE
template void launch_convert<float, double>(float*, const double*, long long, hipStream_t);
template void launch_convert<double, float>(double*, const float*, long long, hipStream_t);
