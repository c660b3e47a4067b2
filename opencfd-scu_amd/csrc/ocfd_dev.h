// ocfd_dev.h - device-side definitions shared by the gfx950 kernels.
//
// Layout: every field (primitives, pressure, metrics, conserved state, RHS,
// viscous fluxes) is one padded box of (nx+8) x (ny+8) x (nz+8) doubles with
// the reference's LAP=4 ghost layers (head/cuda_utility.h:16-27), x fastest.
// The row pitch ldx is a multiple of 16 doubles and the base pointer is offset
// so that interior x=0 of every row is 128-byte aligned.  Because all fields
// share one layout, a grid point has the same linear index in every array.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OCFD_LAPD 4

struct Geo {
    int nx, ny, nz;       // interior sizes
    int ldx, ldy;         // row pitch (elements), rows per plane (ny+8)
    long long sz;         // plane stride = ldx*ldy
    long long off0;       // linear index of interior (0,0,0)
    double hx, hy, hz;
};

// Linear index of interior point (i,j,k); ghosts are i<0 or i>=nx etc.
__host__ __device__ __forceinline__ long long gidx(const Geo& g, int i, int j, int k) {
    return g.off0 + (long long)i + (long long)g.ldx * ((long long)j + (long long)g.ldy * (long long)k);
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2), so consecutive block indices land on
// different L2s and neighbouring tiles - which share 128-byte lines and halo
// rows - would each fetch them from HBM.  xcd_block() maps the hardware block
// index to a logical tile index such that every XCD walks one contiguous run
// of tiles (x fastest): neighbours then share that XCD's L2.  It is a
// permutation of [0, nblocks) whatever the hardware mapping is.
struct Blk { int x, y, z; };
__device__ __forceinline__ Blk xcd_block() {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned nb = gx * gy * gridDim.z;
    const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
#ifdef OCFD_NO_XCD_REMAP
    const unsigned t = b;  // hardware order (for comparison builds)
    (void)nb;
#else
    const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, s = b >> 3;
    const unsigned t = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
#endif
    Blk o;
    o.x = (int)(t % gx);
    o.y = (int)((t / gx) % gy);
    o.z = (int)(t / (gx * gy));
    return o;
}

// The 10 metric fields kept on the device (cofactors and Ajac,
// src/OCFD_Comput_Jacobian3d.cu:285-295).
enum { M_KX = 0, M_KY, M_KZ, M_IX, M_IY, M_IZ, M_SX, M_SY, M_SZ, M_JAC, M_COUNT };

struct Consts {
    double gamma, ama, cv, cp, tsb, amu_c0, tmp0, split_c1, split_c3, vis_c;
    double epsl_sw, p00;                     // p00 = 1/(gamma*Ma^2)
    double rama;                             // 1/Ma
    double hy_i1, hy_i2;                     // HY_DP_INTV
    double tv_lim, tv_max;                   // WENO_TV_Limiter / WENO_TV_MAX
};

// --------------------------------------------------------------------------
// Half-node reconstructions (device restatement of src/OCFD_Schemes.cu and
// src/OCFD_bound_Scheme.cu).  s points at the first node the reference kernel
// gathers; expressions keep the reference operand order.
// --------------------------------------------------------------------------
template <typename R> __device__ __forceinline__ R d_sgn1(R x) { return x >= R(0) ? R(1) : R(-1); }
template <typename R> __device__ __forceinline__ R d_minmod2(R a, R b) {
    return R(0.5) * (d_sgn1(a) + d_sgn1(b)) * fmin(fabs(a), fabs(b));
}
template <typename R> __device__ __forceinline__ R d_minmod4(R a, R b, R c, R d) {
    R m = R(0.5) * (d_sgn1(a) + d_sgn1(b));
    m = m * fabs(R(0.5) * (d_sgn1(a) + d_sgn1(c)));
    m = m * fabs(R(0.5) * (d_sgn1(a) + d_sgn1(d)));
    return m * fmin(fmin(a, b), fmin(c, d));
}

// WENO7 upwind (P) half-node value from nodes s[0..6] (src/OCFD_Schemes.cu:616-663).
// R: type of the stencil values and smoothness indicators; W: type of the
// epsilon/weight arithmetic (the mixed-precision path keeps W = double).
template <typename R, typename W = R> __device__ __forceinline__ R weno7_up(const R* s) {
    R a, b, c, S0, S1, S2, S3;
    a = R(-2) * s[0] + R(9) * s[1] - R(18) * s[2] + R(11) * s[3];
    b = R(-6) * s[0] + R(24) * s[1] - R(30) * s[2] + R(12) * s[3];
    c = R(-6) * s[0] + R(18) * (s[1] - s[2]) + R(6) * s[3];
    S0 = R(960) * a * a; S0 += R(1040) * b * b; S0 += R(1043) * c * c; S0 += R(80) * (a * c);
    a = s[1] - R(6) * s[2] + R(3) * s[3] + R(2) * s[4];
    b = R(6) * s[2] - R(12) * s[3] + R(6) * s[4];
    c = R(-6) * s[1] + R(18) * (s[2] - s[3]) + R(6) * s[4];
    S1 = R(960) * a * a; S1 += R(1040) * b * b; S1 += R(1043) * c * c; S1 += R(80) * (a * c);
    a = R(-2) * s[2] - R(3) * s[3] + R(6) * s[4] - s[5];
    b = R(6) * s[3] - R(12) * s[4] + R(6) * s[5];
    c = R(-6) * s[2] + R(18) * (s[3] - s[4]) + R(6) * s[5];
    S2 = R(960) * a * a; S2 += R(1040) * b * b; S2 += R(1043) * c * c; S2 += R(80) * (a * c);
    a = R(-11) * s[3] + R(18) * s[4] - R(9) * s[5] + R(2) * s[6];
    b = R(12) * s[3] - R(30) * s[4] + R(24) * s[5] - R(6) * s[6];
    c = R(-6) * s[3] + R(18) * (s[4] - s[5]) + R(6) * s[6];
    S3 = R(960) * a * a; S3 += R(1040) * b * b; S3 += R(1043) * c * c; S3 += R(80) * (a * c);
    const W e = W(3.456e-4);
    W q0 = (e + W(S0)) * (e + W(S0)), q1 = (e + W(S1)) * (e + W(S1)), q2 = (e + W(S2)) * (e + W(S2)),
      q3 = (e + W(S3)) * (e + W(S3));
    W w0 = q1 * q2 * q3;
    W w1 = W(12) * q0 * q2 * q3;
    W w2 = W(18) * q0 * q1 * q3;
    W w3 = W(4) * q0 * q1 * q2;
    W am = w0 + w1 + w2 + w3;
    W r = w0 * W(R(-3) * s[0] + R(13) * s[1] - R(23) * s[2] + R(25) * s[3]);
    r += w1 * W(s[1] - R(5) * s[2] + R(13) * s[3] + R(3) * s[4]);
    r += w2 * W(-s[2] + R(7) * s[3] + R(7) * s[4] - s[5]);
    r += w3 * W(R(3) * s[3] + R(13) * s[4] - R(5) * s[5] + s[6]);
    return R(r / (W(12) * am));
}
// WENO7 M side: the same reconstruction on the mirrored stencil
// (src/OCFD_Schemes.cu:692-741); s[0..6].
template <typename R, typename W = R> __device__ __forceinline__ R weno7_dn(const R* s) {
    R m[7] = {s[6], s[5], s[4], s[3], s[2], s[1], s[0]};
    return weno7_up<R, W>(m);
}

// UP7 (src/OCFD_Schemes.cu:865-931).  The division by 420 is a multiplication
// by its reciprocal (an ulp-level difference; a division costs ten FP64
// instructions, the whole reconstruction eight).
template <typename R> __device__ __forceinline__ R up7_up(const R* s) {
    return (R(-3) * s[0] + R(25) * s[1] - R(101) * s[2] + R(319) * s[3] + R(214) * s[4] - R(38) * s[5] +
            R(4) * s[6]) * R(1.0 / 420.0);
}
template <typename R> __device__ __forceinline__ R up7_dn(const R* s) {  // s[1..7]
    return (R(-3) * s[7] + R(25) * s[6] - R(101) * s[5] + R(319) * s[4] + R(214) * s[3] - R(38) * s[2] +
            R(4) * s[1]) * R(1.0 / 420.0);
}

// WENO5 Jiang-Shu (src/OCFD_Schemes.cu:1118-1210) and the one-sided subsets
// used by the closures (src/OCFD_bound_Scheme.cu:557-735).
template <typename R> __device__ __forceinline__ R w5_beta_l(const R* s) {  // stencil s0..s2 (leftmost)
    R t = s[0] - R(2) * s[1] + s[2], u = s[0] - R(4) * s[1] + R(3) * s[2];
    R S = R(13) * t * t; S += R(3) * u * u; return S;
}
template <typename R> __device__ __forceinline__ R w5_beta_c(const R* s) {  // s1..s3
    R t = s[1] - R(2) * s[2] + s[3], u = s[1] - s[3];
    R S = R(13) * t * t; S += R(3) * u * u; return S;
}
template <typename R> __device__ __forceinline__ R w5_beta_r(const R* s) {  // s2..s4
    R t = s[2] - R(2) * s[3] + s[4], u = R(3) * s[2] - R(4) * s[3] + s[4];
    R S = R(13) * t * t; S += R(3) * u * u; return S;
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_up(const R* s) {
    const W e = W(12.0) * W(1e-6);  // 12.0*ep with ep = 1e-6, as the reference writes it
    const W S0 = w5_beta_r(s), S1 = w5_beta_c(s), S2 = w5_beta_l(s);
    R q03 = R(2) * s[2] + R(5) * s[3] - s[4];
    R q13 = -s[1] + R(5) * s[2] + R(2) * s[3];
    R q23 = R(2) * s[0] - R(7) * s[1] + R(11) * s[2];
    W a0 = W(3) * ((e + S1) * (e + S1)) * ((e + S2) * (e + S2));
    W a1 = W(6) * ((e + S0) * (e + S0)) * ((e + S2) * (e + S2));
    W a2 = ((e + S0) * (e + S0)) * ((e + S1) * (e + S1));
    return R((a0 * W(q03) + a1 * W(q13) + a2 * W(q23)) / (W(6) * (a0 + a1 + a2)));
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_dn(const R* s) {
    R m[5] = {s[4], s[3], s[2], s[1], s[0]};
    return weno5_up<R, W>(m);
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_up_left(const R* s) {  // OCFD_weno5_kernel_P_left
    const W e = W(12.0) * W(1e-6);  // 12.0*ep with ep = 1e-6, as the reference writes it
    const W S0 = w5_beta_r(s);
    W a0 = W(1) / ((e + S0) * (e + S0));
    return R(a0 * W(R(2) * s[2] + R(5) * s[3] - s[4]) / (W(6) * a0));
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_up_leftplus(const R* s) {
    const W e = W(12.0) * W(1e-6);  // 12.0*ep with ep = 1e-6, as the reference writes it
    const W S0 = w5_beta_r(s), S1 = w5_beta_c(s);
    W a0 = W(3) / ((e + S0) * (e + S0));
    W a1 = W(6) / ((e + S1) * (e + S1));
    R q03 = R(2) * s[2] + R(5) * s[3] - s[4];
    R q13 = -s[1] + R(5) * s[2] + R(2) * s[3];
    return R((a0 * W(q03) + a1 * W(q13)) / (W(6) * (a0 + a1)));
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_up_rightplus(const R* s) {
    const W e = W(12.0) * W(1e-6);  // 12.0*ep with ep = 1e-6, as the reference writes it
    const W S1 = w5_beta_c(s), S2 = w5_beta_l(s);
    W a1 = W(6) / ((e + S1) * (e + S1));
    W a2 = W(1) / ((e + S2) * (e + S2));
    R q13 = -s[1] + R(5) * s[2] + R(2) * s[3];
    R q23 = R(2) * s[0] - R(7) * s[1] + R(11) * s[2];
    return R((a1 * W(q13) + a2 * W(q23)) / (W(6) * (a1 + a2)));
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_dn_left(const R* s) {  // OCFD_weno5_kernel_M_left
    R m[5] = {s[4], s[3], s[2], s[1], s[0]};
    const W e = W(12.0) * W(1e-6);  // 12.0*ep with ep = 1e-6, as the reference writes it
    const W S2 = w5_beta_l(m);
    W a2 = W(1) / ((e + S2) * (e + S2));
    return R(a2 * W(R(2) * m[0] - R(7) * m[1] + R(11) * m[2]) / (W(6) * a2));
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_dn_leftplus(const R* s) {
    R m[5] = {s[4], s[3], s[2], s[1], s[0]};
    return weno5_up_rightplus<R, W>(m);
}
template <typename R, typename W = R> __device__ __forceinline__ R weno5_dn_rightplus(const R* s) {
    R m[5] = {s[4], s[3], s[2], s[1], s[0]};
    return weno5_up_leftplus<R, W>(m);
}

// WENO7-SYMBO (src/OCFD_Schemes.cu:391-589); P side with optional TV
// limiter; the M side never takes the linear branch (TVmin stays 0).
template <typename R, typename W = R> __device__ __forceinline__ R symbo_core(const R* s, bool linear) {
    W S0, S1, S2, S3, S4;
    if (linear) {
        S0 = W(0.0401954833730); S1 = W(0.2493800006710); S2 = W(0.4802686256260);
        S3 = W(0.2009775476730); S4 = W(0.0291783426580);
    } else {
        R t, B0, B1, B2, B3, B4;
        t = R(-2) * s[0] + R(9) * s[1] - R(18) * s[2] + R(11) * s[3]; B0 = R(720) * t * t;
        t = s[1] - R(6) * s[2] + R(3) * s[3] + R(2) * s[4]; B1 = R(720) * t * t;
        t = R(-2) * s[2] - R(3) * s[3] + R(6) * s[4] - s[5]; B2 = R(720) * t * t;
        t = R(-11) * s[3] + R(18) * s[4] - R(9) * s[5] + R(2) * s[6]; B3 = R(720) * t * t;
        t = R(-26) * s[4] + R(57) * s[5] - R(42) * s[6] + R(11) * s[7]; B4 = R(720) * t * t;
        t = R(-6) * s[0] + R(24) * s[1] - R(30) * s[2] + R(12) * s[3]; B0 += R(780) * t * t;
        t = R(6) * s[2] - R(12) * s[3] + R(6) * s[4]; B1 += R(780) * t * t;
        t = R(6) * s[3] - R(12) * s[4] + R(6) * s[5]; B2 += R(780) * t * t;
        t = R(12) * s[3] - R(30) * s[4] + R(24) * s[5] - R(6) * s[6]; B3 += R(780) * t * t;
        t = R(18) * s[4] - R(48) * s[5] + R(42) * s[6] - R(12) * s[7]; B4 += R(780) * t * t;
        t = R(-6) * s[0] + R(18) * (s[1] - s[2]) + R(6) * s[3]; B0 += R(781) * t * t;
        t = R(-6) * s[1] + R(18) * (s[2] - s[3]) + R(6) * s[4]; B1 += R(781) * t * t;
        t = R(-6) * s[2] + R(18) * (s[3] - s[4]) + R(6) * s[5]; B2 += R(781) * t * t;
        t = R(-6) * s[3] + R(18) * (s[4] - s[5]) + R(6) * s[6]; B3 += R(781) * t * t;
        t = R(-6) * s[4] + R(18) * (s[5] - s[6]) + R(6) * s[7]; B4 += R(781) * t * t;
        B4 = fmax(fmax(fmax(B0, B1), fmax(B2, B3)), B4);
        const W e = W(2.592e-6);
        W e0 = e + W(B0), e1 = e + W(B1), e2 = e + W(B2), e3 = e + W(B3), e4 = e + W(B4);
        S0 = W(0.0401954833730) * e1 * e2 * e3 * e4;
        S1 = W(0.2493800006710) * e0 * e2 * e3 * e4;
        S2 = W(0.4802686256260) * e0 * e1 * e3 * e4;
        S3 = W(0.2009775476730) * e0 * e1 * e2 * e4;
        S4 = W(0.0291783426580) * e0 * e1 * e2 * e3;
    }
    W am = S0 + S1 + S2 + S3 + S4;
    W r = S0 * W(R(-3) * s[0] + R(13) * s[1] - R(23) * s[2] + R(25) * s[3]);
    r += S1 * W(s[1] - R(5) * s[2] + R(13) * s[3] + R(3) * s[4]);
    r += S2 * W(-s[2] + R(7) * s[3] + R(7) * s[4] - s[5]);
    r += S3 * W(R(3) * s[3] + R(13) * s[4] - R(5) * s[5] + s[6]);
    r += S4 * W(R(25) * s[4] - R(23) * s[5] + R(13) * s[6] - R(3) * s[7]);
    return R(r / (W(12) * am));
}
template <typename R, typename W = R> __device__ __forceinline__ R symbo_up(const R* s, int lmt, R tvlim, R tvmax) {
    bool lin = false;
    if (lmt) {
        R a0 = fabs(s[1] - s[0]) + fabs(s[2] - s[1]) + fabs(s[3] - s[2]);
        R a1 = fabs(s[2] - s[1]) + fabs(s[3] - s[2]) + fabs(s[4] - s[3]);
        R a2 = fabs(s[3] - s[2]) + fabs(s[4] - s[3]) + fabs(s[5] - s[4]);
        R a3 = fabs(s[4] - s[3]) + fabs(s[5] - s[4]) + fabs(s[6] - s[5]);
        R a4 = fabs(s[5] - s[4]) + fabs(s[6] - s[5]) + fabs(s[7] - s[6]);
        R tvmin = fmin(fmin(fmin(a0, a1), fmin(a2, a3)), a4);
        R tvmx = fmax(fmax(fmax(a0, a1), fmax(a2, a3)), a4);
        lin = (tvmx < tvlim * tvmin) && (tvmx < tvmax);
    } else {
        lin = (R(1) < tvlim * R(0)) && (R(1) < tvmax);
    }
    return symbo_core<R, W>(s, lin);
}
template <typename R, typename W = R> __device__ __forceinline__ R symbo_dn(const R* s) {
    R m[8] = {s[7], s[6], s[5], s[4], s[3], s[2], s[1], s[0]};
    return symbo_core<R, W>(m, false);
}

// OMP6 (src/OCFD_Schemes.cu:936-1092).  The M limiter is kept as in the
// reference (it is not the mirror image of the P limiter).
template <typename R> __device__ __forceinline__ void omp6_mn(int fl, R& mid, R& m) {
    R mm, nn;
    if (fl == 1) { mm = R(0.001); nn = R(0); }
    else if (fl == 2) { mm = R(0); nn = R(-1.0 / 140.0); }
    else { mm = R(0.015); nn = R(0); }
    mid = R(0.5) * (mm + nn);
    m = R(0.5) * (mm - nn);
}
template <typename R> __device__ __forceinline__ R omp6_lin(const R* s, R mid, R m) {
    return (R(60) * mid * s[7] + (R(1) - R(60) * m - R(360) * mid) * s[6] +
            (R(-8) + R(360) * m + R(900) * mid) * s[5] + (R(37) - R(900) * m - R(1200) * mid) * s[4] +
            (R(37) + R(1200) * m + R(900) * mid) * s[3] + (R(-8) - R(900) * m - R(360) * mid) * s[2] +
            (R(1) + R(360) * m + R(60) * mid) * s[1] - R(60) * m * s[0]) / R(60);
}
template <typename R> __device__ __forceinline__ R omp6_up(int fl, const R* s) {
    R mid, m, n;
    omp6_mn(fl, mid, m);
    mid = omp6_lin(s, mid, m);
    m = s[3] + d_minmod2(s[4] - s[3], s[3] - s[2]);
    if ((mid - s[3]) * (mid - m) >= R(1e-10)) {
        m = s[2] + s[4] - R(2) * s[3];
        n = s[3] + s[5] - R(2) * s[4];
        R t = R(0.5) * (s[3] + s[4]) - R(0.5) * d_minmod4(R(4) * m - n, R(4) * n - m, n, m);
        R nb = s[1] + s[3] - R(2) * s[2];
        R t1 = s[3] + R(0.5) * (s[3] - s[2]) + R(4) * d_minmod4(R(4) * nb - m, R(4) * m - nb, m, nb) / R(3);
        R lo = fmax(fmin(fmin(s[3], s[4]), t), fmin(fmin(s[3], s[3] + R(4) * (s[3] - s[2])), t1));
        R hi = fmin(fmax(s[4], fmax(s[3], t)), fmax(fmax(s[3], s[3] + R(4) * (s[3] - s[2])), t1));
        mid = mid + d_minmod2(hi - mid, lo - mid);
    }
    return mid;
}
template <typename R> __device__ __forceinline__ R omp6_dn(int fl, const R* s) {
    R mid, m, n;
    omp6_mn(fl, mid, m);
    R rv[8] = {s[7], s[6], s[5], s[4], s[3], s[2], s[1], s[0]};
    mid = omp6_lin(rv, mid, m);
    m = s[4] + d_minmod2(s[3] - s[4], s[4] - s[5]);
    if ((mid - s[4]) * (mid - m) >= R(1e-10)) {
        m = s[5] + s[3] - R(2) * s[4];
        n = s[4] + s[2] - R(2) * s[3];
        R t = R(0.5) * (s[4] + s[3]) - R(0.5) * d_minmod4(R(4) * m - n, R(4) * n - m, n, m);
        R nb = s[6] + s[4] - R(2) * s[5];
        R t1 = s[4] + R(0.5) * (s[4] - s[3]) + R(4) * d_minmod4(R(4) * nb - m, R(4) * m - nb, m, nb) / R(3);
        R lo = fmax(fmin(fmin(s[5], s[3]), t), fmin(fmin(s[4], s[4] + R(4) * (s[4] - s[5])), t1));
        R hi = fmin(fmax(s[3], fmax(s[4], t)), fmax(fmax(s[4], s[4] + R(4) * (s[4] - s[5])), t1));
        mid = mid + d_minmod2(hi - mid, lo - mid);
    }
    return mid;
}

// NND2 (src/OCFD_Schemes.cu:800-837); s = &stencil[1].
template <typename R> __device__ __forceinline__ R nnd2_up(const R* s) {
    return s[1] + R(0.5) * d_minmod2(s[2] - s[1], s[1] - s[0]);
}
template <typename R> __device__ __forceinline__ R nnd2_dn(const R* s) {
    return s[1] + R(0.5) * d_minmod2(s[0] - s[1], s[1] - s[2]);
}

// Stencil extents (ia1, ib1) of each scheme's kernels.
template <int SCH, int SIDE> struct Ext {
    static constexpr int ka = (SCH == 303) ? -3 : (SCH == 302 || SCH == 306) ? -2 : (SIDE == 0 ? -3 : -4);
    static constexpr int kb = (SCH == 303) ? 3 : (SCH == 302 || SCH == 306) ? 2 : (SIDE == 0 ? 4 : 3);
};
