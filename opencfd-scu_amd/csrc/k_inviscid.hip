// k_inviscid.hip - fused inviscid sweep per direction (gfx950).
//
// One kernel per direction replaces the reference's split kernel
// (split_Jac3d_Stager_Warming_ker, src/OCFD_split.cu:145-198) and its ten
// derivative launches (OCFD_d?1 / OCFD_d?2, src/OCFD_Schemes_Choose.cu:88-765,
// each launching one P or M kernel per variable with an atomicAdd into du,
// src/OCFD_Schemes.cu:293-381):
//
// x (k_sweep, LDS tile):
//   phase 1  load rho,u,v,w,T and the direction's cofactors for a tile of
//            nodes (+4 ghost nodes each side along the sweep), compute the
//            sound speed and the Steger-Warming f+/f- of all 5 variables,
//            stage them in LDS (they never touch HBM);
//   phase 2  one thread per half node: P and M reconstructions of all 5
//            variables with the reference's boundary closures;
//   phase 3  one thread per node: du += -(hP(i+1/2)-hP(i-1/2))/h, then += the
//            M term, in the reference's order (x P, x M, y P, ...), with no
//            atomics.  The reference multiplies every term by the node's Ajac;
//            here the RK stage multiplies their sum once (k_stage), so the
//            sweeps do not read Ajac.
//
// y and z (k_tile): the same three phases on tiles of x-consecutive columns;
// k_march (the same steps per node while marching along the sweep, with the
// split fluxes in a register queue) is the alternative for the float path.
//
// S (double or float) is the type of the primitives, the split fluxes, the
// reconstructions and du; the WENO weight arithmetic and the Jacobian factor
// with the accumulation into du are always double.
#include <type_traits>

#include "mlp_dev.h"
#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

// Square root for the sound speed: correctly rounded in FP64; on the FP32
// path (whose sweeps are VALU-bound) the hardware square root, 1 ulp, inside
// the FP32 rounding the mixed path already carries (measured on the
// 1024x384x256 swept ramp: y and z sweeps 2.5 % faster).
template <typename R>
__device__ __forceinline__ R sqrt_s(R x) {
    if constexpr (std::is_same<R, float>::value) return __builtin_amdgcn_sqrtf(x);
    else return sqrt(x);
}

// Steger-Warming split of one node (src/OCFD_split.cu:92-142).  FP32: |A|
// and 1/|A| from the hardware reciprocal square root (1 ulp) instead of a
// correctly rounded square root and division.
template <typename R>
__device__ __forceinline__ void sw_split(R tmp0d, R u, R v, R w, R cc, R A1, R A2, R A3, const Consts& c,
                                         R* fp, R* fm) {
    const R s2 = A1 * A1 + A2 * A2 + A3 * A3;
    R ss, rs;
    if constexpr (std::is_same<R, float>::value) {
        rs = __builtin_amdgcn_rsqf(s2);
        ss = s2 * rs;
    } else {
        ss = sqrt(s2);
        rs = R(1.0) / ss;
    }
    R E1 = A1 * u + A2 * v + A3 * w;
    R E2 = E1 - cc * ss;
    R E3 = E1 + cc * ss;
    ss = rs;
    A1 *= ss; A2 *= ss; A3 *= ss;
    const R eps = R(c.epsl_sw);
    R E1P, E2P, E3P;
    if (eps == R(0)) {
        // sqrt(x*x) == |x| in binary floating point with round-to-nearest
        // (no underflow of x*x): the same values without three square roots
        E1P = (E1 + fabs(E1)) * R(0.50);
        E2P = (E2 + fabs(E2)) * R(0.50);
        E3P = (E3 + fabs(E3)) * R(0.50);
    } else {
        E1P = (E1 + sqrt(E1 * E1 + eps * eps)) * R(0.50);
        E2P = (E2 + sqrt(E2 * E2 + eps * eps)) * R(0.50);
        E3P = (E3 + sqrt(E3 * E3 + eps * eps)) * R(0.50);
    }
    R E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    R uc1 = u - cc * A1, uc2 = u + cc * A1;
    R vc1 = v - cc * A2, vc2 = v + cc * A2;
    R wc1 = w - cc * A3, wc2 = w + cc * A3;
    R vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * R(0.50);
    R vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * R(0.50);
    R vv = (R(c.gamma) - R(1.0)) * (u * u + v * v + w * w);
    R W2 = R(c.split_c3) * cc * cc;
    const R C1 = R(c.split_c1);
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

// P and M half-node values at face j+1/2 from gp/gm = f+/f- at nodes j-3..j+4.
// c = j+1, r = j-n (the reference's coords and coords+start-end).
template <typename R, typename W, int SCH>
__device__ __forceinline__ void face_pm(const R* gp, const R* gm, int c, int r, int bl, int br, int hy,
                                        R tvlim, R tvmax, R& hp, R& hm) {
    constexpr int kaP = Ext<SCH, 0>::ka, kbP = Ext<SCH, 0>::kb;
    constexpr int kaM = Ext<SCH, 1>::ka, kbM = Ext<SCH, 1>::kb;
    // ---- P (OCFD_bound_scheme_kernel_p, src/OCFD_bound_Scheme.cu:739-819)
    if (bl && c <= -kaP) {
        if (c == 0) hp = gp[4];
        else if (c == 1) hp = weno5_up_left<R, W>(gp + 1);
        else if (c == 2) hp = weno5_up_leftplus<R, W>(gp + 1);
        else hp = weno5_up<R, W>(gp + 1);
    } else if (br && r >= -kbP) {
        if (r == -1) hp = gp[3] + R(0.5) * d_minmod2(gp[3] - gp[2], gp[3] - gp[2]);
        else if (r == -2) hp = weno5_up_rightplus<R, W>(gp + 1);
        else hp = weno5_up<R, W>(gp + 1);
    } else {
        if (SCH == OCFD_WENO7) hp = weno7_up<R, W>(gp);
        else if (SCH == OCFD_UP7) hp = up7_up(gp);
        else if (SCH == OCFD_WENO5) hp = weno5_up<R, W>(gp + 1);
        else if (SCH == OCFD_NND2) hp = nnd2_up(gp + 2);
        else if (SCH == OCFD_WENO7_SYMBO) hp = symbo_up<R, W>(gp, 0, tvlim, tvmax);
        else if (SCH == OCFD_WENO7_SYMBO_LIM) hp = symbo_up<R, W>(gp, 1, tvlim, tvmax);
        else if (SCH == OCFD_OMP6_HR) hp = omp6_up(0, gp);
        else if (SCH == OCFD_OMP6_LD) hp = omp6_up(1, gp);
        else if (SCH == OCFD_OMP6_CD8) hp = omp6_up(2, gp);
        else {  // hybrid_auto, src/OCFD_Schemes_hybrid_auto.cu:911-919
            if (hy == 1) hp = up7_up(gp);
            else if (hy == 2) hp = weno7_up<R, W>(gp);
            else hp = weno5_up<R, W>(gp + 1);
        }
    }
    // ---- M (OCFD_bound_scheme_kernel_m, src/OCFD_bound_Scheme.cu:855-944)
    if (bl && c < -kaM) {
        if (c == 0) hm = weno5_dn_left<R, W>(gm + 2);
        else if (c == 1) hm = weno5_dn_leftplus<R, W>(gm + 2);
        else hm = weno5_dn<R, W>(gm + 2);
    } else if (br && r >= -kbM - 1) {
        if (r == -1) hm = gm[3];
        else if (r == -2) hm = gm[4] - R(0.5) * d_minmod2(gm[4] - gm[3], gm[4] - gm[3]);
        else if (r == -3) hm = weno5_dn_rightplus<R, W>(gm + 2);
        else hm = weno5_dn<R, W>(gm + 2);
    } else {
        if (SCH == OCFD_WENO7) hm = weno7_dn<R, W>(gm + 1);
        else if (SCH == OCFD_UP7) hm = up7_dn(gm);
        else if (SCH == OCFD_WENO5) hm = weno5_dn<R, W>(gm + 2);
        else if (SCH == OCFD_NND2) hm = nnd2_dn(gm + 3);
        else if (SCH == OCFD_WENO7_SYMBO || SCH == OCFD_WENO7_SYMBO_LIM) hm = symbo_dn<R, W>(gm);
        else if (SCH == OCFD_OMP6_HR) hm = omp6_dn(0, gm);
        else if (SCH == OCFD_OMP6_LD) hm = omp6_dn(1, gm);
        else if (SCH == OCFD_OMP6_CD8) hm = omp6_dn(2, gm);
        else {  // src/OCFD_Schemes_hybrid_auto.cu:1008-1016
            if (hy == 1) hm = up7_dn(gm);
            else if (hy == 2) hm = weno7_dn<R, W>(gm + 1);
            else hm = weno5_dn<R, W>(gm + 2);
        }
    }
}

// The same selection for all five variables of a face: the closure/scheme
// case of the P and the M side is decided once per face, then each case loops
// over the variables (ld(v, k) = f+ of variable v (v < 5) or f- of variable
// v-5 at stencil node k, k = 0..7 for nodes j-3..j+4).  Same arithmetic as
// face_pm, with the per-variable branching hoisted.
template <typename R, typename W, int SCH, typename LD>
__device__ __forceinline__ void face_all_p(LD&& ld, int c, int r, int bl, int br, int hy, R tvlim, R tvmax, R* hp) {
    constexpr int kaP = Ext<SCH, 0>::ka, kbP = Ext<SCH, 0>::kb;
    // P: 0..3 left closures (c), 4..6 right closures, 7 scheme, 8 WENO7, 9 WENO5
    int pc;
    if (bl && c <= -kaP) pc = c;
    else if (br && r >= -kbP) pc = r == -1 ? 4 : r == -2 ? 5 : 6;
    else if (SCH == OCFD_HYBRIDAUTO) pc = hy == 1 ? 7 : hy == 2 ? 8 : 9;
    else pc = 7;
    if (pc == 7) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 0; k < 8; k++) g[k] = ld(m, k);
            if (SCH == OCFD_WENO7) hp[m] = weno7_up<R, W>(g);
            else if (SCH == OCFD_UP7 || SCH == OCFD_HYBRIDAUTO) hp[m] = up7_up(g);
            else if (SCH == OCFD_WENO5) hp[m] = weno5_up<R, W>(g + 1);
            else if (SCH == OCFD_NND2) hp[m] = nnd2_up(g + 2);
            else if (SCH == OCFD_WENO7_SYMBO) hp[m] = symbo_up<R, W>(g, 0, tvlim, tvmax);
            else if (SCH == OCFD_WENO7_SYMBO_LIM) hp[m] = symbo_up<R, W>(g, 1, tvlim, tvmax);
            else if (SCH == OCFD_OMP6_HR) hp[m] = omp6_up(0, g);
            else if (SCH == OCFD_OMP6_LD) hp[m] = omp6_up(1, g);
            else hp[m] = omp6_up(2, g);
        }
    } else if (pc == 8) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 0; k < 7; k++) g[k] = ld(m, k);
            hp[m] = weno7_up<R, W>(g);
        }
    } else if (pc == 9 || pc == 3 || pc == 6) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 1; k < 6; k++) g[k] = ld(m, k);
            hp[m] = weno5_up<R, W>(g + 1);
        }
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 0; k < 8; k++) g[k] = ld(m, k);
            if (pc == 0) hp[m] = g[4];
            else if (pc == 1) hp[m] = weno5_up_left<R, W>(g + 1);
            else if (pc == 2) hp[m] = weno5_up_leftplus<R, W>(g + 1);
            else if (pc == 4) hp[m] = g[3] + R(0.5) * d_minmod2(g[3] - g[2], g[3] - g[2]);
            else hp[m] = weno5_up_rightplus<R, W>(g + 1);
        }
    }
}
// M half-node values (ld(5+m, k) = f- of variable m at node j-3+k).
template <typename R, typename W, int SCH, typename LD>
__device__ __forceinline__ void face_all_m(LD&& ld, int c, int r, int bl, int br, int hy, R* hm) {
    constexpr int kaM = Ext<SCH, 1>::ka, kbM = Ext<SCH, 1>::kb;
    // M: 0..2 left closures (c), 3..6 right closures, 7 scheme, 8 WENO7, 9 WENO5
    int mc;
    if (bl && c < -kaM) mc = c < 2 ? c : 2;  // c >= 2: plain WENO5
    else if (br && r >= -kbM - 1) mc = r == -1 ? 3 : r == -2 ? 4 : r == -3 ? 5 : 6;
    else if (SCH == OCFD_HYBRIDAUTO) mc = hy == 1 ? 7 : hy == 2 ? 8 : 9;
    else mc = 7;
    if (mc == 7) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 0; k < 8; k++) g[k] = ld(5 + m, k);
            if (SCH == OCFD_WENO7) hm[m] = weno7_dn<R, W>(g + 1);
            else if (SCH == OCFD_UP7 || SCH == OCFD_HYBRIDAUTO) hm[m] = up7_dn(g);
            else if (SCH == OCFD_WENO5) hm[m] = weno5_dn<R, W>(g + 2);
            else if (SCH == OCFD_NND2) hm[m] = nnd2_dn(g + 3);
            else if (SCH == OCFD_WENO7_SYMBO || SCH == OCFD_WENO7_SYMBO_LIM) hm[m] = symbo_dn<R, W>(g);
            else if (SCH == OCFD_OMP6_HR) hm[m] = omp6_dn(0, g);
            else if (SCH == OCFD_OMP6_LD) hm[m] = omp6_dn(1, g);
            else hm[m] = omp6_dn(2, g);
        }
    } else if (mc == 8) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 1; k < 8; k++) g[k] = ld(5 + m, k);
            hm[m] = weno7_dn<R, W>(g + 1);
        }
    } else if (mc == 9 || mc == 2 || mc == 6) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 2; k < 7; k++) g[k] = ld(5 + m, k);
            hm[m] = weno5_dn<R, W>(g + 2);
        }
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            R g[8];
#pragma unroll
            for (int k = 0; k < 8; k++) g[k] = ld(5 + m, k);
            if (mc == 0) hm[m] = weno5_dn_left<R, W>(g + 2);
            else if (mc == 1) hm[m] = weno5_dn_leftplus<R, W>(g + 2);
            else if (mc == 3) hm[m] = g[3];
            else if (mc == 4) hm[m] = g[4] - R(0.5) * d_minmod2(g[4] - g[3], g[4] - g[3]);
            else hm[m] = weno5_dn_rightplus<R, W>(g + 2);
        }
    }
}
template <typename R, typename W, int SCH, typename LD>
__device__ __forceinline__ void face_all(LD&& ld, int c, int r, int bl, int br, int hy, R tvlim, R tvmax, R* hp,
                                         R* hm) {
    face_all_p<R, W, SCH>(ld, c, r, bl, br, hy, tvlim, tvmax, hp);
    face_all_m<R, W, SCH>(ld, c, r, bl, br, hy, hm);
}

// The closure/scheme case of a face's P value (0..3 left closures by c, 4..6
// right closures, 7 the scheme, 8 WENO7, 9 WENO5) and M value (0..2 left
// closures, 3..6 right closures, 7 the scheme, 8 WENO7, 9 WENO5): the
// selection of face_all_p / face_all_m.
template <int SCH>
__device__ __forceinline__ int case_p(int c, int r, int bl, int br, int hy) {
    constexpr int kaP = Ext<SCH, 0>::ka, kbP = Ext<SCH, 0>::kb;
    if (bl && c <= -kaP) return c;
    if (br && r >= -kbP) return r == -1 ? 4 : r == -2 ? 5 : 6;
    if (SCH == OCFD_HYBRIDAUTO) return hy == 1 ? 7 : hy == 2 ? 8 : 9;
    return 7;
}
template <int SCH>
__device__ __forceinline__ int case_m(int c, int r, int bl, int br, int hy) {
    constexpr int kaM = Ext<SCH, 1>::ka, kbM = Ext<SCH, 1>::kb;
    if (bl && c < -kaM) return c < 2 ? c : 2;
    if (br && r >= -kbM - 1) return r == -1 ? 3 : r == -2 ? 4 : r == -3 ? 5 : 6;
    if (SCH == OCFD_HYBRIDAUTO) return hy == 1 ? 7 : hy == 2 ? 8 : 9;
    return 7;
}
// One P value (case pc) or M value (case mc) of hybrid_auto from the stencil
// g[0..7] of nodes j-3..j+4: the formulas of face_all_p / face_all_m case by
// case.
template <typename R, typename W>
__device__ __forceinline__ R recon_p_hybrid(int pc, const R* g) {
    switch (pc) {
        case 0: return g[4];
        case 1: return weno5_up_left<R, W>(g + 1);
        case 2: return weno5_up_leftplus<R, W>(g + 1);
        case 4: return g[3] + R(0.5) * d_minmod2(g[3] - g[2], g[3] - g[2]);
        case 5: return weno5_up_rightplus<R, W>(g + 1);
        case 7: return up7_up(g);
        case 8: return weno7_up<R, W>(g);
        default: return weno5_up<R, W>(g + 1);  // 3, 6, 9
    }
}
template <typename R, typename W>
__device__ __forceinline__ R recon_m_hybrid(int mc, const R* g) {
    switch (mc) {
        case 0: return weno5_dn_left<R, W>(g + 2);
        case 1: return weno5_dn_leftplus<R, W>(g + 2);
        case 3: return g[3];
        case 4: return g[4] - R(0.5) * d_minmod2(g[4] - g[3], g[4] - g[3]);
        case 5: return weno5_dn_rightplus<R, W>(g + 2);
        case 7: return up7_dn(g);
        case 8: return weno7_dn<R, W>(g + 1);
        default: return weno5_dn<R, W>(g + 2);  // 2, 6, 9
    }
}

// hybrid_auto sends most faces to UP7 and a few (at shocks, and closures near
// DIF_BOUND faces) to WENO7 / WENO5.  A wave whose lanes each ran their own
// face's ten reconstructions would run the WENO code ten times over for a
// single WENO face while the other lanes wait.  Instead the UP7 faces are
// computed by their own lanes and the ten P and M values of the other
// ("special") faces are spread over the wave, one value per lane: up to six
// special faces per wave in one pass; a wave with more takes the per-face
// path.  Same formulas, so the same values.  Used in the x sweep, which is
// VALU-bound (91 % VALU busy on the 1024x384x256 swept ramp; 2.40 -> 2.23 ms
// per launch); the y and z tiles, at 48-60 % VALU busy, measured the same or
// 2 % slower with it and keep the per-face path.
#ifndef OCFD_SWEEP_COMPACT
#define OCFD_SWEEP_COMPACT 1
#endif
constexpr int COMPACT_MAX = 6;  // special faces per wave spread in one pass (10 values each)
// The lane of the n-th special face of the wave (n-th set bit of mask).
__device__ __forceinline__ int nth_lane(unsigned long long mask, int n) {
    for (int t = 0; t < n; t++) mask &= mask - 1;
    return mask ? __builtin_ctzll(mask) : 0;
}

// x sweep: a block owns XA rows of 56 output nodes (one wave per row); the
// f+/f- of the 64 nodes of a row (4 ghost nodes each side) are staged in LDS.
#ifndef OCFD_XA
// rows per x-sweep block (one wave each): measured 146 us with 1 row, 156 us
// with 2 and 158 us with 4 (FP64, 512x128x64 ramp, kernel alone)
#define OCFD_XA 1
#endif
constexpr int XS = 56, XA = OCFD_XA, XN = XS + 8, XF = XS + 1;

// LEARN (first RK stage with the learned sensor, hybrid_auto): before phase
// 1 the wave of a row runs the learned sensor on its 64 nodes s0-4..s0+59
// (features, two BF16-MFMA layers, the flag rule; mlp_dev.h), writes the
// flags and probabilities of the nodes its faces use (s0-1..s0+55) and takes
// each face's flag from the lane of its node instead of loading it.
template <int SCH, bool FIRST, typename S, bool LEARN = false>
__global__ void __launch_bounds__(64 * XA) k_sweep(SweepArgsT<S> A) {
    static_assert(XF <= 64 && XN == 64, "one row per wave; one node and one face per thread");
    __shared__ S sh[10 * XN * XA];
    const Geo& g = A.g;
    const int tid = threadIdx.x;
    const Blk B = xcd_block();
    const int s0 = B.x * XS, a0 = B.y * XA, k2 = B.z;
    const int n_dir = g.nx;
    // LDS element (v, n, a): node-contiguous so consecutive lanes hold
    // consecutive x
    auto at = [](int v, int n, int a) { return (v * XA + a) * XN + n; };

    // phases 2 and 3: wave a handles row a, lane = position along the row
    // (57 faces, 56 nodes), so that the LDS accesses of a wave are
    // consecutive doubles of one row (no bank conflicts).  The hybrid flag of
    // this thread's face is fetched first, so that its latency hides behind
    // phase 1
    const int lane = tid & 63, wrow = tid >> 6;
    const bool fthread = lane < XF;
    const int f_loc = fthread ? lane : 0, a_loc = wrow;
    const int jf = s0 - 1 + f_loc, acf = a0 + a_loc;
    const bool fvalid = fthread && (jf <= n_dir - 1) && (acf < g.ny);
    int hy = 0;
    if constexpr (LEARN) {
        static_assert(SCH == OCFD_HYBRIDAUTO, "the learned sensor selects hybrid_auto's schemes");
        __shared__ unsigned short sXm[XA][64 * mlp::LS];
        mlp::FeatIn<S> F;
        F.g = g;
        F.d = A.d; F.u = A.u; F.v = A.v; F.w = A.w; F.T = A.T;
        for (int m = 0; m < M_COUNT; m++) F.M[m] = A.mlp.M[m];
        F.p00 = A.c.p00;
        F.rama = A.c.rama;
        mlp::Lane L;
        mlp::lane_weights_packed(A.mlp.P, lane, L);
        const int sn = s0 - 4 + lane;  // this lane's node
        const int ic = sn < -1 ? -1 : (sn > n_dir - 1 ? n_dir - 1 : sn);
        const int jr = acf < g.ny ? acf : g.ny - 1;
        float f[mlp::NF];
        mlp::features(F, ic, jr, k2, f);
        const float prob = mlp::wave_prob(f, sXm[wrow], L, lane);
        const long long qn = gidx(g, ic, jr, k2);
        const int fl = mlp::flag_rule(prob, mlp::jameson_kp(F, qn, A.c.hy_i1, A.c.hy_i2), A.mlp.lo, A.mlp.hi);
        if (lane >= 3 && lane <= 3 + XS && sn == ic && acf < g.ny) {
            A.mlp.fx[qn] = (int8_t)fl;
            if (sn >= 0) {
                A.mlp.fy[qn] = (int8_t)fl;
                A.mlp.fz[qn] = (int8_t)fl;
            }
            if (A.mlp.prob) A.mlp.prob[qn] = prob;
        }
        hy = __shfl(fl, f_loc + 3);  // face jf+1/2 has the flag of node jf = s0-4 + (f_loc+3)
    } else {
        if (SCH == OCFD_HYBRIDAUTO && fvalid) hy = A.flag[gidx(g, jf, acf, k2)];
    }
    const bool nthread = lane < XS;
    const int n3 = nthread ? lane : 0, a3 = wrow;
    const int i3 = s0 + n3, ac3 = a0 + a3;
    const bool nvalid = nthread && i3 < n_dir && ac3 < g.ny;
    const long long q3 = nvalid ? gidx(g, i3, ac3, k2) : 0;

    // ---- phase 1: split of XN x XA nodes into LDS
    {
        const int n = tid % XN, a = tid / XN;
        const int s = s0 - 4 + n, ac = a0 + a;
        S fp[5], fm[5];
        if (s <= n_dir + 3 && ac < g.ny) {
            const long long q = gidx(g, s, ac, k2);
            const S d = A.d[q], u = A.u[q], v = A.v[q], w = A.w[q], T = A.T[q];
            const S cc = sqrt_s(T) * S(A.c.rama);  // sound_speed_kernel, src/commen_kernel.cu:144-153
            sw_split<S>(d * S(A.c.tmp0), u, v, w, cc, S(A.m1[q]), S(A.m2[q]), S(A.m3[q]), A.c, fp, fm);
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) { fp[m] = S(0); fm[m] = S(0); }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[at(m, n, a)] = fp[m];
            sh[at(5 + m, n, a)] = fm[m];
        }
    }
    __syncthreads();

    // ---- phase 2: one half node per thread (face jf+1/2)
    S hp[5], hm[5];
    if (SCH == OCFD_HYBRIDAUTO && OCFD_SWEEP_COMPACT) {
        int pc = 7, mc = 7;
        if (fvalid) {
            const int c = jf + 1, r = jf - n_dir;
            pc = case_p<SCH>(c, r, A.bl, A.br, hy);
            mc = case_m<SCH>(c, r, A.bl, A.br, hy);
        }
        const bool special = fvalid && (pc != 7 || mc != 7);
        if (fvalid && !special) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                S gp[8], gm[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    gp[k] = sh[at(m, f_loc + k, a_loc)];
                    gm[k] = sh[at(5 + m, f_loc + k, a_loc)];
                }
                hp[m] = up7_up(gp);
                hm[m] = up7_dn(gm);
            }
        }
        const unsigned long long smask = __ballot(special);
        const int nspec = __popcll(smask);
        const bool spread = nspec <= COMPACT_MAX;
        // job of this lane: value v (P of variable v < 5, M of variable v-5)
        // of the special face held by lane fl
        const bool job = spread && lane < 10 * nspec;
        const int fl = nth_lane(smask, job ? lane / 10 : 0), v = lane % 10;
        const int pcj = __shfl(pc, fl), mcj = __shfl(mc, fl);
        S res = S(0);
        if (job) {
            S gg[8];
#pragma unroll
            for (int k = 0; k < 8; k++) gg[k] = sh[at(v, fl + k, a_loc)];
            res = v < 5 ? recon_p_hybrid<S, double>(pcj, gg) : recon_m_hybrid<S, double>(mcj, gg);
        }
        if (special && !spread) {
            const int c = jf + 1, r = jf - n_dir;
            face_all<S, double, SCH>([&](int vv, int k) { return sh[at(vv, f_loc + k, a_loc)]; }, c, r, A.bl, A.br,
                                     hy, S(A.c.tv_lim), S(A.c.tv_max), hp, hm);
        }
        __syncthreads();
        if (fthread && !(special && spread)) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                sh[(m * XA + a_loc) * XF + f_loc] = fvalid ? hp[m] : S(0);
                sh[((5 + m) * XA + a_loc) * XF + f_loc] = fvalid ? hm[m] : S(0);
            }
        }
        if (job) sh[(v * XA + a_loc) * XF + fl] = res;
    } else {
        if (fvalid) {
            const int c = jf + 1, r = jf - n_dir;
            face_all<S, double, SCH>([&](int v, int k) { return sh[at(v, f_loc + k, a_loc)]; }, c, r, A.bl, A.br,
                                     hy, S(A.c.tv_lim), S(A.c.tv_max), hp, hm);
        }
        __syncthreads();
        if (fthread) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                sh[(m * XA + a_loc) * XF + f_loc] = fvalid ? hp[m] : S(0);
                sh[((5 + m) * XA + a_loc) * XF + f_loc] = fvalid ? hm[m] : S(0);
            }
        }
    }
    __syncthreads();

    // ---- phase 3: du at nodes, in the reference's order P then M; the
    // accumulation is in double
    if (nvalid) {
        const int n = n3, a = a3, i = i3;
        {
            const long long q = q3;
            const bool zp = A.nrp && i == 0;
            const bool zm = A.nrm && i == n_dir - 1;
            const double rh = 1.0 / A.h;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const S pr = sh[(m * XA + a) * XF + n + 1], pl = sh[(m * XA + a) * XF + n];
                const S mr = sh[((5 + m) * XA + a) * XF + n + 1], ml = sh[((5 + m) * XA + a) * XF + n];
                const double tP = zp ? 0.0 : -double(pr - pl) * rh;
                const double tM = zm ? 0.0 : -double(mr - ml) * rh;
                S* dv = A.du + (long long)m * A.fs;
                double acc = FIRST ? 0.0 : double(dv[q]);
                acc = acc + tP;
                acc = acc + tM;
                // streamed: the next sweep reads it back, not this one
                // (measured: the y sweep 182 -> 176 us)
                __builtin_nontemporal_store(S(acc), &dv[q]);
            }
        }
    }
}

// y and z sweeps with LDS tiles: a block owns NA x-consecutive columns (one
// NA*8-byte row segment per node) and NS output nodes along the sweep, for one
// value of the third index.  Same three phases as the x sweep; a thread may
// handle up to two faces and two nodes.
#ifdef OCFD_TILE_WPE
#define TILE_ATTR __attribute__((amdgpu_waves_per_eu(OCFD_TILE_WPE, 8)))
#else
#define TILE_ATTR
#endif
template <int DIR, int SCH, int NS, int NA, typename S>
__global__ void __launch_bounds__(256) TILE_ATTR k_tile(SweepArgsT<S> A) {
    static_assert(DIR == 1 || DIR == 2, "x is swept by k_sweep");
    constexpr int NN = NS + 8, NF = NS + 1;
    constexpr int FPT = (NF * NA + 255) / 256;
    static_assert(FPT <= 2 && NN * NA >= NF * NA, "at most two faces per thread");
    __shared__ S sh[10 * NN * NA];
    const Geo& g = A.g;
    const int tid = threadIdx.x;
    const Blk B = xcd_block();
    const int a0 = B.x * NA, s0 = B.y * NS, k2 = B.z;
    const int n_dir = DIR == 1 ? g.ny : g.nz;
    auto lin = [&](int s, int a) -> long long { return DIR == 1 ? gidx(g, a, s, k2) : gidx(g, a, k2, s); };
    auto at = [](int v, int n, int a) { return (v * NN + n) * NA + a; };

    // the du of this thread's output nodes is fetched first, so that its
    // latency hides behind phases 1 and 2
    constexpr int NPT = (NS * NA + 255) / 256;
    double pre_du[NPT][5];
#pragma unroll
    for (int t = 0; t < NPT; t++) {
        const int idx = tid + 256 * t;
        const int a = idx % NA, n = idx / NA;
        const int i = s0 + n, ac = a0 + a;
        if (idx < NS * NA && i < n_dir && ac < g.nx) {
            const long long q = lin(i, ac);
#pragma unroll
            for (int m = 0; m < 5; m++) pre_du[t][m] = double(A.du[(long long)m * A.fs + q]);
        }
    }

    // hybrid flags of this thread's faces, fetched before phase 1 as well
    int pre_hy[FPT];
#pragma unroll
    for (int t = 0; t < FPT; t++) {
        pre_hy[t] = 0;
        const int idx = tid + 256 * t;
        const int j = s0 - 1 + idx / NA, ac = a0 + idx % NA;
        if (SCH == OCFD_HYBRIDAUTO && idx < NF * NA && j <= n_dir - 1 && ac < g.nx) pre_hy[t] = A.flag[lin(j, ac)];
    }

    // ---- phase 1: split of NN x NA nodes into LDS
    for (int idx = tid; idx < NN * NA; idx += 256) {
        const int a = idx % NA, n = idx / NA;
        const int s = s0 - 4 + n, ac = a0 + a;
        S fp[5], fm[5];
        if (s <= n_dir + 3 && ac < g.nx) {
            const long long q = lin(s, ac);
            const S d = A.d[q], u = A.u[q], v = A.v[q], w = A.w[q], T = A.T[q];
            const S cc = sqrt_s(T) * S(A.c.rama);  // sound_speed_kernel, src/commen_kernel.cu:144-153
            sw_split<S>(d * S(A.c.tmp0), u, v, w, cc, S(A.m1[q]), S(A.m2[q]), S(A.m3[q]), A.c, fp, fm);
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) { fp[m] = S(0); fm[m] = S(0); }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[at(m, n, a)] = fp[m];
            sh[at(5 + m, n, a)] = fm[m];
        }
    }
    __syncthreads();

    // ---- phase 2: half nodes
    S hp[FPT][5], hm[FPT][5];
#pragma unroll
    for (int t = 0; t < FPT; t++) {
        const int idx = tid + 256 * t;
        if (idx >= NF * NA) break;
        const int a_loc = idx % NA, f_loc = idx / NA;
        const int j = s0 - 1 + f_loc;  // face j+1/2
        const int ac = a0 + a_loc;
        if (j <= n_dir - 1 && ac < g.nx) {
            const int hy = pre_hy[t];
            const int c = j + 1, r = j - n_dir;
            face_all<S, double, SCH>([&](int v, int k) { return sh[at(v, f_loc + k, a_loc)]; }, c, r, A.bl, A.br, hy,
                                     S(A.c.tv_lim), S(A.c.tv_max), hp[t], hm[t]);
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) { hp[t][m] = S(0); hm[t][m] = S(0); }
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < FPT; t++) {
        const int idx = tid + 256 * t;
        if (idx >= NF * NA) break;
        const int a_loc = idx % NA, f_loc = idx / NA;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[(m * NF + f_loc) * NA + a_loc] = hp[t][m];
            sh[((5 + m) * NF + f_loc) * NA + a_loc] = hm[t][m];
        }
    }
    __syncthreads();

    // ---- phase 3: du at nodes (P then M; the sum in double)
    const double rh = 1.0 / A.h;
#pragma unroll
    for (int t = 0; t < NPT; t++) {
        const int idx = tid + 256 * t;
        const int a = idx % NA, n = idx / NA;
        const int i = s0 + n, ac = a0 + a;
        if (idx < NS * NA && i < n_dir && ac < g.nx) {
            const long long q = lin(i, ac);
            const bool zp = A.nrp && i == 0;
            const bool zm = A.nrm && i == n_dir - 1;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const S pr = sh[(m * NF + n + 1) * NA + a], pl = sh[(m * NF + n) * NA + a];
                const S mr = sh[((5 + m) * NF + n + 1) * NA + a], ml = sh[((5 + m) * NF + n) * NA + a];
                const double tP = zp ? 0.0 : -double(pr - pl) * rh;
                const double tM = zm ? 0.0 : -double(mr - ml) * rh;
                double acc = pre_du[t][m];
                acc = acc + tP;
                acc = acc + tM;
                A.du[(long long)m * A.fs + q] = S(acc);
            }
        }
    }
}

// y and z sweeps: one thread per (x, across) column marching along the sweep
// direction over a chunk of L output nodes.  The f+/f- of the 8 nodes a face
// needs (j-3..j+4) live in a register queue: each node is split once per chunk
// (plus the 8 warm-up nodes), rows of 64 x-consecutive nodes are read by one
// wave (512 contiguous bytes in FP64), and there is no LDS and no barrier.
// The next node's inputs are fetched one iteration ahead.
template <int DIR, int SCH, int L, typename S>
__global__ void __launch_bounds__(256) k_march(SweepArgsT<S> A) {
    static_assert(DIR == 1 || DIR == 2, "x is swept by k_sweep");
    const Geo& g = A.g;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int a = blockIdx.y * 4 + threadIdx.y;  // z for the y sweep, y for the z sweep
    const int n_dir = DIR == 1 ? g.ny : g.nz;
    const int n_acr = DIR == 1 ? g.nz : g.ny;
    const int s0 = blockIdx.z * L;
    const int s1 = min(s0 + L, n_dir);
    const bool own = i < g.nx && a < n_acr;
    const int ic = min(i, g.nx - 1), acl = min(a, n_acr - 1);
    const long long st = DIR == 1 ? (long long)g.ldx : g.sz;
    const long long base = DIR == 1 ? gidx(g, ic, 0, acl) : gidx(g, ic, acl, 0);
    const S* __restrict__ pd = A.d;
    const S* __restrict__ pu = A.u;
    const S* __restrict__ pv = A.v;
    const S* __restrict__ pw = A.w;
    const S* __restrict__ pT = A.T;
    const double* __restrict__ p1 = A.m1;
    const double* __restrict__ p2 = A.m2;
    const double* __restrict__ p3 = A.m3;
    const double rh = 1.0 / A.h;

    S qp[5][8], qm[5][8];
    S in[8];
    auto fetch = [&](int s) {
        const long long q = base + (long long)s * st;
        in[0] = pd[q]; in[1] = pu[q]; in[2] = pv[q]; in[3] = pw[q]; in[4] = pT[q];
        in[5] = S(p1[q]); in[6] = S(p2[q]); in[7] = S(p3[q]);
    };
    auto split_into = [&](int slot) {
        const S cc = sqrt_s(in[4]) * S(A.c.rama);  // sound_speed_kernel, src/commen_kernel.cu:144-153
        S fp[5], fm[5];
        sw_split<S>(in[0] * S(A.c.tmp0), in[1], in[2], in[3], cc, in[5], in[6], in[7], A.c, fp, fm);
#pragma unroll
        for (int m = 0; m < 5; m++) { qp[m][slot] = fp[m]; qm[m][slot] = fm[m]; }
    };
#pragma unroll
    for (int t = 0; t < 7; t++) {
        fetch(s0 - 4 + t);
        split_into(t);
    }
    fetch(s0 + 3);
    S pp[5] = {0, 0, 0, 0, 0}, pm[5] = {0, 0, 0, 0, 0};
    for (int j = s0 - 1; j < s1; j++) {
        split_into(7);
        if (j + 5 <= n_dir + 3) fetch(j + 5);
        int hy = 0;
        if (SCH == OCFD_HYBRIDAUTO) hy = A.flag[base + (long long)j * st];
        S hp[5], hm[5];
#pragma unroll
        for (int m = 0; m < 5; m++)
            face_pm<S, double, SCH>(qp[m], qm[m], j + 1, j - n_dir, A.bl, A.br, hy, S(A.c.tv_lim), S(A.c.tv_max),
                                    hp[m], hm[m]);
        if (j >= s0 && own) {
            const long long q = base + (long long)j * st;
            const bool zp = A.nrp && j == 0;
            const bool zm = A.nrm && j == n_dir - 1;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const double tP = zp ? 0.0 : -double(hp[m] - pp[m]) * rh;
                const double tM = zm ? 0.0 : -double(hm[m] - pm[m]) * rh;
                S* dv = A.du + (long long)m * A.fs;
                double acc = double(dv[q]);
                acc = acc + tP;
                acc = acc + tM;
                dv[q] = S(acc);
            }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            pp[m] = hp[m];
            pm[m] = hm[m];
#pragma unroll
            for (int t = 0; t < 7; t++) { qp[m][t] = qp[m][t + 1]; qm[m][t] = qm[m][t + 1]; }
        }
    }
}

// ---------------------------------------------------------------------------
// Characteristic-space reconstruction (CHARTERIC = 1):
// OCFD_weno7_SYMBO_character_P/M_kernel and
// OCFD_HybridAuto_character_P/M_Jameson_kernel (src/OCFD_flux_charteric.cu:
// 754-908, 1000-1268).  Same three phases as the physical sweeps, on one
// tile shape for all directions; per face j+1/2 in phase 2:
//   * the face state (means of u, v, w, c and the cofactor row over nodes j
//     and j+1) gives the left and right eigenvector matrices L and R
//     (get_para_charteric_*, flux_charteric_ptoc/ctop_kernel, :321-752);
//   * the eight f+ vectors of the stencil are projected with L, the five
//     characteristic variables reconstructed with the closures of the
//     physical path, and R maps the result back; the same for f-.  A hybrid
//     face with flag 1 (UP7) stays in physical space.
// The face state, L and R are FP64 on both paths.  Phase 3 follows
// put_du_character_p/m_kernel (:16-119): with a closure on one face only, the
// boundary node of that face takes a physical-space difference of the split
// flux (OCFD_bound_character_kernel_p/m, :121-319) instead of the face
// difference, the non-reflecting zero applies only there, and the low-side
// M difference is not taken in y (the reference's test compares the y flag
// with the z value).
// ---------------------------------------------------------------------------
struct CharFace {
    double L[25], R[25];
};
__device__ __forceinline__ void char_face(double u0, double v0, double w0, double c0, double a10, double a20,
                                          double a30, double u1, double v1, double w1, double c1, double a11,
                                          double a21, double a31, double gamma, CharFace& F) {
    const double u = (u0 + u1) * 0.5, v = (v0 + v1) * 0.5, w = (w0 + w1) * 0.5, c = (c0 + c1) * 0.5;
    const double a1 = (a10 + a11) * 0.5, a2 = (a20 + a21) * 0.5, a3 = (a30 + a31) * 0.5;
    double ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
    const double n1 = a1 / ss, n2 = a2 / ss, n3 = a3 / ss;
    double l1, l2, l3;
    if (fabs(n3) <= fabs(n2)) {
        ss = sqrt(n1 * n1 + n2 * n2);
        l1 = -n2 / ss; l2 = n1 / ss; l3 = 0.0;
    } else {
        ss = sqrt(n1 * n1 + n3 * n3);
        l1 = -n3 / ss; l2 = 0.0; l3 = n1 / ss;
    }
    const double m1 = n2 * l3 - n3 * l2, m2 = n3 * l1 - n1 * l3, m3 = n1 * l2 - n2 * l1;
    const double KK = (gamma - 1.0) / (c * c);
    const double X1 = 1.0 / (2.0 * c);
    const double un = u * n1 + v * n2 + w * n3;
    const double v2 = (u * u + v * v + w * w) * 0.5;
    double* L = F.L;
    L[0] = 1.0 - KK * v2; L[1] = KK * u; L[2] = KK * v; L[3] = KK * w; L[4] = -KK;
    L[5] = -(u * l1 + v * l2 + w * l3); L[6] = l1; L[7] = l2; L[8] = l3; L[9] = 0.0;
    L[10] = -(u * m1 + v * m2 + w * m3); L[11] = m1; L[12] = m2; L[13] = m3; L[14] = 0.0;
    L[15] = KK * v2 * 0.5 + X1 * un;
    L[16] = -X1 * n1 - KK * u * 0.5; L[17] = -X1 * n2 - KK * v * 0.5; L[18] = -X1 * n3 - KK * w * 0.5;
    L[19] = KK * 0.5;
    L[20] = L[15] - 2 * X1 * un;
    L[21] = 2 * X1 * n1 + L[16]; L[22] = 2 * X1 * n2 + L[17]; L[23] = 2 * X1 * n3 + L[18]; L[24] = L[19];
    const double H = v2 + 1.0 / KK;
    double* R = F.R;
    R[0] = 1.0; R[1] = 0.0; R[2] = 0.0; R[3] = 1.0; R[4] = 1.0;
    R[5] = u; R[6] = l1; R[7] = m1; R[8] = u - c * n1; R[9] = 2 * u - R[8];
    R[10] = v; R[11] = l2; R[12] = m2; R[13] = v - c * n2; R[14] = 2 * v - R[13];
    R[15] = w; R[16] = l3; R[17] = m3; R[18] = w - c * n3; R[19] = 2 * w - R[18];
    R[20] = v2; R[21] = u * l1 + v * l2 + w * l3; R[22] = u * m1 + v * m2 + w * m3;
    R[23] = H - c * un; R[24] = 2 * H - R[23];
}
__device__ __forceinline__ void mv5(const double* M, const double* x, double* y) {
#pragma unroll
    for (int i = 0; i < 5; i++)
        y[i] = M[5 * i] * x[0] + M[5 * i + 1] * x[1] + M[5 * i + 2] * x[2] + M[5 * i + 3] * x[3] + M[5 * i + 4] * x[4];
}

template <int DIR> struct CharTile { static constexpr int NS = 24, NA = 8; };
template <> struct CharTile<0> { static constexpr int NS = 56, NA = 4; };

template <int DIR, int SCH, bool FIRST, typename S>
__global__ void __launch_bounds__(256) k_char(SweepArgsT<S> A) {
    constexpr int NS = CharTile<DIR>::NS, NA = CharTile<DIR>::NA, NN = NS + 8, NF = NS + 1;
    static_assert(NN * NA == 256 && NF * NA <= 256 && (DIR != 0 || (NA == 4 && NF <= 64)),
                  "one node and at most one face per thread; x: one row per wave");
    __shared__ S sh[10 * NN * NA];
    __shared__ S shb[2][10][NA];  // boundary differences: [low/high][P 0..4, M 5..9][column]
    const Geo& g = A.g;
    const int tid = threadIdx.x;
    const Blk B = xcd_block();
    const int s0 = (DIR == 0 ? B.x : B.y) * NS, a0 = (DIR == 0 ? B.y : B.x) * NA, k2 = B.z;
    const int n_dir = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int n_acr = DIR == 0 ? g.ny : g.nx;
    auto lin = [&](int s, int a) -> long long {
        return DIR == 0 ? gidx(g, s, a, k2) : DIR == 1 ? gidx(g, a, s, k2) : gidx(g, a, k2, s);
    };
    // lanes run along x: along the sweep for x, across it for y and z
    auto at = [](int v, int n, int a) { return DIR == 0 ? (v * NA + a) * NN + n : (v * NN + n) * NA + a; };
    auto fat = [](int v, int f, int a) { return DIR == 0 ? (v * NA + a) * NF + f : (v * NF + f) * NA + a; };
    const int fyc = A.bl && A.br ? 3 : A.bl ? 1 : A.br ? 2 : 0;  // closures: both, low, high, none

    // ---- phase 1: split into LDS
    {
        const int n = DIR == 0 ? tid % NN : tid / NA, a = DIR == 0 ? tid / NN : tid % NA;
        const int s = s0 - 4 + n, ac = a0 + a;
        S fp[5], fm[5];
        if (s <= n_dir + 3 && ac < n_acr) {
            const long long q = lin(s, ac);
            const S d = A.d[q], u = A.u[q], v = A.v[q], w = A.w[q], T = A.T[q];
            const S cc = sqrt_s(T) * S(A.c.rama);
            sw_split<S>(d * S(A.c.tmp0), u, v, w, cc, S(A.m1[q]), S(A.m2[q]), S(A.m3[q]), A.c, fp, fm);
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) { fp[m] = S(0); fm[m] = S(0); }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[at(m, n, a)] = fp[m];
            sh[at(5 + m, n, a)] = fm[m];
        }
    }
    __syncthreads();

    // ---- phase 2: one face per thread
    S hp[5], hm[5];
    // x: wave a handles row a (lane = face along the row); y, z: lanes across
    const bool fthread = DIR == 0 ? (tid & 63) < NF : tid < NF * NA;
    const int f_loc = DIR == 0 ? tid & 63 : tid / NA, a_loc = DIR == 0 ? tid >> 6 : tid % NA;
    const int j = s0 - 1 + f_loc, ac = a0 + a_loc;
    const bool fvalid = fthread && j <= n_dir - 1 && ac < n_acr;
    if (fvalid) {
        auto ld = [&](int v, int k) { return sh[at(v, f_loc + k, a_loc)]; };
        int hy = 0;
        if (SCH == OCFD_HYBRIDAUTO) hy = A.flag[lin(j, ac)];
        const int c = j + 1, r = j - n_dir;
        if (SCH == OCFD_HYBRIDAUTO && hy == 1) {
            face_all<S, double, SCH>(ld, c, r, A.bl, A.br, hy, S(A.c.tv_lim), S(A.c.tv_max), hp, hm);
        } else {
            const long long q0 = lin(j, ac), q1 = lin(j + 1, ac);
            const double rama = A.c.rama;
            CharFace F;
            char_face(double(A.u[q0]), double(A.v[q0]), double(A.w[q0]), sqrt(double(A.T[q0])) * rama, A.m1[q0],
                      A.m2[q0], A.m3[q0], double(A.u[q1]), double(A.v[q1]), double(A.w[q1]),
                      sqrt(double(A.T[q1])) * rama, A.m1[q1], A.m2[q1], A.m3[q1], A.c.gamma, F);
            S cs[5][8];
            double x[5], y[5], hc[5];
            S hcs[5];
#pragma unroll
            for (int k = 0; k < 8; k++) {
#pragma unroll
                for (int m = 0; m < 5; m++) x[m] = double(ld(m, k));
                mv5(F.L, x, y);
#pragma unroll
                for (int m = 0; m < 5; m++) cs[m][k] = S(y[m]);
            }
            face_all_p<S, double, SCH>([&](int v, int k) { return cs[v][k]; }, c, r, A.bl, A.br, hy,
                                       S(A.c.tv_lim), S(A.c.tv_max), hcs);
#pragma unroll
            for (int m = 0; m < 5; m++) hc[m] = double(hcs[m]);
            mv5(F.R, hc, y);
#pragma unroll
            for (int m = 0; m < 5; m++) hp[m] = S(y[m]);
#pragma unroll
            for (int k = 0; k < 8; k++) {
#pragma unroll
                for (int m = 0; m < 5; m++) x[m] = double(ld(5 + m, k));
                mv5(F.L, x, y);
#pragma unroll
                for (int m = 0; m < 5; m++) cs[m][k] = S(y[m]);
            }
            face_all_m<S, double, SCH>([&](int v, int k) { return cs[v - 5][k]; }, c, r, A.bl, A.br, hy, hcs);
#pragma unroll
            for (int m = 0; m < 5; m++) hc[m] = double(hcs[m]);
            mv5(F.R, hc, y);
#pragma unroll
            for (int m = 0; m < 5; m++) hm[m] = S(y[m]);
        }
        // physical-space differences of the boundary nodes (node j-3+k at k)
        if (j == 0 && fyc == 1) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                shb[0][m][a_loc] = ld(m, 4) - ld(m, 3);
                const S f0 = ld(5 + m, 3), f1 = ld(5 + m, 4), f2 = ld(5 + m, 5);
                shb[0][5 + m][a_loc] =
                    (f1 - S(0.5) * d_minmod2(f2 - f1, f1 - f0)) - (f0 - S(0.5) * d_minmod2(f1 - f0, f1 - f0));
            }
        }
        if (j == n_dir - 1 && fyc == 2) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const S a = ld(m, 3), b = ld(m, 2), cc = ld(m, 1);
                shb[1][m][a_loc] = (a + S(0.5) * d_minmod2(a - b, a - b)) - (b + S(0.5) * d_minmod2(a - b, b - cc));
                shb[1][5 + m][a_loc] = ld(5 + m, 3) - ld(5 + m, 2);
            }
        }
    }
    __syncthreads();
    if (fthread) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[fat(m, f_loc, a_loc)] = fvalid ? hp[m] : S(0);
            sh[fat(5 + m, f_loc, a_loc)] = fvalid ? hm[m] : S(0);
        }
    }
    __syncthreads();

    // ---- phase 3: du at nodes, P then M
    if (DIR == 0 ? (tid & 63) < NS : tid < NS * NA) {
        const int n = DIR == 0 ? tid & 63 : tid / NA, a = DIR == 0 ? tid >> 6 : tid % NA;
        const int i = s0 + n, ai = a0 + a;
        if (i < n_dir && ai < n_acr) {
            const long long q = lin(i, ai);
            const double rh = 1.0 / A.h;
            const bool lowP = fyc == 1 && i == 0, highP = fyc == 2 && i == n_dir - 1;
            const bool highM = fyc == 2 && i == n_dir - 1, lowM = fyc == 1 && i == 0 && DIR != 1;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                double tP, tM;
                if (lowP) tP = A.nrp ? 0.0 : -double(shb[0][m][a]) * rh;
                else if (highP) tP = -double(shb[1][m][a]) * rh;
                else tP = -double(sh[fat(m, n + 1, a)] - sh[fat(m, n, a)]) * rh;
                if (highM) tM = A.nrm ? 0.0 : -double(shb[1][5 + m][a]) * rh;
                else if (lowM) tM = -double(shb[0][5 + m][a]) * rh;
                else tM = -double(sh[fat(5 + m, n + 1, a)] - sh[fat(5 + m, n, a)]) * rh;
                S* dv = A.du + (long long)m * A.fs;
                double acc = FIRST ? 0.0 : double(dv[q]);
                acc = acc + tP;
                acc = acc + tM;
                dv[q] = S(acc);
            }
        }
    }
}

template <int DIR, int SCH, typename S>
void launch_char_dir(const SweepArgsT<S>& a, bool first, hipStream_t st) {
    constexpr int NS = CharTile<DIR>::NS, NA = CharTile<DIR>::NA;
    const Geo& g = a.g;
    const int n_dir = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int n_acr = DIR == 0 ? g.ny : g.nx;
    const int third = DIR == 0 ? g.nz : DIR == 1 ? g.nz : g.ny;
    const dim3 grid = DIR == 0 ? dim3((n_dir + NS - 1) / NS, (n_acr + NA - 1) / NA, third)
                               : dim3((n_acr + NA - 1) / NA, (n_dir + NS - 1) / NS, third);
    if (first) hipLaunchKernelGGL((k_char<DIR, SCH, true, S>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_char<DIR, SCH, false, S>), grid, dim3(256), 0, st, a);
}
template <int SCH, typename S>
void launch_char_scheme(int dir, const SweepArgsT<S>& a, bool first, hipStream_t st) {
    if (dir == 0) launch_char_dir<0, SCH>(a, first, st);
    else if (dir == 1) launch_char_dir<1, SCH>(a, first, st);
    else launch_char_dir<2, SCH>(a, first, st);
}

#ifndef OCFD_LMARCH
#define OCFD_LMARCH 32  // output nodes per marching chunk (k_march)
#endif
// y/z tile shapes (output nodes x columns).  Measured on the 512x128x64 ramp
// (FP64, 28 x 8 tiles): NS = 32 with 320 threads (no partial tile on lines
// of 64 and 128 nodes) is 10-20 % slower than NS = 28 with 256 threads; a
// per-block queue that hands
// the rare WENO faces of hybrid_auto to all threads, and a form with the face
// pass fused into the node pass by cross-lane shuffles (one more resident
// wave per SIMD), were no faster.
// FP64: 14 nodes x 16 columns (128-byte row segments): on the 1024x384x256
// swept ramp the y sweep 3.58-3.74 -> 3.43-3.53 ms and the z sweep
// 3.91-4.00 -> 3.53-3.58 ms against 28 x 8 (12 x 16 in between; 56 x 4, 4.4
// and 6.2 ms), although it splits 1.57 nodes per output node against 1.29.
#ifndef OCFD_YZ_NS_D
#define OCFD_YZ_NS_D 14
#endif
#ifndef OCFD_YZ_NA_D
#define OCFD_YZ_NA_D 16
#endif
#ifndef OCFD_YZ_NS_F
#define OCFD_YZ_NS_F 15
#endif
#ifndef OCFD_YZ_NA_F
#define OCFD_YZ_NA_F 16
#endif
#ifndef OCFD_MARCH_F32
#define OCFD_MARCH_F32 0  // y/z sweeps of the float path: 1 = k_march, 0 = k_tile
#endif
constexpr int LMARCH = OCFD_LMARCH;
template <typename S> struct YZ { static constexpr int NS = OCFD_YZ_NS_D, NA = OCFD_YZ_NA_D; };
template <> struct YZ<float> { static constexpr int NS = OCFD_YZ_NS_F, NA = OCFD_YZ_NA_F; };

template <int DIR, int SCH, typename S>
void launch_dir(const SweepArgsT<S>& a, bool first, hipStream_t st) {
    const Geo& g = a.g;
    if constexpr (DIR == 0) {
        dim3 grid((g.nx + XS - 1) / XS, (g.ny + XA - 1) / XA, g.nz);
        if constexpr (SCH == OCFD_HYBRIDAUTO) {
            if (a.mlp.P && first) {
                hipLaunchKernelGGL((k_sweep<SCH, true, S, true>), grid, dim3(64 * XA), 0, st, a);
                return;
            }
        }
        if (first) hipLaunchKernelGGL((k_sweep<SCH, true, S>), grid, dim3(64 * XA), 0, st, a);
        else hipLaunchKernelGGL((k_sweep<SCH, false, S>), grid, dim3(64 * XA), 0, st, a);
    } else {
        const int n_dir = DIR == 1 ? g.ny : g.nz, n_acr = DIR == 1 ? g.nz : g.ny;
        if constexpr (std::is_same<S, float>::value && OCFD_MARCH_F32) {
            dim3 grid((g.nx + 63) / 64, (n_acr + 3) / 4, (n_dir + LMARCH - 1) / LMARCH);
            hipLaunchKernelGGL((k_march<DIR, SCH, LMARCH, S>), grid, dim3(64, 4), 0, st, a);
        } else {
            constexpr int NS = YZ<S>::NS, NA = YZ<S>::NA;
            dim3 grid((g.nx + NA - 1) / NA, (n_dir + NS - 1) / NS, n_acr);
            hipLaunchKernelGGL((k_tile<DIR, SCH, NS, NA, S>), grid, dim3(256), 0, st, a);
        }
    }
}

template <int SCH, typename S>
void launch_scheme(int dir, const SweepArgsT<S>& a, bool first, hipStream_t st) {
    if (dir == 0) launch_dir<0, SCH>(a, first, st);
    else if (dir == 1) launch_dir<1, SCH>(a, first, st);
    else launch_dir<2, SCH>(a, first, st);
}

}  // namespace

template <typename S>
int launch_sweep(int scheme, int dir, const SweepArgsT<S>& a, bool first, hipStream_t st) {
    switch (scheme) {
    case OCFD_UP7: launch_scheme<OCFD_UP7>(dir, a, first, st); break;
    case OCFD_WENO5: launch_scheme<OCFD_WENO5>(dir, a, first, st); break;
    case OCFD_WENO7: launch_scheme<OCFD_WENO7>(dir, a, first, st); break;
    case OCFD_WENO7_SYMBO: launch_scheme<OCFD_WENO7_SYMBO>(dir, a, first, st); break;
    case OCFD_WENO7_SYMBO_LIM: launch_scheme<OCFD_WENO7_SYMBO_LIM>(dir, a, first, st); break;
    case OCFD_NND2: launch_scheme<OCFD_NND2>(dir, a, first, st); break;
    case OCFD_OMP6_HR: launch_scheme<OCFD_OMP6_HR>(dir, a, first, st); break;
    case OCFD_OMP6_LD: launch_scheme<OCFD_OMP6_LD>(dir, a, first, st); break;
    case OCFD_OMP6_CD8: launch_scheme<OCFD_OMP6_CD8>(dir, a, first, st); break;
    case OCFD_HYBRIDAUTO: launch_scheme<OCFD_HYBRIDAUTO>(dir, a, first, st); break;
    default: return -1;
    }
    return 0;
}
template int launch_sweep<double>(int, int, const SweepArgsT<double>&, bool, hipStream_t);
template int launch_sweep<float>(int, int, const SweepArgsT<float>&, bool, hipStream_t);

template <typename S>
int launch_sweep_char(int scheme, int dir, const SweepArgsT<S>& a, bool first, hipStream_t st) {
    switch (scheme) {
    case OCFD_WENO7_SYMBO: launch_char_scheme<OCFD_WENO7_SYMBO>(dir, a, first, st); break;
    case OCFD_WENO7_SYMBO_LIM: launch_char_scheme<OCFD_WENO7_SYMBO_LIM>(dir, a, first, st); break;
    case OCFD_HYBRIDAUTO: launch_char_scheme<OCFD_HYBRIDAUTO>(dir, a, first, st); break;
    default: return -1;
    }
    return 0;
}
template int launch_sweep_char<double>(int, int, const SweepArgsT<double>&, bool, hipStream_t);
template int launch_sweep_char<float>(int, int, const SweepArgsT<float>&, bool, hipStream_t);
