// k_inviscid.hip - fused inviscid sweep per direction (gfx950).
//
// One kernel per direction replaces the reference's split kernel
// (split_Jac3d_Stager_Warming_ker, src/OCFD_split.cu:145-198) and its ten
// derivative launches (OCFD_d?1 / OCFD_d?2, src/OCFD_Schemes_Choose.cu:88-765,
// each launching one P or M kernel per variable with an atomicAdd into du,
// src/OCFD_Schemes.cu:293-381):
//
//   phase 1  load rho,u,v,w,T and the direction's cofactors for a tile of
//            nodes (+4 ghost nodes each side along the sweep), compute the
//            sound speed and the Steger-Warming f+/f- of all 5 variables,
//            stage them in LDS (they never touch HBM);
//   phase 2  one thread per half node: P and M reconstructions of all 5
//            variables with the reference's boundary closures;
//   phase 3  one thread per node: du += -Ajac*(hP(i+1/2)-hP(i-1/2))/h, then
//            += the M term, in the reference's order (x P, x M, y P, ...), with
//            no atomics.
//
// Tiles: x sweep 56 outputs x 4 rows (64 nodes per row = one wave), y and z
// sweeps 16 x-lanes (one 128-byte row segment) x 15 outputs.
#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

template <int DIR> struct Tile;
template <> struct Tile<0> { static constexpr int NS = 56, NA = 4; };
template <> struct Tile<1> { static constexpr int NS = 15, NA = 16; };
template <> struct Tile<2> { static constexpr int NS = 15, NA = 16; };

// LDS element (v, n, a): node-contiguous for the x sweep (consecutive lanes
// hold consecutive x), lane-contiguous for y/z sweeps.
template <int DIR, int NN, int NA>
__device__ __forceinline__ int lds_at(int v, int n, int a) {
    if (DIR == 0) return (v * NA + a) * NN + n;
    return (v * NN + n) * NA + a;
}

// Steger-Warming split of one node (src/OCFD_split.cu:92-142).
template <typename R>
__device__ __forceinline__ void sw_split(R tmp0d, R u, R v, R w, R cc, R A1, R A2, R A3, const Consts& c,
                                         R* fp, R* fm) {
    R ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
    R E1 = A1 * u + A2 * v + A3 * w;
    R E2 = E1 - cc * ss;
    R E3 = E1 + cc * ss;
    ss = R(1.0) / ss;
    A1 *= ss; A2 *= ss; A3 *= ss;
    const R eps = R(c.epsl_sw);
    R E1P = (E1 + sqrt(E1 * E1 + eps * eps)) * R(0.50);
    R E2P = (E2 + sqrt(E2 * E2 + eps * eps)) * R(0.50);
    R E3P = (E3 + sqrt(E3 * E3 + eps * eps)) * R(0.50);
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
template <typename R, int SCH>
__device__ __forceinline__ void face_pm(const R* gp, const R* gm, int c, int r, int bl, int br, int hy,
                                        R tvlim, R tvmax, R& hp, R& hm) {
    constexpr int kaP = Ext<SCH, 0>::ka, kbP = Ext<SCH, 0>::kb;
    constexpr int kaM = Ext<SCH, 1>::ka, kbM = Ext<SCH, 1>::kb;
    // ---- P (OCFD_bound_scheme_kernel_p, src/OCFD_bound_Scheme.cu:739-819)
    if (bl && c <= -kaP) {
        if (c == 0) hp = gp[4];
        else if (c == 1) hp = weno5_up_left(gp + 1);
        else if (c == 2) hp = weno5_up_leftplus(gp + 1);
        else hp = weno5_up(gp + 1);
    } else if (br && r >= -kbP) {
        if (r == -1) hp = gp[3] + R(0.5) * d_minmod2(gp[3] - gp[2], gp[3] - gp[2]);
        else if (r == -2) hp = weno5_up_rightplus(gp + 1);
        else hp = weno5_up(gp + 1);
    } else {
        if (SCH == OCFD_WENO7) hp = weno7_up(gp);
        else if (SCH == OCFD_UP7) hp = up7_up(gp);
        else if (SCH == OCFD_WENO5) hp = weno5_up(gp + 1);
        else if (SCH == OCFD_NND2) hp = nnd2_up(gp + 2);
        else if (SCH == OCFD_WENO7_SYMBO) hp = symbo_up(gp, 0, tvlim, tvmax);
        else if (SCH == OCFD_WENO7_SYMBO_LIM) hp = symbo_up(gp, 1, tvlim, tvmax);
        else if (SCH == OCFD_OMP6_HR) hp = omp6_up(0, gp);
        else if (SCH == OCFD_OMP6_LD) hp = omp6_up(1, gp);
        else if (SCH == OCFD_OMP6_CD8) hp = omp6_up(2, gp);
        else {  // hybrid_auto, src/OCFD_Schemes_hybrid_auto.cu:911-919
            if (hy == 1) hp = up7_up(gp);
            else if (hy == 2) hp = weno7_up(gp);
            else hp = weno5_up(gp + 1);
        }
    }
    // ---- M (OCFD_bound_scheme_kernel_m, src/OCFD_bound_Scheme.cu:855-944)
    if (bl && c < -kaM) {
        if (c == 0) hm = weno5_dn_left(gm + 2);
        else if (c == 1) hm = weno5_dn_leftplus(gm + 2);
        else hm = weno5_dn(gm + 2);
    } else if (br && r >= -kbM - 1) {
        if (r == -1) hm = gm[3];
        else if (r == -2) hm = gm[4] - R(0.5) * d_minmod2(gm[4] - gm[3], gm[4] - gm[3]);
        else if (r == -3) hm = weno5_dn_rightplus(gm + 2);
        else hm = weno5_dn(gm + 2);
    } else {
        if (SCH == OCFD_WENO7) hm = weno7_dn(gm + 1);
        else if (SCH == OCFD_UP7) hm = up7_dn(gm);
        else if (SCH == OCFD_WENO5) hm = weno5_dn(gm + 2);
        else if (SCH == OCFD_NND2) hm = nnd2_dn(gm + 3);
        else if (SCH == OCFD_WENO7_SYMBO || SCH == OCFD_WENO7_SYMBO_LIM) hm = symbo_dn(gm);
        else if (SCH == OCFD_OMP6_HR) hm = omp6_dn(0, gm);
        else if (SCH == OCFD_OMP6_LD) hm = omp6_dn(1, gm);
        else if (SCH == OCFD_OMP6_CD8) hm = omp6_dn(2, gm);
        else {  // src/OCFD_Schemes_hybrid_auto.cu:1008-1016
            if (hy == 1) hm = up7_dn(gm);
            else if (hy == 2) hm = weno7_dn(gm + 1);
            else hm = weno5_dn(gm + 2);
        }
    }
}

template <int DIR, int SCH, bool FIRST>
__global__ void __launch_bounds__(256) k_sweep(SweepArgs A) {
    constexpr int NS = Tile<DIR>::NS, NA = Tile<DIR>::NA, NN = NS + 8, NF = NS + 1;
    static_assert(NF * NA <= 256, "one face per thread");
    __shared__ double sh[10 * NN * NA];
    const Geo& g = A.g;
    const int tid = threadIdx.x;
    // tile origin: s0 along the sweep, a0 across; k2 = the third index
    int s0, a0, k2;
    if (DIR == 0) { s0 = blockIdx.x * NS; a0 = blockIdx.y * NA; k2 = blockIdx.z; }
    else { a0 = blockIdx.x * NA; s0 = blockIdx.y * NS; k2 = blockIdx.z; }
    const int n_dir = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int n_acr = DIR == 0 ? g.ny : g.nx;
    const long long sstride = DIR == 0 ? 1 : DIR == 1 ? (long long)g.ldx : g.sz;
    // point (along s, across a) -> linear index
    auto lin = [&](int s, int a) -> long long {
        if (DIR == 0) return gidx(g, s, a, k2);
        if (DIR == 1) return gidx(g, a, s, k2);
        return gidx(g, a, k2, s);
    };

    // ---- phase 1: split of NN x NA nodes into LDS
    for (int idx = tid; idx < NN * NA; idx += 256) {
        int n, a;
        if (DIR == 0) { n = idx % NN; a = idx / NN; } else { a = idx % NA; n = idx / NA; }
        const int s = s0 - 4 + n, ac = a0 + a;
        double fp[5], fm[5];
        if (s <= n_dir + 3 && ac < n_acr) {
            const long long q = lin(s, ac);
            const double d = A.d[q], u = A.u[q], v = A.v[q], w = A.w[q], T = A.T[q];
            const double cc = sqrt(T) / A.c.ama;  // sound_speed_kernel, src/commen_kernel.cu:144-153
            sw_split<double>(d * A.c.tmp0, u, v, w, cc, A.m1[q], A.m2[q], A.m3[q], A.c, fp, fm);
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) { fp[m] = 0.0; fm[m] = 0.0; }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[lds_at<DIR, NN, NA>(m, n, a)] = fp[m];
            sh[lds_at<DIR, NN, NA>(5 + m, n, a)] = fm[m];
        }
    }
    __syncthreads();

    // ---- phase 2: one half node per thread
    double hp[5], hm[5];
    int f_loc = 0, a_loc = 0;
    bool fvalid = false;
    if (tid < NF * NA) {
        if (DIR == 0) { f_loc = tid % NF; a_loc = tid / NF; } else { a_loc = tid % NA; f_loc = tid / NA; }
        const int j = s0 - 1 + f_loc;  // face j+1/2
        const int ac = a0 + a_loc;
        fvalid = (j <= n_dir - 1) && (ac < n_acr);
        if (fvalid) {
            int hy = 0;
            if (SCH == OCFD_HYBRIDAUTO) hy = A.flag[lin(j, ac)];
            const int c = j + 1, r = j - n_dir;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                double gp[8], gm[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    gp[k] = sh[lds_at<DIR, NN, NA>(m, f_loc + k, a_loc)];
                    gm[k] = sh[lds_at<DIR, NN, NA>(5 + m, f_loc + k, a_loc)];
                }
                face_pm<double, SCH>(gp, gm, c, r, A.bl, A.br, hy, A.c.tv_lim, A.c.tv_max, hp[m], hm[m]);
            }
        }
    }
    __syncthreads();
    if (tid < NF * NA) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            sh[(m * NF + f_loc) * NA + a_loc] = fvalid ? hp[m] : 0.0;
            sh[((5 + m) * NF + f_loc) * NA + a_loc] = fvalid ? hm[m] : 0.0;
        }
    }
    __syncthreads();

    // ---- phase 3: du at nodes, in the reference's order P then M
    if (tid < NS * NA) {
        int n, a;
        if (DIR == 0) { n = tid % NS; a = tid / NS; } else { a = tid % NA; n = tid / NA; }
        const int i = s0 + n, ac = a0 + a;
        if (i < n_dir && ac < n_acr) {
            const long long q = lin(i, ac);
            const double ajac = A.jac[q];
            const bool zp = A.nrp && i == 0;
            const bool zm = A.nrm && i == n_dir - 1;
            const double rh = 1.0 / A.h;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const double pr = sh[(m * NF + n + 1) * NA + a], pl = sh[(m * NF + n) * NA + a];
                const double mr = sh[((5 + m) * NF + n + 1) * NA + a], ml = sh[((5 + m) * NF + n) * NA + a];
                const double tP = zp ? 0.0 : -ajac * (pr - pl) * rh;
                const double tM = zm ? 0.0 : -ajac * (mr - ml) * rh;
                double* dv = A.du + (long long)m * A.fs;
                double acc = FIRST ? 0.0 : dv[q];
                acc = acc + tP;
                acc = acc + tM;
                dv[q] = acc;
            }
        }
    }
    (void)sstride;
}

// y and z sweeps: one thread per (x, across) column marching along the sweep
// direction over a chunk of L output nodes.  The f+/f- of the 8 nodes a face
// needs (j-3..j+4) live in a register queue: each node is split once per chunk
// (plus the 8 warm-up nodes), rows of 64 x-consecutive nodes are read by one
// wave (512 contiguous bytes), and there is no LDS and no barrier.  The next
// node's inputs are fetched one iteration ahead.
template <int DIR, int SCH, int L>
__global__ void __launch_bounds__(256) k_march(SweepArgs A) {
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
    const double* __restrict__ pd = A.d;
    const double* __restrict__ pu = A.u;
    const double* __restrict__ pv = A.v;
    const double* __restrict__ pw = A.w;
    const double* __restrict__ pT = A.T;
    const double* __restrict__ p1 = A.m1;
    const double* __restrict__ p2 = A.m2;
    const double* __restrict__ p3 = A.m3;
    const double rh = 1.0 / A.h;

    double qp[5][8], qm[5][8];
    double in[8];
    auto fetch = [&](int s) {
        const long long q = base + (long long)s * st;
        in[0] = pd[q]; in[1] = pu[q]; in[2] = pv[q]; in[3] = pw[q]; in[4] = pT[q];
        in[5] = p1[q]; in[6] = p2[q]; in[7] = p3[q];
    };
    auto split_into = [&](int slot) {
        const double cc = sqrt(in[4]) / A.c.ama;  // sound_speed_kernel, src/commen_kernel.cu:144-153
        double fp[5], fm[5];
        sw_split<double>(in[0] * A.c.tmp0, in[1], in[2], in[3], cc, in[5], in[6], in[7], A.c, fp, fm);
#pragma unroll
        for (int m = 0; m < 5; m++) { qp[m][slot] = fp[m]; qm[m][slot] = fm[m]; }
    };
#pragma unroll
    for (int t = 0; t < 7; t++) {
        fetch(s0 - 4 + t);
        split_into(t);
    }
    fetch(s0 + 3);
    double pp[5] = {0, 0, 0, 0, 0}, pm[5] = {0, 0, 0, 0, 0};
    for (int j = s0 - 1; j < s1; j++) {
        split_into(7);
        if (j + 5 <= n_dir + 3) fetch(j + 5);
        int hy = 0;
        if (SCH == OCFD_HYBRIDAUTO) hy = A.flag[base + (long long)j * st];
        double hp[5], hm[5];
#pragma unroll
        for (int m = 0; m < 5; m++)
            face_pm<double, SCH>(qp[m], qm[m], j + 1, j - n_dir, A.bl, A.br, hy, A.c.tv_lim, A.c.tv_max, hp[m],
                                 hm[m]);
        if (j >= s0 && own) {
            const long long q = base + (long long)j * st;
            const double ajac = A.jac[q];
            const bool zp = A.nrp && j == 0;
            const bool zm = A.nrm && j == n_dir - 1;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const double tP = zp ? 0.0 : -ajac * (hp[m] - pp[m]) * rh;
                const double tM = zm ? 0.0 : -ajac * (hm[m] - pm[m]) * rh;
                double* dv = A.du + (long long)m * A.fs;
                double acc = dv[q];
                acc = acc + tP;
                acc = acc + tM;
                dv[q] = acc;
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

#ifndef OCFD_LMARCH
#define OCFD_LMARCH 32  // output nodes per marching chunk (y and z sweeps)
#endif
constexpr int LMARCH = OCFD_LMARCH;

template <int DIR, int SCH>
void launch_dir(const SweepArgs& a, bool first, hipStream_t st) {
    const Geo& g = a.g;
    if constexpr (DIR == 0) {
        constexpr int NS = Tile<0>::NS, NA = Tile<0>::NA;
        dim3 grid((g.nx + NS - 1) / NS, (g.ny + NA - 1) / NA, g.nz);
        if (first) hipLaunchKernelGGL((k_sweep<0, SCH, true>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_sweep<0, SCH, false>), grid, dim3(256), 0, st, a);
    } else {
        const int n_dir = DIR == 1 ? g.ny : g.nz, n_acr = DIR == 1 ? g.nz : g.ny;
        dim3 grid((g.nx + 63) / 64, (n_acr + 3) / 4, (n_dir + LMARCH - 1) / LMARCH);
        hipLaunchKernelGGL((k_march<DIR, SCH, LMARCH>), grid, dim3(64, 4), 0, st, a);
    }
}

template <int SCH>
void launch_scheme(int dir, const SweepArgs& a, bool first, hipStream_t st) {
    if (dir == 0) launch_dir<0, SCH>(a, first, st);
    else if (dir == 1) launch_dir<1, SCH>(a, first, st);
    else launch_dir<2, SCH>(a, first, st);
}

}  // namespace

int launch_sweep(int scheme, int dir, const SweepArgs& a, bool first, hipStream_t st) {
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
