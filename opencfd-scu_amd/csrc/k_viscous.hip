// k_viscous.hip - viscous terms (gfx950).
//
// Restates du_vis_Jacobian3d (src/OCFD_Stream.cu:307-335) in two kernels:
//
//   k_visc_gf  : the 12 derivatives of u,v,w,T along xi/eta/zeta (CD6/CD8 with
//                the D0 closures: du_viscous_Jacobian3d_init,
//                src/OCFD_NS_Jacobian3d.cu:58-77; OCFD_CD6/CD8_kernel,
//                src/OCFD_Schemes.cu:96-156; OCFD_D0bound_scheme_kernel,
//                src/OCFD_bound_Scheme.cu:79-186) and, at the same point, the
//                contravariant viscous fluxes Ev1..4 of all three directions
//                (vis_flux_ker, src/OCFD_NS_Jacobian3d.cu:80-281).  The
//                gradients never leave registers.
//   k_visc_div : du[1..4] += Ajac*d(Ev)/dxi, then eta, then zeta (OCFD_d?0 of
//                Ev + YF_Pe_XF, src/OCFD_NS_Jacobian3d.cu:304-440,
//                src/commen_kernel.cu:158-169), no atomics; on the FP64 path
//                the three terms go to their own buffer dv instead and the
//                RK stage adds du + dv, so that the divergence need not wait
//                for the inviscid sweeps.
//   k_qcrit    : the Q criterion of the analysis events (same derivatives).
//
// S (double or float) is the type of u,v,w,T, mu, the derivatives, the
// fluxes, du and dv; the Jacobian factor of the divergence and its sums are
// double.
//
// Both use 2.5-D blocking: a block owns a TX x TY tile of (x,y) columns and
// marches through a chunk of z planes (32 on the FP64 path; 16 on the FP32
// path).  x and y neighbours come from the
// current plane staged in LDS with its 4-point halo; z neighbours come from a
// 9-plane queue each thread keeps in registers for its own column.
#include <type_traits>

#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

#ifndef OCFD_VISC_KZ_D
#define OCFD_VISC_KZ_D 32  // z planes per block, FP64 path (16 / 24 / 32 / 64: 4.57 / 4.55 / 4.46 /
                           // 4.74 ms per launch on the 1024x384x256 swept ramp)
#endif
#ifndef OCFD_VISC_KZ_F
#define OCFD_VISC_KZ_F 16  // z planes per block, FP32 path: 8 is the fastest alone, 16 the faster
                           // step beside the sweeps (2.016 -> 2.008 ms on the 512x128x64 ramp)
#endif
#ifndef OCFD_VISC_TX
#define OCFD_VISC_TX 32  // x columns per block
#endif
#ifndef OCFD_VISC_TY_D
#define OCFD_VISC_TY_D 8  // y columns per block of the gradient kernel, FP64 path (16 measured 4.45 ->
                          // 4.41 ms per launch on the 1024x384x256 swept ramp, the step 64.47 -> 64.18
                          // ms, and passes the parity tests; 2 waves per SIMD either way)
#endif
#ifndef OCFD_VISC_TY_F
#define OCFD_VISC_TY_F 8  // FP32 path: a 512-thread block would cost it its third wave per SIMD
#endif
#ifndef OCFD_VISC_PREFETCH
#define OCFD_VISC_PREFETCH 1  // fetch plane k+1 while plane k is computed
#endif
constexpr int TX = OCFD_VISC_TX;
// y columns per block of the gradient/flux kernel
template <typename S> struct TY_ { static constexpr int v = OCFD_VISC_TY_D; };
template <> struct TY_<float> { static constexpr int v = OCFD_VISC_TY_F; };
#ifndef OCFD_VDIV_KZ_D
#define OCFD_VDIV_KZ_D 32  // z planes per block of the divergence, FP64 path: 32 against 16 on the
                           // 1024x384x256 swept ramp, the divergence's share of the step 13.65 ->
                           // 13.44 % and the step 64.34 -> 64.17 ms (bench runs alternated); no
                           // change on the 512x128x64 and 1024x256x128 ramps
#endif
#ifndef OCFD_VDIV_KZ_F
#define OCFD_VDIV_KZ_F 16  // z planes per block of the divergence, FP32 path
#endif
#ifndef OCFD_VDIV_TY
#define OCFD_VDIV_TY 8  // y columns per block of the divergence
#endif
constexpr int TYD = OCFD_VDIV_TY;
// z planes per block: KZ_<S, false> for the gradient/flux kernel, KZ_<S,
// true> for the divergence
template <typename S, bool DIV = false> struct KZ_ { static constexpr int v = DIV ? OCFD_VDIV_KZ_D : OCFD_VISC_KZ_D; };
template <bool DIV> struct KZ_<float, DIV> { static constexpr int v = DIV ? OCFD_VDIV_KZ_F : OCFD_VISC_KZ_F; };
constexpr bool VISC_PREFETCH = OCFD_VISC_PREFETCH;
#ifndef OCFD_VISC_RING
#define OCFD_VISC_RING 1  // FP64 gradient kernel: halo strips fetched five planes ahead into an LDS ring
                          // (the FP32 form would lose its third wave per SIMD to the registers)
#endif
#ifndef OCFD_VISC_SKIP_CORNERS
#define OCFD_VISC_SKIP_CORNERS 1  // the gradient kernel does not load the tile's halo corners
#endif

// Central derivative at interior index i (0..n-1) from s[0..8] = values at
// i-4..i+4 (OCFD_CD6/CD8_kernel + D0 closures + put_d0_kernel).  rh = 1/h; the
// reference's divisions by 12, 60, 840 and h are multiplications by the
// reciprocals here (a difference of an ulp or two per derivative).  CL =
// false compiles the interior formula alone, for callers that know that no
// closure is in reach.
template <int SV, typename R, bool CL = true>
__device__ __forceinline__ R cd9(const R* s, int i, int n, int bl, int br, R rh) {
    const R* c = s + 4;
    R tmp;
    if (CL && bl && i <= 3) {
        if (i == 0) tmp = (c[1] - c[0]);
        else if (i == 1) tmp = (c[1] - c[-1]) * R(0.5);
        else if (i == 2) tmp = (c[-2] - R(8) * c[-1] + R(8) * c[1] - c[2]) * R(1.0 / 12.0);
        else tmp = (c[3] - c[-3] - R(9) * (c[2] - c[-2]) + R(45) * (c[1] - c[-1])) * R(1.0 / 60.0);
    } else if (CL && br && i >= n - 4) {
        const int t1 = i - n;
        if (t1 == -1) tmp = (c[0] - c[-1]);
        else if (t1 == -2) tmp = (c[1] - c[-1]) * R(0.5);
        else if (t1 == -3) tmp = (c[-2] - R(8) * c[-1] + R(8) * c[1] - c[2]) * R(1.0 / 12.0);
        else tmp = (c[3] - c[-3] - R(9) * (c[2] - c[-2]) + R(45) * (c[1] - c[-1])) * R(1.0 / 60.0);
    } else if (SV == OCFD_CD6) {
        tmp = (R(45) * (c[1] - c[-1]) - R(9) * (c[2] - c[-2]) + (c[3] - c[-3])) * R(1.0 / 60.0);
    } else {
        tmp = (R(672) * (c[1] - c[-1]) - R(168) * (c[2] - c[-2]) + R(32) * (c[3] - c[-3]) - R(3) * (c[4] - c[-4])) *
              R(1.0 / 840.0);
    }
    return tmp * rh;
}

// Tile bookkeeping shared by both kernels.
struct Tile25 {
    int i0, j0, i, j, kbeg, kend;
    bool own;
};
template <typename S, bool DIV, int TYB = TY_<S>::v>
__device__ __forceinline__ Tile25 tile_of(const Geo& g) {
    constexpr int KZ = KZ_<S, DIV>::v;
    Tile25 t;
    const Blk B = xcd_block();
    t.i0 = B.x * TX;
    t.j0 = B.y * TYB;
    t.i = t.i0 + threadIdx.x;
    t.j = t.j0 + threadIdx.y;
    t.kbeg = B.z * KZ;
    t.kend = min(t.kbeg + KZ, g.nz);
    t.own = t.i < g.nx && t.j < g.ny;
    return t;
}

// Gradients of u, v, w, T at (t.i, t.j, k) from the plane tile pl (the
// current plane with its x and y strips, row pitch PW) and the thread's z
// queue q, and at the same point the viscous fluxes Ev1..4 of the three
// directions (vis_flux_s_ker / vis_flux_e_ker / vis_flus_ev_ker).  CL =
// false: no derivative of the block's plane reaches a closure (the caller
// decides per block and plane), and the body is straight-line code.
template <int SV, typename S, int PW, bool CL>
__device__ __forceinline__ void gf_point(const ViscArgsT<S>& A, const Tile25& t, int k,
                                         const S (*pl)[PW * (TY_<S>::v + 8)],
                                         const S (*q)[9], int tx, int ty, S rhx, S rhy, S rhz) {
    const Geo& g = A.g;
    // The point's metrics are loaded first, so that the loads are in flight
    // while the derivatives are formed.  They are read once here and the
    // fluxes are written once: all go through the caches as streaming
    // (non-temporal) accesses, which leaves L2 to the plane strips that
    // neighbouring tiles share (measured 203 -> 196 us, and the divergence
    // that reads Ev next 42 -> 40 us).
    const long long qi = gidx(g, t.i, t.j, k), fs = A.fs;
    auto ld = [&](int m) { return S(__builtin_nontemporal_load(&A.M[m][qi])); };
    const S Akx = ld(M_KX), Aky = ld(M_KY), Akz = ld(M_KZ);
    const S Aix = ld(M_IX), Aiy = ld(M_IY), Aiz = ld(M_IZ);
    const S Asx = ld(M_SX), Asy = ld(M_SY), Asz = ld(M_SZ);
    const S Aj1 = ld(M_JAC);
    // Sutherland's law at the point (get_Amu_kernal, src/commen_kernel.cu:
    // 122-131) from its T (the centre of the z queue), in FP64 as the RK stage
    // and the boundary kernels form it: the same value as a viscosity field
    // written after each stage, without its 8 B/pt store and reload
    const double Tp = double(q[3][4]);
    const S Amu = S(A.c.amu_c0 * sqrt(Tp * Tp * Tp) / (A.c.tsb + Tp));
    S gr[4][3];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        S s[9];
#pragma unroll
        for (int m = 0; m < 9; m++) s[m] = pl[f][(ty + 4) * PW + tx + m];
        gr[f][0] = cd9<SV, S, CL>(s, t.i, g.nx, A.bnd[0], A.bnd[1], rhx);
#pragma unroll
        for (int m = 0; m < 9; m++) s[m] = pl[f][(ty + m) * PW + tx + 4];
        gr[f][1] = cd9<SV, S, CL>(s, t.j, g.ny, A.bnd[2], A.bnd[3], rhy);
        gr[f][2] = cd9<SV, S, CL>(q[f], k, g.nz, A.bnd[4], A.bnd[5], rhz);
    }
    // vis_flux_s_ker / vis_flux_e_ker / vis_flus_ev_ker
    const S uk = gr[0][0], ui = gr[0][1], us = gr[0][2];
    const S vk = gr[1][0], vi = gr[1][1], vs = gr[1][2];
    const S wk = gr[2][0], wi = gr[2][1], ws = gr[2][2];
    const S Tk = gr[3][0], Ti = gr[3][1], Ts = gr[3][2];
    const S ux = uk * Akx + ui * Aix + us * Asx;
    const S vx = vk * Akx + vi * Aix + vs * Asx;
    const S wx = wk * Akx + wi * Aix + ws * Asx;
    const S uy = uk * Aky + ui * Aiy + us * Asy;
    const S vy = vk * Aky + vi * Aiy + vs * Asy;
    const S wy = wk * Aky + wi * Aiy + ws * Asy;
    const S uz = uk * Akz + ui * Aiz + us * Asz;
    const S vz = vk * Akz + vi * Aiz + vs * Asz;
    const S wz = wk * Akz + wi * Aiz + ws * Asz;
    const S div = ux + vy + wz;
    const S two3 = S(2.0 / 3.0);
    const S s11 = (S(2) * ux - two3 * div) * Amu;
    const S s22 = (S(2) * vy - two3 * div) * Amu;
    const S s33 = (S(2) * wz - two3 * div) * Amu;
    const S s12 = (uy + vx) * Amu;
    const S s13 = (uz + wx) * Amu;
    const S s23 = (vz + wy) * Amu;
    const S Amuk = Amu * S(A.c.vis_c);
    const S Tx = Tk * Akx + Ti * Aix + Ts * Asx;
    const S Ty = Tk * Aky + Ti * Aiy + Ts * Asy;
    const S Tz = Tk * Akz + Ti * Aiz + Ts * Asz;
    const S u = q[0][4], v = q[1][4], w = q[2][4];
    const S E1 = u * s11 + v * s12 + w * s13 + Amuk * Tx;
    const S E2 = u * s12 + v * s22 + w * s23 + Amuk * Ty;
    const S E3 = u * s13 + v * s23 + w * s33 + Amuk * Tz;
    const S ax[3] = {Akx, Aix, Asx}, ay[3] = {Aky, Aiy, Asy}, az[3] = {Akz, Aiz, Asz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const S akx = ax[d] * Aj1, aky = ay[d] * Aj1, akz = az[d] * Aj1;
        S* E = A.ev + (long long)(4 * d) * fs + qi;
        __builtin_nontemporal_store(S(akx * s11 + aky * s12 + akz * s13), &E[0]);
        __builtin_nontemporal_store(S(akx * s12 + aky * s22 + akz * s23), &E[fs]);
        __builtin_nontemporal_store(S(akx * s13 + aky * s23 + akz * s33), &E[2 * fs]);
        __builtin_nontemporal_store(S(akx * E1 + aky * E2 + akz * E3), &E[3 * fs]);
    }
    
}

// Whether a central-derivative closure is in reach of the block's plane k:
// the same for every thread of the block.
template <int TYB, typename S>
__device__ __forceinline__ bool near_closure(const ViscArgsT<S>& A, const Tile25& t, int k) {
    const Geo& g = A.g;
    return (A.bnd[0] && t.i0 <= 3) || (A.bnd[1] && t.i0 + TX > g.nx - 4) || (A.bnd[2] && t.j0 <= 3) ||
           (A.bnd[3] && t.j0 + TYB > g.ny - 4) || (A.bnd[4] && k <= 3) || (A.bnd[5] && k >= g.nz - 4);
}

// The plane loads of both kernels go through per-thread slots: the slot's
// LDS position and its global offset in plane 0 (-1: outside the ghost
// layers, stored as 0).  Loads for plane k+1 are issued before plane k is
// computed.
template <int SV, typename S>
__global__ void __launch_bounds__(TX * TY_<S>::v) k_visc_gf(ViscArgsT<S> A) {
    constexpr int TY = TY_<S>::v, NT = TX * TY;
    constexpr int PW = TX + 8, PN = PW * (TY + 8), NSL = (PN + NT - 1) / NT;
    const Geo& g = A.g;
    const Tile25 t = tile_of<S, false>(g);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = t.i0, j0 = t.j0;
    const S rhx = S(1.0 / g.hx), rhy = S(1.0 / g.hy), rhz = S(1.0 / g.hz);
    __shared__ S pl[4][PN];
    const S* src[4] = {A.u, A.v, A.w, A.T};
    const int ic = min(t.i, g.nx - 1), jc = min(t.j, g.ny - 1);  // clamped column for queue loads
    const long long col = gidx(g, ic, jc, 0);
    // halo slots: positions outside the TX x TY interior of the tile that a
    // derivative reads (the x strips beside the interior rows and the y
    // strips beside the interior columns; the corners are never read)
    int sl[NSL];
    long long so[NSL];
#pragma unroll
    for (int r = 0; r < NSL; r++) {
        const int p = tid + NT * r;
        const int hx = p % PW, hy = p / PW;
        const bool xin = hx >= 4 && hx < TX + 4, yin = hy >= 4 && hy < TY + 4;
        sl[r] = (p < PN && (xin != yin || !OCFD_VISC_SKIP_CORNERS) && !(xin && yin)) ? p : -1;
        const int gi = i0 - 4 + hx, gj = j0 - 4 + hy;
        so[r] = (gi <= g.nx + 3 && gj <= g.ny + 3) ? gidx(g, gi, gj, 0) : -1;
    }
    // the thread's own tile position: from the queue when owned, else from
    // memory (x/y ghosts inside the tile) or 0
    const long long co = t.own ? -1 : (t.i <= g.nx + 3 && t.j <= g.ny + 3) ? gidx(g, t.i, t.j, 0) : -1;
    S hv[NSL][4], cv[4];
    auto load_plane = [&](int k) {
        const long long pk = (long long)k * g.sz;
#pragma unroll
        for (int r = 0; r < NSL; r++)
#pragma unroll
            for (int f = 0; f < 4; f++) hv[r][f] = (sl[r] >= 0 && so[r] >= 0) ? src[f][so[r] + pk] : S(0);
#pragma unroll
        for (int f = 0; f < 4; f++) cv[f] = co >= 0 ? src[f][co + pk] : S(0);
    };
    // z queue of the own column: q[f][m] = field f at plane k-4+m
    S q[4][9], nq[4];
#pragma unroll
    for (int m = 0; m < 8; m++)
#pragma unroll
        for (int f = 0; f < 4; f++) q[f][m] = src[f][col + (long long)(t.kbeg - 4 + m) * g.sz];
#pragma unroll
    for (int f = 0; f < 4; f++) nq[f] = src[f][col + (long long)(t.kbeg + 4) * g.sz];
    if constexpr (OCFD_VISC_RING && std::is_same<S, double>::value) {
    // The halo strips of plane k+5 are fetched in step k, when the
    // neighbouring tiles fetch the same rows as their own z-queue plane k+5
    // (so one of the two reads hits L2), and wait in an LDS ring of RING
    // planes until step k+5.  Each thread keeps its own halo positions (a
    // dense numbering h of the strips: x strips first, then y strips), so
    // the ring needs no barrier of its own.
    constexpr int NHALO = 8 * TY + 8 * TX, HPT = (NHALO + NT - 1) / NT, RING = 5;
    __shared__ S ring[RING][4][NHALO];
    int hpos[HPT];
    long long hoff[HPT];
#pragma unroll
    for (int r = 0; r < HPT; r++) {
        const int h = tid + NT * r;
        int hx = 0, hy = 0;
        if (h < 8 * TY) {
            const int c = h % 8;
            hx = c < 4 ? c : TX + c;
            hy = 4 + h / 8;
        } else {
            const int h2 = h - 8 * TY, rw = h2 / TX;
            hx = 4 + h2 % TX;
            hy = rw < 4 ? rw : TY + rw;
        }
        const int gi = i0 - 4 + hx, gj = j0 - 4 + hy;
        hpos[r] = h < NHALO ? hy * PW + hx : -1;
        hoff[r] = (h < NHALO && gi <= g.nx + 3 && gj <= g.ny + 3) ? gidx(g, gi, gj, 0) : -1;
    }
    S hn[HPT][4];
    auto load_strips = [&](int k) {
        const long long pk = (long long)k * g.sz;
#pragma unroll
        for (int r = 0; r < HPT; r++)
#pragma unroll
            for (int f = 0; f < 4; f++) hn[r][f] = (hoff[r] >= 0 && k < t.kend) ? src[f][hoff[r] + pk] : S(0);
    };
    auto store_strips = [&](int k) {
        const int sl = (k - t.kbeg) % RING;
#pragma unroll
        for (int r = 0; r < HPT; r++)
            if (hpos[r] >= 0)
#pragma unroll
                for (int f = 0; f < 4; f++) ring[sl][f][tid + NT * r] = hn[r][f];
    };
#pragma unroll
    for (int p = 0; p < RING; p++) {
        load_strips(t.kbeg + p);
        store_strips(t.kbeg + p);
    }
#pragma unroll
    for (int f = 0; f < 4; f++) cv[f] = co >= 0 ? src[f][co + (long long)t.kbeg * g.sz] : S(0);
    for (int k = t.kbeg; k < t.kend; k++) {
        const int sl = (k - t.kbeg) % RING;
#pragma unroll
        for (int f = 0; f < 4; f++) {
            q[f][8] = nq[f];
            pl[f][(ty + 4) * PW + tx + 4] = t.own ? q[f][4] : cv[f];
#pragma unroll
            for (int r = 0; r < HPT; r++)
                if (hpos[r] >= 0) pl[f][hpos[r]] = ring[sl][f][tid + NT * r];
        }
        __syncthreads();
        if (k + 1 < t.kend) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                nq[f] = src[f][col + (long long)(k + 5) * g.sz];
                cv[f] = co >= 0 ? src[f][co + (long long)(k + 1) * g.sz] : S(0);
            }
            load_strips(k + 5);
        }
        if (t.own) {
            if (near_closure<TY>(A, t, k)) gf_point<SV, S, PW, true>(A, t, k, pl, q, tx, ty, rhx, rhy, rhz);
            else gf_point<SV, S, PW, false>(A, t, k, pl, q, tx, ty, rhx, rhy, rhz);
        }
        // plane k+5 into the slot plane k has just left
        if (k + 1 < t.kend) store_strips(k + 5);
#pragma unroll
        for (int f = 0; f < 4; f++)
#pragma unroll
            for (int m = 0; m < 8; m++) q[f][m] = q[f][m + 1];
        __syncthreads();
    }
    (void)sl;
    (void)so;
    (void)hv;
    (void)load_plane;
    } else {
    load_plane(t.kbeg);
    for (int k = t.kbeg; k < t.kend; k++) {
        if (!VISC_PREFETCH && k > t.kbeg) {
#pragma unroll
            for (int f = 0; f < 4; f++) nq[f] = src[f][col + (long long)(k + 4) * g.sz];
            load_plane(k);
        }
#pragma unroll
        for (int f = 0; f < 4; f++) {
            q[f][8] = nq[f];
            pl[f][(ty + 4) * PW + tx + 4] = t.own ? q[f][4] : cv[f];
#pragma unroll
            for (int r = 0; r < NSL; r++)
                if (sl[r] >= 0) pl[f][sl[r]] = hv[r][f];
        }
        __syncthreads();
        if (VISC_PREFETCH && k + 1 < t.kend) {
#pragma unroll
            for (int f = 0; f < 4; f++) nq[f] = src[f][col + (long long)(k + 5) * g.sz];
            load_plane(k + 1);
        }
        if (t.own) {
            if (near_closure<TY>(A, t, k)) gf_point<SV, S, PW, true>(A, t, k, pl, q, tx, ty, rhx, rhy, rhz);
            else gf_point<SV, S, PW, false>(A, t, k, pl, q, tx, ty, rhx, rhy, rhz);
        }
#pragma unroll
        for (int f = 0; f < 4; f++)
#pragma unroll
            for (int m = 0; m < 8; m++) q[f][m] = q[f][m + 1];
        __syncthreads();
    }
    }
}

// Ev_x (4 fields) needs x neighbours, Ev_y y neighbours, Ev_z z neighbours.
// One launch handles NC of the four flux components, starting at c0 (each
// component is independent; fewer components per launch need fewer
// registers).
constexpr int NTD = TX * TYD;
template <int SV, typename S, int NC, bool SEP>
__global__ void __launch_bounds__(NTD) k_visc_div(ViscArgsT<S> A, int c0) {
    constexpr int XW = TX + 8, XN = TYD * XW, YN = (TYD + 8) * TX;
    constexpr int NSX = (XN + NTD - 1) / NTD, NSY = (YN + NTD - 1) / NTD;
    const Geo& g = A.g;
    const Tile25 t = tile_of<S, true, TYD>(g);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = t.i0, j0 = t.j0;
    const S rhx = S(1.0 / g.hx), rhy = S(1.0 / g.hy), rhz = S(1.0 / g.hz);
    __shared__ S ex[NC][XN];
    __shared__ S ey[NC][YN];
    const long long fs = A.fs;
    const S* EX = A.ev + c0 * fs;
    const S* EY = A.ev + (4 + c0) * fs;
    const S* EZ = A.ev + (8 + c0) * fs;
    const int ic = min(t.i, g.nx - 1), jc = min(t.j, g.ny - 1);
    const long long col = gidx(g, ic, jc, 0);
    int slx[NSX], sly[NSY];
    long long sox[NSX], soy[NSY];
#pragma unroll
    for (int r = 0; r < NSX; r++) {
        const int p = tid + NTD * r;
        const int hx = p % XW, hy = p / XW;
        const int gi = i0 - 4 + hx, gj = j0 + hy;
        slx[r] = p < XN ? p : -1;
        sox[r] = (gi <= g.nx + 3 && gj < g.ny) ? gidx(g, gi, gj, 0) : -1;
    }
#pragma unroll
    for (int r = 0; r < NSY; r++) {
        const int p = tid + NTD * r;
        const int hx = p % TX, hy = p / TX;
        const int gi = i0 + hx, gj = j0 - 4 + hy;
        sly[r] = p < YN ? p : -1;
        soy[r] = (gi < g.nx && gj <= g.ny + 3) ? gidx(g, gi, gj, 0) : -1;
    }
    S xv[NSX][NC], yv[NSY][NC];
    auto load_plane = [&](int k) {
        const long long pk = (long long)k * g.sz;
#pragma unroll
        for (int r = 0; r < NSX; r++)
#pragma unroll
            for (int f = 0; f < NC; f++) xv[r][f] = (slx[r] >= 0 && sox[r] >= 0) ? EX[f * fs + sox[r] + pk] : S(0);
#pragma unroll
        for (int r = 0; r < NSY; r++)
#pragma unroll
            for (int f = 0; f < NC; f++) yv[r][f] = (sly[r] >= 0 && soy[r] >= 0) ? EY[f * fs + soy[r] + pk] : S(0);
    };
    S q[NC][9], nq[NC];
#pragma unroll
    for (int m = 0; m < 8; m++)
#pragma unroll
        for (int f = 0; f < NC; f++) q[f][m] = EZ[f * fs + col + (long long)(t.kbeg - 4 + m) * g.sz];
#pragma unroll
    for (int f = 0; f < NC; f++) nq[f] = EZ[f * fs + col + (long long)(t.kbeg + 4) * g.sz];
    load_plane(t.kbeg);
    for (int k = t.kbeg; k < t.kend; k++) {
        if (!VISC_PREFETCH && k > t.kbeg) {
#pragma unroll
            for (int f = 0; f < NC; f++) nq[f] = EZ[f * fs + col + (long long)(k + 4) * g.sz];
            load_plane(k);
        }
#pragma unroll
        for (int f = 0; f < NC; f++) {
            q[f][8] = nq[f];
#pragma unroll
            for (int r = 0; r < NSX; r++)
                if (slx[r] >= 0) ex[f][slx[r]] = xv[r][f];
#pragma unroll
            for (int r = 0; r < NSY; r++)
                if (sly[r] >= 0) ey[f][sly[r]] = yv[r][f];
        }
        __syncthreads();
        if (VISC_PREFETCH && k + 1 < t.kend) {
#pragma unroll
            for (int f = 0; f < NC; f++) nq[f] = EZ[f * fs + col + (long long)(k + 5) * g.sz];
            load_plane(k + 1);
        }
        // CL = false: no closure in reach of the block's plane (straight-line code)
        auto point = [&](auto cl) {
            constexpr bool CL = decltype(cl)::value;
            const long long qi = gidx(g, t.i, t.j, k);
#pragma unroll
            for (int m = 0; m < NC; m++) {
                S s[9];
#pragma unroll
                for (int r = 0; r < 9; r++) s[r] = ex[m][ty * XW + tx + r];
                const S vx = cd9<SV, S, CL>(s, t.i, g.nx, A.bnd[0], A.bnd[1], rhx);
#pragma unroll
                for (int r = 0; r < 9; r++) s[r] = ey[m][(ty + r) * TX + tx];
                const S vy = cd9<SV, S, CL>(s, t.j, g.ny, A.bnd[2], A.bnd[3], rhy);
                const S vz = cd9<SV, S, CL>(q[m], k, g.nz, A.bnd[4], A.bnd[5], rhz);
                // the sums are double; the node's Ajac, by which the
                // reference multiplies each of these terms, is applied to the
                // whole right-hand side by the RK stage
                if (SEP) {
                    double acc = double(vx);
                    acc = acc + double(vy);
                    acc = acc + double(vz);
                    A.dv[(long long)(c0 + m) * fs + qi] = S(acc);
                } else {
                    S* du = A.du + (long long)(c0 + m + 1) * fs + qi;
                    double acc = double(*du);
                    acc = acc + double(vx);
                    acc = acc + double(vy);
                    acc = acc + double(vz);
                    *du = S(acc);
                }
            }
        };
        if (t.own) {
            if (near_closure<TYD>(A, t, k)) point(std::true_type{});
            else point(std::false_type{});
        }
#pragma unroll
        for (int f = 0; f < NC; f++)
#pragma unroll
            for (int m = 0; m < 8; m++) q[f][m] = q[f][m + 1];
        __syncthreads();
    }
}

// Q criterion of the velocity (get_Q / get_Q_kernal, src/OCFD_ana.cu:
// 299-420): the nine derivatives of u, v, w along xi/eta/zeta with the
// viscous scheme and its D0 closures (OCFD_dx0/dy0/dz0), the Cartesian
// gradient from the cofactors, Q = (ux vy + ux wz + vy wz - uy vx - uz wx
// - vz wy) Ajac^2.  A diagnostic run once: one thread per point, stencils read
// straight from memory, FP64 arithmetic on both paths.
template <int SV, typename S>
__global__ void __launch_bounds__(256) k_qcrit(ViscArgsT<S> A, double* __restrict__ qout) {
    const Geo& g = A.g;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const long long qi = gidx(g, i, j, k);
    const double rhx = 1.0 / g.hx, rhy = 1.0 / g.hy, rhz = 1.0 / g.hz;
    const S* src[3] = {A.u, A.v, A.w};
    double gr[3][3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        double s[9];
#pragma unroll
        for (int m = 0; m < 9; m++) s[m] = double(src[f][qi + (m - 4)]);
        gr[f][0] = cd9<SV>(s, i, g.nx, A.bnd[0], A.bnd[1], rhx);
#pragma unroll
        for (int m = 0; m < 9; m++) s[m] = double(src[f][qi + (long long)(m - 4) * g.ldx]);
        gr[f][1] = cd9<SV>(s, j, g.ny, A.bnd[2], A.bnd[3], rhy);
#pragma unroll
        for (int m = 0; m < 9; m++) s[m] = double(src[f][qi + (long long)(m - 4) * g.sz]);
        gr[f][2] = cd9<SV>(s, k, g.nz, A.bnd[4], A.bnd[5], rhz);
    }
    const double Akx = A.M[M_KX][qi], Aky = A.M[M_KY][qi], Akz = A.M[M_KZ][qi];
    const double Aix = A.M[M_IX][qi], Aiy = A.M[M_IY][qi], Aiz = A.M[M_IZ][qi];
    const double Asx = A.M[M_SX][qi], Asy = A.M[M_SY][qi], Asz = A.M[M_SZ][qi];
    const double ux = gr[0][0] * Akx + gr[0][1] * Aix + gr[0][2] * Asx;
    const double vx = gr[1][0] * Akx + gr[1][1] * Aix + gr[1][2] * Asx;
    const double wx = gr[2][0] * Akx + gr[2][1] * Aix + gr[2][2] * Asx;
    const double uy = gr[0][0] * Aky + gr[0][1] * Aiy + gr[0][2] * Asy;
    const double vy = gr[1][0] * Aky + gr[1][1] * Aiy + gr[1][2] * Asy;
    const double wy = gr[2][0] * Aky + gr[2][1] * Aiy + gr[2][2] * Asy;
    const double uz = gr[0][0] * Akz + gr[0][1] * Aiz + gr[0][2] * Asz;
    const double vz = gr[1][0] * Akz + gr[1][1] * Aiz + gr[1][2] * Asz;
    const double wz = gr[2][0] * Akz + gr[2][1] * Aiz + gr[2][2] * Asz;
    const double aj = A.M[M_JAC][qi];
    qout[qi] = (ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy) * aj * aj;
}

template <typename S, bool DIV>
dim3 grid25(const Geo& g) {
    constexpr int KZ = KZ_<S, DIV>::v, TYB = DIV ? TYD : TY_<S>::v;
    return dim3((g.nx + TX - 1) / TX, (g.ny + TYB - 1) / TYB, (g.nz + KZ - 1) / KZ);
}

}  // namespace

template <typename S>
void launch_visc_gf(const ViscArgsT<S>& a, hipStream_t st) {
    if (a.svis == OCFD_CD6) hipLaunchKernelGGL((k_visc_gf<OCFD_CD6, S>), (grid25<S, false>(a.g)), dim3(TX, TY_<S>::v), 0, st, a);
    else hipLaunchKernelGGL((k_visc_gf<OCFD_CD8, S>), (grid25<S, false>(a.g)), dim3(TX, TY_<S>::v), 0, st, a);
}
// Flux components per divergence launch: two (measured on the 1024x384x256
// swept ramp: FP64 3.00-3.30 ms per stage for two launches of two against
// 3.34-3.59 ms for four of one and 3.27-3.44 ms for one of four; the mixed
// step 43.5 -> 43.0 ms with two launches of two instead of one of four; on
// the 512x128x64 ramp four of one had been the fastest).
#ifndef OCFD_VDIV_NC_D
#define OCFD_VDIV_NC_D 2  // flux components per divergence launch, FP64 path
#endif
template <typename S> struct NC_ { static constexpr int v = OCFD_VDIV_NC_D; };
#ifndef OCFD_VDIV_NC_F
#define OCFD_VDIV_NC_F 2  // flux components per divergence launch, FP32 path
#endif
template <> struct NC_<float> { static constexpr int v = OCFD_VDIV_NC_F; };
template <typename S>
int visc_div_components() { return NC_<S>::v; }
template <typename S>
void launch_visc_div(const ViscArgsT<S>& a, int c0, hipStream_t st) {
    constexpr int NC = NC_<S>::v;
    // SEP: the terms go to dv (set) or are added to du (null)
    const dim3 grid = grid25<S, true>(a.g), blk(TX, TYD);
    if (a.svis == OCFD_CD6) {
        if (a.dv) hipLaunchKernelGGL((k_visc_div<OCFD_CD6, S, NC, true>), grid, blk, 0, st, a, c0);
        else hipLaunchKernelGGL((k_visc_div<OCFD_CD6, S, NC, false>), grid, blk, 0, st, a, c0);
    } else {
        if (a.dv) hipLaunchKernelGGL((k_visc_div<OCFD_CD8, S, NC, true>), grid, blk, 0, st, a, c0);
        else hipLaunchKernelGGL((k_visc_div<OCFD_CD8, S, NC, false>), grid, blk, 0, st, a, c0);
    }
}
template <typename S>
void launch_qcrit(const ViscArgsT<S>& a, double* q, hipStream_t st) {
    const dim3 grid((a.g.nx + 63) / 64, (a.g.ny + 3) / 4, a.g.nz);
    if (a.svis == OCFD_CD6) hipLaunchKernelGGL((k_qcrit<OCFD_CD6, S>), grid, dim3(64, 4), 0, st, a, q);
    else hipLaunchKernelGGL((k_qcrit<OCFD_CD8, S>), grid, dim3(64, 4), 0, st, a, q);
}
template void launch_qcrit<double>(const ViscArgsT<double>&, double*, hipStream_t);
template void launch_qcrit<float>(const ViscArgsT<float>&, double*, hipStream_t);
template void launch_visc_gf<double>(const ViscArgsT<double>&, hipStream_t);
template void launch_visc_gf<float>(const ViscArgsT<float>&, hipStream_t);
template void launch_visc_div<double>(const ViscArgsT<double>&, int, hipStream_t);
template void launch_visc_div<float>(const ViscArgsT<float>&, int, hipStream_t);
template int visc_div_components<double>();
template int visc_div_components<float>();
