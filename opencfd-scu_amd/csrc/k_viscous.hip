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
//                src/commen_kernel.cu:158-169), no atomics.
//
// Both use 2.5-D blocking: a block owns a 32x8 tile of (x,y) columns and
// marches through a chunk of z planes.  x and y neighbours come from the
// current plane staged in LDS with its 4-point halo; z neighbours come from a
// 9-plane queue each thread keeps in registers for its own column.
#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

#ifndef OCFD_VISC_KZ
#define OCFD_VISC_KZ 16  // z planes per block
#endif
constexpr int TX = 32, TY = 8, KZ = OCFD_VISC_KZ;

// Central derivative at interior index i (0..n-1) from s[0..8] = values at
// i-4..i+4 (OCFD_CD6/CD8_kernel + D0 closures + put_d0_kernel).  rh = 1/h; the
// reference's divisions by 12, 60, 840 and h are multiplications by the
// reciprocals here (a difference of an ulp or two per derivative).
template <int SV>
__device__ __forceinline__ double cd9(const double* s, int i, int n, int bl, int br, double rh) {
    const double* c = s + 4;
    double tmp;
    if (bl && i <= 3) {
        if (i == 0) tmp = (c[1] - c[0]);
        else if (i == 1) tmp = (c[1] - c[-1]) * 0.5;
        else if (i == 2) tmp = (c[-2] - 8.0 * c[-1] + 8.0 * c[1] - c[2]) * (1.0 / 12.0);
        else tmp = (c[3] - c[-3] - 9.0 * (c[2] - c[-2]) + 45.0 * (c[1] - c[-1])) * (1.0 / 60.0);
    } else if (br && i >= n - 4) {
        const int t1 = i - n;
        if (t1 == -1) tmp = (c[0] - c[-1]);
        else if (t1 == -2) tmp = (c[1] - c[-1]) * 0.5;
        else if (t1 == -3) tmp = (c[-2] - 8.0 * c[-1] + 8.0 * c[1] - c[2]) * (1.0 / 12.0);
        else tmp = (c[3] - c[-3] - 9.0 * (c[2] - c[-2]) + 45.0 * (c[1] - c[-1])) * (1.0 / 60.0);
    } else if (SV == OCFD_CD6) {
        tmp = (45.0 * (c[1] - c[-1]) - 9.0 * (c[2] - c[-2]) + (c[3] - c[-3])) * (1.0 / 60.0);
    } else {
        tmp = (672.0 * (c[1] - c[-1]) - 168.0 * (c[2] - c[-2]) + 32.0 * (c[3] - c[-3]) - 3.0 * (c[4] - c[-4])) *
              (1.0 / 840.0);
    }
    return tmp * rh;
}

// Tile bookkeeping shared by both kernels.
struct Tile25 {
    int i, j, kbeg, kend;
    bool own;
};
__device__ __forceinline__ Tile25 tile_of(const Geo& g) {
    Tile25 t;
    t.i = blockIdx.x * TX + threadIdx.x;
    t.j = blockIdx.y * TY + threadIdx.y;
    t.kbeg = blockIdx.z * KZ;
    t.kend = min(t.kbeg + KZ, g.nz);
    t.own = t.i < g.nx && t.j < g.ny;
    return t;
}

// The plane loads of both kernels go through per-thread slots: the slot's
// LDS position and its global offset in plane 0 (-1: outside the ghost
// layers, stored as 0).  Loads for plane k+1 are issued before plane k is
// computed.
constexpr int NT = TX * TY;

template <int SV>
__global__ void __launch_bounds__(NT) k_visc_gf(ViscArgs A) {
    constexpr int PW = TX + 8, PN = PW * (TY + 8), NSL = (PN + NT - 1) / NT;
    const Geo& g = A.g;
    const Tile25 t = tile_of(g);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const double rhx = 1.0 / g.hx, rhy = 1.0 / g.hy, rhz = 1.0 / g.hz;
    __shared__ double pl[4][PN];
    const double* src[4] = {A.u, A.v, A.w, A.T};
    const int ic = min(t.i, g.nx - 1), jc = min(t.j, g.ny - 1);  // clamped column for queue loads
    const long long col = gidx(g, ic, jc, 0);
    // halo slots (positions outside the 32x8 interior of the tile)
    int sl[NSL];
    long long so[NSL];
#pragma unroll
    for (int r = 0; r < NSL; r++) {
        const int p = tid + NT * r;
        const int hx = p % PW, hy = p / PW;
        const bool inner = hx >= 4 && hx < TX + 4 && hy >= 4 && hy < TY + 4;
        sl[r] = (p < PN && !inner) ? p : -1;
        const int gi = i0 - 4 + hx, gj = j0 - 4 + hy;
        so[r] = (gi <= g.nx + 3 && gj <= g.ny + 3) ? gidx(g, gi, gj, 0) : -1;
    }
    // the thread's own tile position: from the queue when owned, else from
    // memory (x/y ghosts inside the tile) or 0
    const long long co = t.own ? -1 : (t.i <= g.nx + 3 && t.j <= g.ny + 3) ? gidx(g, t.i, t.j, 0) : -1;
    double hv[NSL][4], cv[4];
    auto load_plane = [&](int k) {
        const long long pk = (long long)k * g.sz;
#pragma unroll
        for (int r = 0; r < NSL; r++)
#pragma unroll
            for (int f = 0; f < 4; f++) hv[r][f] = (sl[r] >= 0 && so[r] >= 0) ? src[f][so[r] + pk] : 0.0;
#pragma unroll
        for (int f = 0; f < 4; f++) cv[f] = co >= 0 ? src[f][co + pk] : 0.0;
    };
    // z queue of the own column: q[f][m] = field f at plane k-4+m
    double q[4][9], nq[4];
#pragma unroll
    for (int m = 0; m < 8; m++)
#pragma unroll
        for (int f = 0; f < 4; f++) q[f][m] = src[f][col + (long long)(t.kbeg - 4 + m) * g.sz];
#pragma unroll
    for (int f = 0; f < 4; f++) nq[f] = src[f][col + (long long)(t.kbeg + 4) * g.sz];
    load_plane(t.kbeg);
    for (int k = t.kbeg; k < t.kend; k++) {
#pragma unroll
        for (int f = 0; f < 4; f++) {
            q[f][8] = nq[f];
            pl[f][(ty + 4) * PW + tx + 4] = t.own ? q[f][4] : cv[f];
#pragma unroll
            for (int r = 0; r < NSL; r++)
                if (sl[r] >= 0) pl[f][sl[r]] = hv[r][f];
        }
        __syncthreads();
        if (k + 1 < t.kend) {
#pragma unroll
            for (int f = 0; f < 4; f++) nq[f] = src[f][col + (long long)(k + 5) * g.sz];
            load_plane(k + 1);
        }
        if (t.own) {
            double gr[4][3];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                double s[9];
#pragma unroll
                for (int m = 0; m < 9; m++) s[m] = pl[f][(ty + 4) * PW + tx + m];
                gr[f][0] = cd9<SV>(s, t.i, g.nx, A.bnd[0], A.bnd[1], rhx);
#pragma unroll
                for (int m = 0; m < 9; m++) s[m] = pl[f][(ty + m) * PW + tx + 4];
                gr[f][1] = cd9<SV>(s, t.j, g.ny, A.bnd[2], A.bnd[3], rhy);
                gr[f][2] = cd9<SV>(q[f], k, g.nz, A.bnd[4], A.bnd[5], rhz);
            }
            // vis_flux_s_ker / vis_flux_e_ker / vis_flus_ev_ker
            const long long qi = gidx(g, t.i, t.j, k), fs = A.fs;
            const double Akx = A.M[M_KX][qi], Aky = A.M[M_KY][qi], Akz = A.M[M_KZ][qi];
            const double Aix = A.M[M_IX][qi], Aiy = A.M[M_IY][qi], Aiz = A.M[M_IZ][qi];
            const double Asx = A.M[M_SX][qi], Asy = A.M[M_SY][qi], Asz = A.M[M_SZ][qi];
            const double Amu = A.amu[qi];
            const double uk = gr[0][0], ui = gr[0][1], us = gr[0][2];
            const double vk = gr[1][0], vi = gr[1][1], vs = gr[1][2];
            const double wk = gr[2][0], wi = gr[2][1], ws = gr[2][2];
            const double Tk = gr[3][0], Ti = gr[3][1], Ts = gr[3][2];
            const double ux = uk * Akx + ui * Aix + us * Asx;
            const double vx = vk * Akx + vi * Aix + vs * Asx;
            const double wx = wk * Akx + wi * Aix + ws * Asx;
            const double uy = uk * Aky + ui * Aiy + us * Asy;
            const double vy = vk * Aky + vi * Aiy + vs * Asy;
            const double wy = wk * Aky + wi * Aiy + ws * Asy;
            const double uz = uk * Akz + ui * Aiz + us * Asz;
            const double vz = vk * Akz + vi * Aiz + vs * Asz;
            const double wz = wk * Akz + wi * Aiz + ws * Asz;
            const double div = ux + vy + wz;
            const double two3 = 2.0 / 3.0;
            const double s11 = (2.0 * ux - two3 * div) * Amu;
            const double s22 = (2.0 * vy - two3 * div) * Amu;
            const double s33 = (2.0 * wz - two3 * div) * Amu;
            const double s12 = (uy + vx) * Amu;
            const double s13 = (uz + wx) * Amu;
            const double s23 = (vz + wy) * Amu;
            const double Amuk = Amu * A.c.vis_c;
            const double Tx = Tk * Akx + Ti * Aix + Ts * Asx;
            const double Ty = Tk * Aky + Ti * Aiy + Ts * Asy;
            const double Tz = Tk * Akz + Ti * Aiz + Ts * Asz;
            const double u = q[0][4], v = q[1][4], w = q[2][4];
            const double E1 = u * s11 + v * s12 + w * s13 + Amuk * Tx;
            const double E2 = u * s12 + v * s22 + w * s23 + Amuk * Ty;
            const double E3 = u * s13 + v * s23 + w * s33 + Amuk * Tz;
            const double Aj1 = A.M[M_JAC][qi];
            const double ax[3] = {Akx, Aix, Asx}, ay[3] = {Aky, Aiy, Asy}, az[3] = {Akz, Aiz, Asz};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double akx = ax[d] * Aj1, aky = ay[d] * Aj1, akz = az[d] * Aj1;
                double* E = A.ev + (long long)(4 * d) * fs + qi;
                E[0] = (akx * s11 + aky * s12 + akz * s13);
                E[fs] = (akx * s12 + aky * s22 + akz * s23);
                E[2 * fs] = (akx * s13 + aky * s23 + akz * s33);
                E[3 * fs] = (akx * E1 + aky * E2 + akz * E3);
            }
        }
#pragma unroll
        for (int f = 0; f < 4; f++)
#pragma unroll
            for (int m = 0; m < 8; m++) q[f][m] = q[f][m + 1];
        __syncthreads();
    }
}

// Ev_x (4 fields) needs x neighbours, Ev_y y neighbours, Ev_z z neighbours.
template <int SV>
__global__ void __launch_bounds__(NT) k_visc_div(ViscArgs A) {
    constexpr int XW = TX + 8, XN = TY * XW, YN = (TY + 8) * TX;
    constexpr int NSX = (XN + NT - 1) / NT, NSY = (YN + NT - 1) / NT;
    const Geo& g = A.g;
    const Tile25 t = tile_of(g);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const double rhx = 1.0 / g.hx, rhy = 1.0 / g.hy, rhz = 1.0 / g.hz;
    __shared__ double ex[4][XN];
    __shared__ double ey[4][YN];
    const long long fs = A.fs;
    const double* EX = A.ev;
    const double* EY = A.ev + 4 * fs;
    const double* EZ = A.ev + 8 * fs;
    const int ic = min(t.i, g.nx - 1), jc = min(t.j, g.ny - 1);
    const long long col = gidx(g, ic, jc, 0);
    int slx[NSX], sly[NSY];
    long long sox[NSX], soy[NSY];
#pragma unroll
    for (int r = 0; r < NSX; r++) {
        const int p = tid + NT * r;
        const int hx = p % XW, hy = p / XW;
        const int gi = i0 - 4 + hx, gj = j0 + hy;
        slx[r] = p < XN ? p : -1;
        sox[r] = (gi <= g.nx + 3 && gj < g.ny) ? gidx(g, gi, gj, 0) : -1;
    }
#pragma unroll
    for (int r = 0; r < NSY; r++) {
        const int p = tid + NT * r;
        const int hx = p % TX, hy = p / TX;
        const int gi = i0 + hx, gj = j0 - 4 + hy;
        sly[r] = p < YN ? p : -1;
        soy[r] = (gi < g.nx && gj <= g.ny + 3) ? gidx(g, gi, gj, 0) : -1;
    }
    double xv[NSX][4], yv[NSY][4];
    auto load_plane = [&](int k) {
        const long long pk = (long long)k * g.sz;
#pragma unroll
        for (int r = 0; r < NSX; r++)
#pragma unroll
            for (int f = 0; f < 4; f++) xv[r][f] = (slx[r] >= 0 && sox[r] >= 0) ? EX[f * fs + sox[r] + pk] : 0.0;
#pragma unroll
        for (int r = 0; r < NSY; r++)
#pragma unroll
            for (int f = 0; f < 4; f++) yv[r][f] = (sly[r] >= 0 && soy[r] >= 0) ? EY[f * fs + soy[r] + pk] : 0.0;
    };
    double q[4][9], nq[4];
#pragma unroll
    for (int m = 0; m < 8; m++)
#pragma unroll
        for (int f = 0; f < 4; f++) q[f][m] = EZ[f * fs + col + (long long)(t.kbeg - 4 + m) * g.sz];
#pragma unroll
    for (int f = 0; f < 4; f++) nq[f] = EZ[f * fs + col + (long long)(t.kbeg + 4) * g.sz];
    load_plane(t.kbeg);
    for (int k = t.kbeg; k < t.kend; k++) {
#pragma unroll
        for (int f = 0; f < 4; f++) {
            q[f][8] = nq[f];
#pragma unroll
            for (int r = 0; r < NSX; r++)
                if (slx[r] >= 0) ex[f][slx[r]] = xv[r][f];
#pragma unroll
            for (int r = 0; r < NSY; r++)
                if (sly[r] >= 0) ey[f][sly[r]] = yv[r][f];
        }
        __syncthreads();
        if (k + 1 < t.kend) {
#pragma unroll
            for (int f = 0; f < 4; f++) nq[f] = EZ[f * fs + col + (long long)(k + 5) * g.sz];
            load_plane(k + 1);
        }
        if (t.own) {
            const long long qi = gidx(g, t.i, t.j, k);
            const double ajac = A.M[M_JAC][qi];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                double s[9];
#pragma unroll
                for (int r = 0; r < 9; r++) s[r] = ex[m][ty * XW + tx + r];
                const double vx = cd9<SV>(s, t.i, g.nx, A.bnd[0], A.bnd[1], rhx);
#pragma unroll
                for (int r = 0; r < 9; r++) s[r] = ey[m][(ty + r) * TX + tx];
                const double vy = cd9<SV>(s, t.j, g.ny, A.bnd[2], A.bnd[3], rhy);
                const double vz = cd9<SV>(q[m], k, g.nz, A.bnd[4], A.bnd[5], rhz);
                double* dv = A.du + (long long)(m + 1) * fs + qi;
                double acc = *dv;
                acc = acc + ajac * vx;
                acc = acc + ajac * vy;
                acc = acc + ajac * vz;
                *dv = acc;
            }
        }
#pragma unroll
        for (int f = 0; f < 4; f++)
#pragma unroll
            for (int m = 0; m < 8; m++) q[f][m] = q[f][m + 1];
        __syncthreads();
    }
}

dim3 grid25(const Geo& g) { return dim3((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, (g.nz + KZ - 1) / KZ); }

}  // namespace

void launch_visc_gf(const ViscArgs& a, hipStream_t st) {
    if (a.svis == OCFD_CD6) hipLaunchKernelGGL(k_visc_gf<OCFD_CD6>, grid25(a.g), dim3(TX, TY), 0, st, a);
    else hipLaunchKernelGGL(k_visc_gf<OCFD_CD8>, grid25(a.g), dim3(TX, TY), 0, st, a);
}
void launch_visc_div(const ViscArgs& a, hipStream_t st) {
    if (a.svis == OCFD_CD6) hipLaunchKernelGGL(k_visc_div<OCFD_CD6>, grid25(a.g), dim3(TX, TY), 0, st, a);
    else hipLaunchKernelGGL(k_visc_div<OCFD_CD8>, grid25(a.g), dim3(TX, TY), 0, st, a);
}
