// k_viscous.hip - viscous terms (gfx950).
//
// Restates du_vis_Jacobian3d (src/OCFD_Stream.cu:307-335):
//   grad : 12 derivatives of u,v,w,T along xi/eta/zeta with CD6/CD8 and the
//          D0 closures (du_viscous_Jacobian3d_init, src/OCFD_NS_Jacobian3d.cu:
//          58-77; OCFD_CD6/CD8_kernel, src/OCFD_Schemes.cu:96-156;
//          OCFD_D0bound_scheme_kernel, src/OCFD_bound_Scheme.cu:79-186);
//   flux : the contravariant viscous fluxes Ev1..4 of all three directions
//          in one pass (vis_flux_ker, src/OCFD_NS_Jacobian3d.cu:80-281);
//   div  : du[1..4] += Ajac*d(Ev)/dxi, then eta, then zeta (OCFD_d?0 of Ev +
//          YF_Pe_XF, src/OCFD_NS_Jacobian3d.cu:304-440, src/commen_kernel.cu:
//          158-169), one thread per point, no atomics.
#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

// Central derivative at interior index i (0..n-1) of the line through p with
// element stride st; p points at the value of node i.
template <int SV>
__device__ __forceinline__ double cd_d0(const double* __restrict__ p, long long st, int i, int n, int bl, int br,
                                        double h) {
    auto F = [&](int k) { return p[(long long)k * st]; };
    double tmp;
    if (bl && i <= 3) {
        if (i == 0) tmp = (F(1) - F(0));
        else if (i == 1) tmp = (F(1) - F(-1)) * 0.5;
        else if (i == 2) tmp = (F(-2) - 8.0 * F(-1) + 8.0 * F(1) - F(2)) / 12.0;
        else tmp = (F(3) - F(-3) - 9.0 * (F(2) - F(-2)) + 45.0 * (F(1) - F(-1))) / 60.0;
    } else if (br && i >= n - 4) {
        const int t1 = i - n;
        if (t1 == -1) tmp = (F(0) - F(-1));
        else if (t1 == -2) tmp = (F(1) - F(-1)) * 0.5;
        else if (t1 == -3) tmp = (F(-2) - 8.0 * F(-1) + 8.0 * F(1) - F(2)) / 12.0;
        else tmp = (F(3) - F(-3) - 9.0 * (F(2) - F(-2)) + 45.0 * (F(1) - F(-1))) / 60.0;
    } else if (SV == OCFD_CD6) {
        tmp = (45.0 * (F(1) - F(-1)) - 9.0 * (F(2) - F(-2)) + (F(3) - F(-3))) / 60.0;
    } else {
        tmp = (672.0 * (F(1) - F(-1)) - 168.0 * (F(2) - F(-2)) + 32.0 * (F(3) - F(-3)) - 3.0 * (F(4) - F(-4))) /
              840.0;
    }
    return tmp / h;
}

__device__ __forceinline__ bool point_of(const Geo& g, int& i, int& j, int& k) {
    i = blockIdx.x * 64 + threadIdx.x;
    j = blockIdx.y * 4 + threadIdx.y;
    k = blockIdx.z;
    return i < g.nx && j < g.ny;
}

template <int SV>
__global__ void __launch_bounds__(256) k_visc_grad(ViscArgs A) {
    int i, j, k;
    if (!point_of(A.g, i, j, k)) return;
    const Geo& g = A.g;
    const long long q = gidx(g, i, j, k);
    const double* src[4] = {A.u, A.v, A.w, A.T};
    const long long st[3] = {1, g.ldx, g.sz};
    const int idx[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    const double h[3] = {g.hx, g.hy, g.hz};
#pragma unroll
    for (int fI = 0; fI < 4; fI++)
#pragma unroll
        for (int d = 0; d < 3; d++)
            A.grad[(long long)(3 * fI + d) * A.fs + q] =
                cd_d0<SV>(src[fI] + q, st[d], idx[d], n[d], A.bnd[2 * d], A.bnd[2 * d + 1], h[d]);
}

// vis_flux_s_ker / vis_flux_e_ker / vis_flus_ev_ker, src/OCFD_NS_Jacobian3d.cu:80-236
__global__ void __launch_bounds__(256) k_visc_flux(ViscArgs A) {
    int i, j, k;
    if (!point_of(A.g, i, j, k)) return;
    const long long q = gidx(A.g, i, j, k), fs = A.fs;
    const double* G = A.grad + q;
    const double uk = G[0], ui = G[fs], us = G[2 * fs], vk = G[3 * fs], vi = G[4 * fs], vs = G[5 * fs];
    const double wk = G[6 * fs], wi = G[7 * fs], ws = G[8 * fs], Tk = G[9 * fs], Ti = G[10 * fs], Ts = G[11 * fs];
    const double Akx = A.M[M_KX][q], Aky = A.M[M_KY][q], Akz = A.M[M_KZ][q];
    const double Aix = A.M[M_IX][q], Aiy = A.M[M_IY][q], Aiz = A.M[M_IZ][q];
    const double Asx = A.M[M_SX][q], Asy = A.M[M_SY][q], Asz = A.M[M_SZ][q];
    const double Amu = A.amu[q];
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
    const double u = A.u[q], v = A.v[q], w = A.w[q];
    const double E1 = u * s11 + v * s12 + w * s13 + Amuk * Tx;
    const double E2 = u * s12 + v * s22 + w * s23 + Amuk * Ty;
    const double E3 = u * s13 + v * s23 + w * s33 + Amuk * Tz;
    const double Aj1 = A.M[M_JAC][q];
    const double ax[3] = {Akx, Aix, Asx}, ay[3] = {Aky, Aiy, Asy}, az[3] = {Akz, Aiz, Asz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double akx = ax[d] * Aj1, aky = ay[d] * Aj1, akz = az[d] * Aj1;
        double* E = A.ev + (long long)(4 * d) * fs + q;
        E[0] = (akx * s11 + aky * s12 + akz * s13);
        E[fs] = (akx * s12 + aky * s22 + akz * s23);
        E[2 * fs] = (akx * s13 + aky * s23 + akz * s33);
        E[3 * fs] = (akx * E1 + aky * E2 + akz * E3);
    }
}

template <int SV>
__global__ void __launch_bounds__(256) k_visc_div(ViscArgs A) {
    int i, j, k;
    if (!point_of(A.g, i, j, k)) return;
    const Geo& g = A.g;
    const long long q = gidx(g, i, j, k), fs = A.fs;
    const double ajac = A.M[M_JAC][q];
    const long long st[3] = {1, g.ldx, g.sz};
    const int idx[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    const double h[3] = {g.hx, g.hy, g.hz};
#pragma unroll
    for (int m = 0; m < 4; m++) {
        double acc = A.du[(long long)(m + 1) * fs + q];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double vis = cd_d0<SV>(A.ev + (long long)(4 * d + m) * fs + q, st[d], idx[d], n[d],
                                         A.bnd[2 * d], A.bnd[2 * d + 1], h[d]);
            acc = acc + ajac * vis;
        }
        A.du[(long long)(m + 1) * fs + q] = acc;
    }
}

dim3 pgrid(const Geo& g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }

}  // namespace

void launch_visc_grad(const ViscArgs& a, hipStream_t st) {
    if (a.svis == OCFD_CD6) hipLaunchKernelGGL(k_visc_grad<OCFD_CD6>, pgrid(a.g), dim3(64, 4), 0, st, a);
    else hipLaunchKernelGGL(k_visc_grad<OCFD_CD8>, pgrid(a.g), dim3(64, 4), 0, st, a);
}
void launch_visc_flux(const ViscArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_visc_flux, pgrid(a.g), dim3(64, 4), 0, st, a);
}
void launch_visc_div(const ViscArgs& a, hipStream_t st) {
    if (a.svis == OCFD_CD6) hipLaunchKernelGGL(k_visc_div<OCFD_CD6>, pgrid(a.g), dim3(64, 4), 0, st, a);
    else hipLaunchKernelGGL(k_visc_div<OCFD_CD8>, pgrid(a.g), dim3(64, 4), 0, st, a);
}
