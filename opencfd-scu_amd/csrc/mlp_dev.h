// mlp_dev.h - device side of the learned hybrid sensor's inference, shared by
// its own pass (k_mlp.hip) and the x sweep that runs it at the first RK stage
// (k_inviscid.hip): node features, the plain Jameson kp, and one wave's
// BF16-MFMA forward of 64 nodes.
//
// Network: 24 -> 32 -> 32 -> 1, GELU (tanh form), sigmoid; parameters
// (1889 floats) W1 [32][24], b1 [32], W2 [32][32], b2 [32], W3 [32], b3.
#pragma once

#include <hip/hip_runtime.h>

#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace mlp {

constexpr int NF = 24, NH = 32;
constexpr int OW1 = 0, OB1 = 768, OW2 = 800, OB2 = 1824, OW3 = 1856, OB3 = 1888;

// GELU for inference: with y = sqrt(2/pi) (x + 0.044715 x^3), 0.5 x (1 +
// tanh y) = x - x / (exp(2y) + 1), evaluated with the hardware exp2 and
// reciprocal (about 1 ulp each, far below the BF16 rounding the activations
// then get); for large |y| the reciprocal is exactly 0 or 1.  exp(2y) =
// 2^(x (a + b x^2)) with a = 2 sqrt(2/pi) log2(e), b = 0.044715 a.
__device__ __forceinline__ float gelu_fast(float x) {
    constexpr float a = 2.0f * 0.7978845608028654f * 1.4426950408889634f;
    constexpr float b = 0.044715f * a;
    const float e = __builtin_amdgcn_exp2f(x * (a + b * (x * x)));
    return x - x * __builtin_amdgcn_rcpf(e + 1.0f);
}

// ---------------------------------------------------------------------------
// Features of node (i, j, k) (interior indices, >= -1).  p = p00*rho*T,
// jumps (b - a)/(b + a) (0 where b + a = 0), U = (A . u)/|A| with the
// direction's cofactor row, c = sqrt(T)/Ma.
// ---------------------------------------------------------------------------
template <typename S>
struct FeatIn {
    Geo g;
    const S *d, *u, *v, *w, *T;
    const double* M[M_COUNT];
    double p00, rama;
};

// FP32 arithmetic with the hardware reciprocal and reciprocal square root
// (about 1 ulp each): the features are FP32 values that the network reads
// in BF16, and they stay within a few 1e-7 of an FP64 evaluation.
__device__ __forceinline__ float jump(float a, float b) {
    const float s = a + b;
    return s != 0.f ? (b - a) * __builtin_amdgcn_rcpf(s) : 0.f;
}

template <typename S>
__device__ __forceinline__ void features(const FeatIn<S>& F, int i, int j, int k, float* out) {
    const Geo& g = F.g;
    const long long q = gidx(g, i, j, k);
    const long long st[3] = {1, (long long)g.ldx, g.sz};
    const float p00 = float(F.p00);
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
        float p[8], r[8];
#pragma unroll
        for (int o = 0; o < 8; o++) {  // nodes n-3..n+4
            const long long qo = q + (long long)(o - 3) * st[dd];
            const float rho = float(F.d[qo]);
            r[o] = rho;
            p[o] = p00 * rho * float(F.T[qo]);
        }
        // node n is p[3]; intervals m = -1, 0, 1 are (p[2],p[3]), (p[3],p[4]), (p[4],p[5])
        out[8 * dd + 0] = jump(p[2], p[3]);
        out[8 * dd + 1] = jump(p[3], p[4]);
        out[8 * dd + 2] = jump(p[4], p[5]);
        out[8 * dd + 3] = jump(r[2], r[3]);
        out[8 * dd + 4] = jump(r[3], r[4]);
        out[8 * dd + 5] = jump(r[4], r[5]);
        float mx = fabsf(jump(p[0], p[1]));
        mx = fmaxf(mx, fabsf(jump(p[1], p[2])));
        mx = fmaxf(mx, fabsf(jump(p[5], p[6])));
        mx = fmaxf(mx, fabsf(jump(p[6], p[7])));
        out[8 * dd + 6] = mx;
        const double* A1 = F.M[3 * dd];
        const double* A2 = F.M[3 * dd + 1];
        const double* A3 = F.M[3 * dd + 2];
        float U[2];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const long long qm = q + m * st[dd];
            const float a1 = float(A1[qm]), a2 = float(A2[qm]), a3 = float(A3[qm]);
            U[m] = (a1 * float(F.u[qm]) + a2 * float(F.v[qm]) + a3 * float(F.w[qm])) *
                   __builtin_amdgcn_rsqf(a1 * a1 + a2 * a2 + a3 * a3);
        }
        const float c = sqrtf(fmaxf(float(F.T[q]), 0.f)) * float(F.rama);
        out[8 * dd + 7] = c != 0.f ? (U[1] - U[0]) * __builtin_amdgcn_rcpf(c) : 0.f;
    }
}

// Plain Jameson kp of node q (the arithmetic of k_sensor).
template <typename S>
__device__ __forceinline__ int jameson_kp(const FeatIn<S>& F, long long q, double i1, double i2) {
    const Geo& g = F.g;
    auto P = [&](long long o) { return F.p00 * double(F.d[o]) * double(F.T[o]); };
    const double pc = P(q);
    const long long st[3] = {1, (long long)g.ldx, g.sz};
    double dp0 = 0.0;
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
        const double pm = P(q - st[dd]), pp = P(q + st[dd]);
        dp0 += fabs(-pm + 2 * pc - pp) / (pm + 2 * pc + pp);
    }
    int kp = 1;
    if (dp0 > i1) kp += 1;
    if (dp0 > i2) kp += 1;
    return kp;
}

// The flag rule: the plain kp where the probability lies in [lo, hi] (low
// confidence), 1 (UP7) below lo and max(kp, 2) above hi.
__device__ __forceinline__ int flag_rule(float prob, int kp, float lo, float hi) {
    int fl = kp;
    if (prob < lo) fl = 1;
    else if (prob > hi) fl = kp > 2 ? kp : 2;
    return fl;
}

// ---------------------------------------------------------------------------
// One wave's forward of 64 nodes with 16x16x16 BF16 MFMA (FP32 accumulation)
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short to_bf16(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

constexpr int LS = 36;  // LDS row stride (bf16 elements) of a wave's 64 x 32 activation tile

// Each layer is computed transposed, Z^T = W . X^T, with the weights as the
// A operand and the activations as the B operand of
// v_mfma_f32_16x16x16bf16 (A: lane l holds A[l%16][4(l/16) .. +3]; B: lane l
// holds B[4(l/16) .. +3][l%16]; D: lane l holds D[4(l/16) .. +3][l%16]).  The
// output of a layer then sits in each lane exactly as the next layer's B
// operand needs it (node 16t + l%16, hidden units 16o + 4(l/16) .. +3), so
// only the input features pass through LDS.
//
// The weights a lane feeds: w1[kk][o] = W1[16o + lane%16][16kk + 4(lane/16)
// .. +3] in BF16 (W1 padded from 24 to 32 inputs with zeros), w2 the same for
// W2; b1[o], b2[o], w3[o] = the entries 16o + 4(lane/16) .. +3 of b1, b2, W3.
struct Lane {
    v4s w1[2][2], w2[2][2];
    v4f b1[2], b2[2], w3[2];
    float b3;
};
__device__ __forceinline__ void lane_weights(const float* __restrict__ P, int lane, Lane& L) {
    const int r = lane & 15, g4 = 4 * (lane >> 4);
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int o = 0; o < 2; o++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = 16 * o + r, col = 16 * kk + g4 + e;
                L.w1[kk][o][e] = (short)to_bf16(col < NF ? P[OW1 + row * NF + col] : 0.f);
                L.w2[kk][o][e] = (short)to_bf16(P[OW2 + row * NH + col]);
            }
#pragma unroll
    for (int o = 0; o < 2; o++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            L.b1[o][e] = P[OB1 + 16 * o + g4 + e];
            L.b2[o][e] = P[OB2 + 16 * o + g4 + e];
            L.w3[o][e] = P[OW3 + 16 * o + g4 + e];
        }
    L.b3 = P[OB3];
}

// The same fragments packed once per parameter set (k_mlp_pack), lane
// fastest, so that a wave reads each with one coalesced access: w1 and w2
// as [kk][o][lane] groups of 4 BF16 (2 floats of storage), b1, b2, w3 as
// [o][lane] groups of 4 floats, then b3.
constexpr int PK_W1 = 0, PK_W2 = 512, PK_B1 = 1024, PK_B2 = 1536, PK_W3 = 2048, PK_B3 = 2560;
static_assert(PK_B3 + 1 <= MLP_PACKED_FLOATS, "packed weight layout");
__device__ __forceinline__ void pack_lane(const Lane& L, int lane, float* __restrict__ pk) {
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int o = 0; o < 2; o++) {
            *(v4s*)&pk[PK_W1 + 2 * ((2 * kk + o) * 64 + lane)] = L.w1[kk][o];
            *(v4s*)&pk[PK_W2 + 2 * ((2 * kk + o) * 64 + lane)] = L.w2[kk][o];
        }
#pragma unroll
    for (int o = 0; o < 2; o++) {
        *(v4f*)&pk[PK_B1 + 4 * (64 * o + lane)] = L.b1[o];
        *(v4f*)&pk[PK_B2 + 4 * (64 * o + lane)] = L.b2[o];
        *(v4f*)&pk[PK_W3 + 4 * (64 * o + lane)] = L.w3[o];
    }
    if (lane == 0) pk[PK_B3] = L.b3;
}
__device__ __forceinline__ void lane_weights_packed(const float* __restrict__ pk, int lane, Lane& L) {
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int o = 0; o < 2; o++) {
            L.w1[kk][o] = *(const v4s*)&pk[PK_W1 + 2 * ((2 * kk + o) * 64 + lane)];
            L.w2[kk][o] = *(const v4s*)&pk[PK_W2 + 2 * ((2 * kk + o) * 64 + lane)];
        }
#pragma unroll
    for (int o = 0; o < 2; o++) {
        L.b1[o] = *(const v4f*)&pk[PK_B1 + 4 * (64 * o + lane)];
        L.b2[o] = *(const v4f*)&pk[PK_B2 + 4 * (64 * o + lane)];
        L.w3[o] = *(const v4f*)&pk[PK_W3 + 4 * (64 * o + lane)];
    }
    L.b3 = pk[PK_B3];
}

// One layer for the wave's 64 nodes: acc[t][o] element e = Z[node 16t +
// lane%16][out 16o + 4(lane/16) + e] from the B fragments x[t][kk] = X[node
// 16t + lane%16][16kk + 4(lane/16) .. +3].
__device__ __forceinline__ void wave_layer(const v4s (&w)[2][2], const v4s (&x)[4][2], v4f acc[4][2]) {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int o = 0; o < 2; o++) acc[t][o] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int o = 0; o < 2; o++)
                acc[t][o] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w[kk][o], x[t][kk], acc[t][o], 0, 0, 0);
}

// Probability of the node of each lane of the wave.  f = the lane's 24
// features; X = the wave's 64 x LS tile in LDS, through which the features
// reach the lanes that feed them to the first layer.  Every thread of the
// block calls it (one block barrier).
__device__ __forceinline__ float wave_prob(const float* f, unsigned short* X, const Lane& L, int lane) {
#pragma unroll
    for (int e = 0; e < 32; e++) X[lane * LS + e] = to_bf16(e < NF ? f[e] : 0.f);
    __syncthreads();
    const int r = lane & 15, g4 = 4 * (lane >> 4);
    v4s x[4][2];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int kk = 0; kk < 2; kk++) x[t][kk] = *(const v4s*)&X[(16 * t + r) * LS + 16 * kk + g4];
    v4f acc[4][2];
    wave_layer(L.w1, x, acc);
    // bias + GELU, rounded to BF16: the second layer's B fragments
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int o = 0; o < 2; o++)
#pragma unroll
            for (int e = 0; e < 4; e++) x[t][o][e] = (short)to_bf16(gelu_fast(acc[t][o][e] + L.b1[o][e]));
    wave_layer(L.w2, x, acc);
    // output layer: the lane's 8 hidden units of node 16t + lane%16, then the
    // sum over the 4 lanes that hold that node's other units
    float z[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float s = 0.f;
#pragma unroll
        for (int o = 0; o < 2; o++)
#pragma unroll
            for (int e = 0; e < 4; e++) s += gelu_fast(acc[t][o][e] + L.b2[o][e]) * L.w3[o][e];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        z[t] = s;
    }
    // this lane's node is 16 (lane/16) + lane%16
    const int tl = lane >> 4;
    const float zl = tl == 0 ? z[0] : tl == 1 ? z[1] : tl == 2 ? z[2] : z[3];
    return 1.f / (1.f + expf(-(zl + L.b3)));
}

}  // namespace mlp
