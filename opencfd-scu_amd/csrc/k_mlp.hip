// k_mlp.hip - learned hybrid sensor (gfx950).
//
// A small MLP (24 -> 32 -> 32 -> 1, GELU, sigmoid) stands beside the
// hybrid_auto style-2 selector (Set_Scheme_HybridAuto / Comput_Scheme_point_
// Jameson, src/OCFD_Schemes_hybrid_auto.cu:21-68, 611-703).  It has no
// counterpart in the reference.
//
//   k_mlp_sensor  features of a node (8 per direction, FP32) -> BF16 -> two
//                 16x16x16 BF16 MFMA layers per wave of 64 nodes (FP32
//                 accumulation), GELU, the output dot product with a
//                 16-lane reduction, sigmoid; the flag is the plain kp where
//                 the probability lies in [lo, hi] (low confidence), 1 (UP7)
//                 below lo and max(kp, 2) above hi.  The plain dp is
//                 computed in the same kernel.  Written once per step, at the
//                 first RK stage, like the plain sensor.
//   k_mlp_collect features and plain labels (kp > 1) of the nodes of a
//                 sub-lattice, for training.
//   k_mlp_fwd / k_mlp_param   training in FP32: per-sample forward and
//                 backward vectors, then one thread per parameter sums the
//                 batch in sample order (no atomics: the gradients and the
//                 Adam update do not depend on scheduling).
//
// Parameters (1889 floats): W1 [32][24], b1 [32], W2 [32][32], b2 [32],
// W3 [32], b3.
#include <hip/hip_runtime.h>

#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

constexpr int NF = 24, NH = 32;
constexpr int OW1 = 0, OB1 = 768, OW2 = 800, OB2 = 1824, OW3 = 1856, OB3 = 1888;
static_assert(OB3 + 1 == MLP_NPARAM, "parameter layout");
constexpr int PER_SAMPLE = 4 * NH + 2;  // dz1, a1, dz2, a2, dz3, loss term

__device__ __forceinline__ float gelu(float x) {
    const float t = tanhf(0.7978845608028654f * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.0f + t);
}
// The same GELU for inference: with y as above, 0.5 x (1 + tanh y) =
// x - x / (exp(2y) + 1), evaluated with the hardware exp2 and reciprocal
// (about 1 ulp each, far below the BF16 rounding the activations then get);
// for large |y| the reciprocal is exactly 0 or 1.  exp(2y) = 2^(x (a + b x^2))
// with a = 2 sqrt(2/pi) log2(e), b = 0.044715 a.
__device__ __forceinline__ float gelu_fast(float x) {
    constexpr float a = 2.0f * 0.7978845608028654f * 1.4426950408889634f;
    constexpr float b = 0.044715f * a;
    const float e = __builtin_amdgcn_exp2f(x * (a + b * (x * x)));
    return x - x * __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float t = tanhf(0.7978845608028654f * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
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
__device__ void features(const FeatIn<S>& F, int i, int j, int k, float* out) {
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

// ---------------------------------------------------------------------------
// Inference
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short to_bf16(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

constexpr int LS = 36;  // LDS row stride (bf16 elements) of 32-wide tiles

// Z(64 x 32) = X(64 x 32) . W^T for one wave: X rows are nodes (bf16, row
// stride LS), W rows are outputs (bf16 [32][LS]).  acc[t][o] element e holds
// Z[16t + 4*(lane/16) + e][16o + lane%16].
__device__ __forceinline__ void wave_layer(const unsigned short* X, const unsigned short* W, int lane,
                                           v4f acc[4][2]) {
    const int r = lane & 15, g4 = 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int o = 0; o < 2; o++) acc[t][o] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        v4s b[2];
#pragma unroll
        for (int o = 0; o < 2; o++) b[o] = *(const v4s*)&W[(16 * o + r) * LS + 16 * kk + g4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const v4s a = *(const v4s*)&X[(16 * t + r) * LS + 16 * kk + g4];
#pragma unroll
            for (int o = 0; o < 2; o++) acc[t][o] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[o], acc[t][o], 0, 0, 0);
        }
    }
}

template <typename S>
__global__ void __launch_bounds__(256) k_mlp_sensor(FeatIn<S> F, const float* __restrict__ P, double i1, double i2,
                                                    float lo, float hi, int8_t* fx, int8_t* fy, int8_t* fz,
                                                    float* prob_out) {
    __shared__ unsigned short sW1[NH * LS], sW2[NH * LS];
    __shared__ float sb1[NH], sb2[NH], sW3[NH];
    __shared__ unsigned short sX[4][64 * LS];
    __shared__ float sZ[4][64];
    const Geo& g = F.g;
    const int lane = threadIdx.x, wv = threadIdx.y, tid = wv * 64 + lane;
    // weights to LDS (W1 padded from 24 to 32 inputs with zeros)
    for (int e = tid; e < NH * 32; e += 256) {
        const int o = e / 32, kx = e % 32;
        sW1[o * LS + kx] = to_bf16(kx < NF ? P[OW1 + o * NF + kx] : 0.f);
        sW2[o * LS + kx] = to_bf16(P[OW2 + o * NH + kx]);
    }
    if (tid < NH) {
        sb1[tid] = P[OB1 + tid];
        sb2[tid] = P[OB2 + tid];
        sW3[tid] = P[OW3 + tid];
    }
    const Blk B = xcd_block();
    const int i = B.x * 64 + lane - 1, j = B.y * 4 + wv - 1, k = B.z - 1;
    const bool own = i < g.nx && j < g.ny;
    const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1);
    float f[NF];
    features(F, ic, jc, k, f);
    unsigned short* X = sX[wv];
#pragma unroll
    for (int e = 0; e < 32; e++) X[lane * LS + e] = to_bf16(e < NF ? f[e] : 0.f);
    __syncthreads();
    v4f acc[4][2];
    wave_layer(X, sW1, lane, acc);
    __syncthreads();
    const int r = lane & 15, g4 = 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int o = 0; o < 2; o++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                X[(16 * t + g4 + e) * LS + 16 * o + r] = to_bf16(gelu_fast(acc[t][o][e] + sb1[16 * o + r]));
    __syncthreads();
    wave_layer(X, sW2, lane, acc);
    // output layer: per node sum over the 32 hidden units (16 lanes x 2 tiles)
    float z[4][4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < 2; o++) s += gelu_fast(acc[t][o][e] + sb2[16 * o + r]) * sW3[16 * o + r];
            s += __shfl_xor(s, 8);
            s += __shfl_xor(s, 4);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 1);
            z[t][e] = s;
        }
    if (r == 0)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int e = 0; e < 4; e++) sZ[wv][16 * t + g4 + e] = z[t][e];
    __syncthreads();
    if (!own) return;
    const float prob = 1.f / (1.f + expf(-(sZ[wv][lane] + P[OB3])));
    const long long q = gidx(g, i, j, k);
    const int kp = jameson_kp(F, q, i1, i2);
    int fl = kp;
    if (prob < lo) fl = 1;
    else if (prob > hi) fl = kp > 2 ? kp : 2;
    if (j >= 0 && k >= 0) fx[q] = (int8_t)fl;
    if (i >= 0 && k >= 0) fy[q] = (int8_t)fl;
    if (i >= 0 && j >= 0) fz[q] = (int8_t)fl;
    if (prob_out) prob_out[q] = prob;
}

// ---------------------------------------------------------------------------
// Training data and training
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_mlp_collect(FeatIn<S> F, int stride, int nlx, int nly, int nlz, double i1,
                                                     double i2, float* X, float* y) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nlx * nly * nlz;
    if (t >= n) return;
    const int a = (int)(t % nlx), b = (int)((t / nlx) % nly), c = (int)(t / ((long long)nlx * nly));
    const int i = a * stride, j = b * stride, k = c * stride;
    float f[NF];
    features(F, i, j, k, f);
#pragma unroll
    for (int e = 0; e < NF; e++) X[t * NF + e] = f[e];
    y[t] = jameson_kp(F, gidx(F.g, i, j, k), i1, i2) > 1 ? 1.f : 0.f;
}

// Per sample of the batch: forward, weighted BCE term and the backward
// vectors dz1, a1, dz2, a2, dz3 (the batch mean is taken in k_mlp_param).
__global__ void __launch_bounds__(256) k_mlp_fwd(const float* __restrict__ P, const float* __restrict__ X,
                                                 const float* __restrict__ y, const int* __restrict__ idx, int nb,
                                                 float wpos, float* __restrict__ out) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= nb) return;
    const long long row = idx ? idx[s] : s;
    float x[NF];
#pragma unroll
    for (int e = 0; e < NF; e++) x[e] = X[row * NF + e];
    float z1[NH], a1[NH], z2[NH], a2[NH];
    for (int o = 0; o < NH; o++) {
        float acc = P[OB1 + o];
#pragma unroll
        for (int e = 0; e < NF; e++) acc += P[OW1 + o * NF + e] * x[e];
        z1[o] = acc;
        a1[o] = gelu(acc);
    }
    for (int o = 0; o < NH; o++) {
        float acc = P[OB2 + o];
#pragma unroll
        for (int h = 0; h < NH; h++) acc += P[OW2 + o * NH + h] * a1[h];
        z2[o] = acc;
        a2[o] = gelu(acc);
    }
    float z3 = P[OB3];
#pragma unroll
    for (int h = 0; h < NH; h++) z3 += P[OW3 + h] * a2[h];
    const float prob = 1.f / (1.f + expf(-z3));
    const float yy = y[row];
    const float w = yy > 0.5f ? wpos : 1.f;
    const float eps = 1e-7f;
    const float lt = -w * (yy * logf(prob + eps) + (1.f - yy) * logf(1.f - prob + eps));
    const float dz3 = w * (prob - yy);
    float* o_ = out + (long long)s * PER_SAMPLE;
    float dz2[NH];
    for (int o = 0; o < NH; o++) dz2[o] = dz3 * P[OW3 + o] * gelu_grad(z2[o]);
    for (int h = 0; h < NH; h++) {
        float da = 0.f;
        for (int o = 0; o < NH; o++) da += P[OW2 + o * NH + h] * dz2[o];
        o_[h] = da * gelu_grad(z1[h]);  // dz1
    }
    for (int h = 0; h < NH; h++) {
        o_[NH + h] = a1[h];
        o_[2 * NH + h] = dz2[h];
        o_[3 * NH + h] = a2[h];
    }
    o_[4 * NH] = dz3;
    o_[4 * NH + 1] = lt;
}

// One thread per parameter: the batch-mean gradient summed in sample order,
// then (adam_t > 0) the Adam update of the parameter.  Thread MLP_NPARAM sums
// the loss.
__global__ void __launch_bounds__(256) k_mlp_param(float* __restrict__ P, const float* __restrict__ X,
                                                   const int* __restrict__ idx, const float* __restrict__ per, int nb,
                                                   float* __restrict__ grad, float* __restrict__ m1,
                                                   float* __restrict__ m2, int adam_t, float lr, float* loss) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > MLP_NPARAM) return;
    const float inv = 1.f / (float)nb;
    float gsum = 0.f;
    if (p == MLP_NPARAM) {
        for (int s = 0; s < nb; s++) gsum += per[(long long)s * PER_SAMPLE + 4 * NH + 1];
        *loss = gsum * inv;
        return;
    }
    if (p < OB1) {  // W1[o][e] = sum dz1[o] x[e]
        const int o = p / NF, e = p % NF;
        for (int s = 0; s < nb; s++) {
            const long long row = idx ? idx[s] : s;
            gsum += per[(long long)s * PER_SAMPLE + o] * X[row * NF + e];
        }
    } else if (p < OW2) {
        const int o = p - OB1;
        for (int s = 0; s < nb; s++) gsum += per[(long long)s * PER_SAMPLE + o];
    } else if (p < OB2) {  // W2[o][h] = sum dz2[o] a1[h]
        const int o = (p - OW2) / NH, h = (p - OW2) % NH;
        for (int s = 0; s < nb; s++) {
            const float* r = per + (long long)s * PER_SAMPLE;
            gsum += r[2 * NH + o] * r[NH + h];
        }
    } else if (p < OW3) {
        const int o = p - OB2;
        for (int s = 0; s < nb; s++) gsum += per[(long long)s * PER_SAMPLE + 2 * NH + o];
    } else if (p < OB3) {  // W3[h] = sum dz3 a2[h]
        const int h = p - OW3;
        for (int s = 0; s < nb; s++) {
            const float* r = per + (long long)s * PER_SAMPLE;
            gsum += r[4 * NH] * r[3 * NH + h];
        }
    } else {
        for (int s = 0; s < nb; s++) gsum += per[(long long)s * PER_SAMPLE + 4 * NH];
    }
    const float gr = gsum * inv;
    if (grad) grad[p] = gr;
    if (adam_t > 0) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mm = b1 * m1[p] + (1.f - b1) * gr;
        const float vv = b2 * m2[p] + (1.f - b2) * gr * gr;
        m1[p] = mm;
        m2[p] = vv;
        const float mh = mm / (1.f - powf(b1, (float)adam_t));
        const float vh = vv / (1.f - powf(b2, (float)adam_t));
        P[p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

template <typename S>
FeatIn<S> feat_in(const MlpSensorArgsT<S>& a) {
    FeatIn<S> F;
    F.g = a.g;
    F.d = a.d; F.u = a.u; F.v = a.v; F.w = a.w; F.T = a.T;
    for (int m = 0; m < M_COUNT; m++) F.M[m] = a.M[m];
    F.p00 = a.c.p00;
    F.rama = a.c.rama;
    return F;
}

}  // namespace

template <typename S>
void launch_mlp_sensor(const MlpSensorArgsT<S>& a, hipStream_t st) {
    const Geo& g = a.g;
    dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, g.nz + 1);
    hipLaunchKernelGGL(k_mlp_sensor<S>, grid, dim3(64, 4), 0, st, feat_in(a), a.params, a.c.hy_i1, a.c.hy_i2, a.lo,
                       a.hi, a.fx, a.fy, a.fz, a.prob);
}

template <typename S>
void launch_mlp_collect(const MlpSensorArgsT<S>& a, int stride, float* X, float* y, hipStream_t st) {
    const Geo& g = a.g;
    const int nlx = (g.nx + stride - 1) / stride, nly = (g.ny + stride - 1) / stride, nlz = (g.nz + stride - 1) / stride;
    const long long n = (long long)nlx * nly * nlz;
    hipLaunchKernelGGL(k_mlp_collect<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, feat_in(a), stride, nlx,
                       nly, nlz, a.c.hy_i1, a.c.hy_i2, X, y);
}

long long mlp_collect_count(const Geo& g, int stride) {
    return (long long)((g.nx + stride - 1) / stride) * ((g.ny + stride - 1) / stride) * ((g.nz + stride - 1) / stride);
}

void launch_mlp_step(float* params, const float* X, const float* y, const int* idx, int nb, float wpos, float* per,
                     float* grad, float* m1, float* m2, int adam_t, float lr, float* loss, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_fwd, dim3((nb + 255) / 256), dim3(256), 0, st, params, X, y, idx, nb, wpos, per);
    hipLaunchKernelGGL(k_mlp_param, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, st, params, X, idx, per, nb,
                       grad, m1, m2, adam_t, lr, loss);
}

size_t mlp_per_sample_floats() { return PER_SAMPLE; }

template void launch_mlp_sensor<double>(const MlpSensorArgsT<double>&, hipStream_t);
template void launch_mlp_sensor<float>(const MlpSensorArgsT<float>&, hipStream_t);
template void launch_mlp_collect<double>(const MlpSensorArgsT<double>&, int, float*, float*, hipStream_t);
template void launch_mlp_collect<float>(const MlpSensorArgsT<float>&, int, float*, float*, hipStream_t);
