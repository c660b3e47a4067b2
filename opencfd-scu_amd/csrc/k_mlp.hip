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
//                 computed in the same kernel (mlp_dev.h holds these steps,
//                 which the x sweep of the first RK stage runs as well).
//                 Flags are written once per step, at the first RK stage,
//                 like the plain sensor's.
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

#include "mlp_dev.h"
#include "ocfd_dev.h"
#include "ocfd_kernels.h"

namespace {

using namespace mlp;
static_assert(OB3 + 1 == MLP_NPARAM, "parameter layout");
constexpr int PER_SAMPLE = 4 * NH + 2;  // dz1, a1, dz2, a2, dz3, loss term

__device__ __forceinline__ float gelu(float x) {
    const float t = tanhf(0.7978845608028654f * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float t = tanhf(0.7978845608028654f * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
}

// ---------------------------------------------------------------------------
// Inference
// ---------------------------------------------------------------------------
// plane = 0: every node with interior indices >= -1 (the whole pass);
// plane = 1: the nodes (i >= 0, j = -1, k >= 0), whose flag only the y sweep
// reads (face -1/2); plane = 2: the nodes (i >= 0, j >= 0, k = -1) for the z
// sweep.  The x sweep of the first RK stage computes the other nodes itself
// (k_sweep with the learned sensor), and planes 1 and 2 complete it.
// The lane fragments of the weights in the layout lane_weights_packed reads
// (one wave).
__global__ void __launch_bounds__(64) k_mlp_pack(const float* __restrict__ P, float* __restrict__ pk) {
    Lane L;
    lane_weights(P, threadIdx.x, L);
    pack_lane(L, threadIdx.x, pk);
}

// P = the packed weights (k_mlp_pack).
template <typename S>
__global__ void __launch_bounds__(256) k_mlp_sensor(FeatIn<S> F, const float* __restrict__ P, double i1, double i2,
                                                    float lo, float hi, int8_t* fx, int8_t* fy, int8_t* fz,
                                                    float* prob_out, int plane) {
    __shared__ unsigned short sX[4][64 * LS];
    const Geo& g = F.g;
    const int lane = threadIdx.x, wv = threadIdx.y;
    Lane L;
    lane_weights_packed(P, lane, L);
    const Blk B = xcd_block();
    int i, j, k;
    bool own;
    if (plane == 0) {
        i = B.x * 64 + lane - 1; j = B.y * 4 + wv - 1; k = B.z - 1;
        own = i < g.nx && j < g.ny;
    } else if (plane == 1) {
        i = B.x * 64 + lane; j = -1; k = B.y * 4 + wv;
        own = i < g.nx && k < g.nz;
    } else {
        i = B.x * 64 + lane; j = B.y * 4 + wv; k = -1;
        own = i < g.nx && j < g.ny;
    }
    const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1), kc = min(k, g.nz - 1);
    float f[NF];
    features(F, ic, jc, kc, f);
    const float prob = wave_prob(f, sX[wv], L, lane);
    if (!own) return;
    const long long q = gidx(g, i, j, k);
    const int fl = flag_rule(prob, jameson_kp(F, q, i1, i2), lo, hi);
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
void launch_mlp_sensor(const MlpSensorArgsT<S>& a, int plane, hipStream_t st) {
    const Geo& g = a.g;
    dim3 grid;
    if (plane == 0) grid = dim3((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, g.nz + 1);
    else if (plane == 1) grid = dim3((g.nx + 63) / 64, (g.nz + 3) / 4, 1);
    else grid = dim3((g.nx + 63) / 64, (g.ny + 3) / 4, 1);
    hipLaunchKernelGGL(k_mlp_sensor<S>, grid, dim3(64, 4), 0, st, feat_in(a), a.params, a.c.hy_i1, a.c.hy_i2, a.lo,
                       a.hi, a.fx, a.fy, a.fz, a.prob, plane);
}

template <typename S>
void launch_mlp_collect(const MlpSensorArgsT<S>& a, int stride, float* X, float* y, hipStream_t st) {
    const Geo& g = a.g;
    const int nlx = (g.nx + stride - 1) / stride, nly = (g.ny + stride - 1) / stride, nlz = (g.nz + stride - 1) / stride;
    const long long n = (long long)nlx * nly * nlz;
    hipLaunchKernelGGL(k_mlp_collect<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, feat_in(a), stride, nlx,
                       nly, nlz, a.c.hy_i1, a.c.hy_i2, X, y);
}

void launch_mlp_pack(const float* params, float* packed, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, st, params, packed);
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

template void launch_mlp_sensor<double>(const MlpSensorArgsT<double>&, int, hipStream_t);
template void launch_mlp_sensor<float>(const MlpSensorArgsT<float>&, int, hipStream_t);
template void launch_mlp_collect<double>(const MlpSensorArgsT<double>&, int, float*, float*, hipStream_t);
template void launch_mlp_collect<float>(const MlpSensorArgsT<float>&, int, float*, float*, hipStream_t);
