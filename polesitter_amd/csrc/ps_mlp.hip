// ps_mlp.hip -- the learned far-field correction (no reference counterpart).
//
// A first-order FMM applies each cell's far field as one constant over the
// target leaf (L2P of the field at the leaf centre, H:1158-1235) and keeps
// only monopole + dipole of the sources (H:1065-1092).  The dominant missing
// term is each level's field gradient times the particle's offset from that
// level's cell centre.  The correction is a small MLP per particle:
//
//   features F (24): for the leaf L, its parent P and grandparent G, the
//     particle's offset from the cell centre in half widths (3 each) and the
//     cell's own M2L field (its local minus its parent's, 3 each), the field
//     arriving from above G (3), the leaf's mass centroid offset (3); fields
//     divided by s = |E_L|, the leaf's total far field.
//   net: 24 -> 64 -> 64 -> 64 -> 3, GELU (exact, erf) on the hidden layers,
//     FP32 weights and accumulation.
//   correction: a_i += s * net(F_i), i.e. f_i += m_i * s * net(F_i).
//
// Parameters are one packed FP32 array (PS_MLP_PARAMS floats):
//   W1[64][24] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3].
// Weight reads are uniform across a wavefront, so they are scalar loads
// (SGPR operands); activations of a thread live in registers between layers,
// staged through an LDS row per thread.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kIn = PS_MLP_IN, kH = PS_MLP_HIDDEN, kOut = 3;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kOut * kH;
static_assert(oB4 + kOut == PS_MLP_PARAMS, "parameter layout");

constexpr int kFwdBlock = 128;
constexpr int kRow = kH + 1;  // LDS row stride (no bank conflicts on columns)

__device__ __forceinline__ float gelu(float x) {
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
           x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// ---------------------------------------------------------------------------
// features of particle i (sorted order); returns s = |E_L| (0: no far field)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mlp_features(size_t i, const float4* __restrict__ pos4,
                                              const uint32_t* __restrict__ leaf_of,
                                              const float4* __restrict__ geom,
                                              const float4* __restrict__ mp,
                                              const float4* __restrict__ loc,
                                              const uint32_t* __restrict__ parent,
                                              float f[kIn]) {
    const float4 p = pos4[i];
    const uint32_t L = leaf_of[i];
    const uint32_t P = parent[L];
    const uint32_t G = P != PS_NONE ? parent[P] : PS_NONE;
    const float4 gL = geom[L], lL = loc[L], mL = mp[L];
    const float4 gP = P != PS_NONE ? geom[P] : gL;
    const float4 lP = P != PS_NONE ? loc[P] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 gG = G != PS_NONE ? geom[G] : gP;
    const float4 lG = G != PS_NONE ? loc[G] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float s = sqrtf(lL.y * lL.y + lL.z * lL.z + lL.w * lL.w);
    const float inv = s > 0.f ? 1.0f / s : 0.f;
    const float hl = 1.0f / gL.w, hp = 1.0f / gP.w, hg = 1.0f / gG.w;
    f[0] = (p.x - gL.x) * hl;
    f[1] = (p.y - gL.y) * hl;
    f[2] = (p.z - gL.z) * hl;
    f[3] = (lL.y - lP.y) * inv;
    f[4] = (lL.z - lP.z) * inv;
    f[5] = (lL.w - lP.w) * inv;
    const bool hasP = P != PS_NONE, hasG = G != PS_NONE;
    f[6] = hasP ? (p.x - gP.x) * hp : 0.f;
    f[7] = hasP ? (p.y - gP.y) * hp : 0.f;
    f[8] = hasP ? (p.z - gP.z) * hp : 0.f;
    f[9] = (lP.y - lG.y) * inv;
    f[10] = (lP.z - lG.z) * inv;
    f[11] = (lP.w - lG.w) * inv;
    f[12] = hasG ? (p.x - gG.x) * hg : 0.f;
    f[13] = hasG ? (p.y - gG.y) * hg : 0.f;
    f[14] = hasG ? (p.z - gG.z) * hg : 0.f;
    // grandparent's own contribution needs its parent: use the whole field
    // above G as one more input instead
    f[15] = lG.y * inv;
    f[16] = lG.z * inv;
    f[17] = lG.w * inv;
    f[18] = 0.f;
    f[19] = 0.f;
    f[20] = 0.f;
    const float cm = mL.x != 0.f ? hl / mL.x : 0.f;
    f[21] = mL.y * cm;
    f[22] = mL.z * cm;
    f[23] = mL.w * cm;
    if (s == 0.f)
#pragma unroll
        for (int k = 0; k < kIn; ++k) f[k] = 0.f;
    return s;
}

// one hidden layer (kH outputs) of the thread's activations; W row-major,
// results go through the thread's LDS row and come back as registers
template <int NIN>
__device__ __forceinline__ void dense_gelu(const float* __restrict__ W, const float* __restrict__ b,
                                           const float in[NIN], float out[kH], float* row,
                                           float* zrow) {
    for (int jb = 0; jb < kH; jb += 8) {
        float acc[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) acc[jj] = b[jb + jj];
#pragma unroll
        for (int k = 0; k < NIN; ++k) {
            const float v = in[k];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) acc[jj] = fmaf(W[(jb + jj) * NIN + k], v, acc[jj]);
        }
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            if (zrow) zrow[jb + jj] = acc[jj];
            row[jb + jj] = gelu(acc[jj]);
        }
    }
#pragma unroll
    for (int k = 0; k < kH; ++k) out[k] = row[k];
}

__device__ __forceinline__ void mlp_net(const float* __restrict__ W, const float f[kIn],
                                        float* row, float o[kOut]) {
    float h[kH], g[kH];
    dense_gelu<kIn>(W + oW1, W + oB1, f, h, row, nullptr);
    dense_gelu<kH>(W + oW2, W + oB2, h, g, row, nullptr);
    dense_gelu<kH>(W + oW3, W + oB3, g, h, row, nullptr);
#pragma unroll
    for (int c = 0; c < kOut; ++c) {
        float acc = W[oB4 + c];
#pragma unroll
        for (int k = 0; k < kH; ++k) acc = fmaf(W[oW4 + c * kH + k], h[k], acc);
        o[c] = acc;
    }
}

// row r of (feat, scale) = features of particle idx[r] (idx null: r); rows
// whose index is not a particle (>= n) get zeros
__global__ __launch_bounds__(kFwdBlock) void mlp_features_kernel(
    size_t n, const uint32_t* __restrict__ idx, size_t rows, const float4* __restrict__ pos4,
    const uint32_t* __restrict__ leaf_of, const float4* __restrict__ geom,
    const float4* __restrict__ mp, const float4* __restrict__ loc,
    const uint32_t* __restrict__ parent, float* __restrict__ feat, float* __restrict__ scale) {
    const size_t r = (size_t)blockIdx.x * kFwdBlock + threadIdx.x;
    if (r >= rows) return;
    const size_t i = idx ? (size_t)idx[r] : r;
    float f[kIn];
    float s = 0.f;
    if (i < n) {
        s = mlp_features(i, pos4, leaf_of, geom, mp, loc, parent, f);
    } else {
#pragma unroll
        for (int k = 0; k < kIn; ++k) f[k] = 0.f;
    }
    scale[r] = s;
#pragma unroll
    for (int k = 0; k < kIn; ++k) feat[r * kIn + k] = f[k];
}

__global__ __launch_bounds__(kFwdBlock) void mlp_forward_kernel(const float* __restrict__ W,
                                                                const float* __restrict__ feat,
                                                                size_t n, float* __restrict__ out) {
    __shared__ float rows[kFwdBlock * kRow];
    const size_t i = (size_t)blockIdx.x * kFwdBlock + threadIdx.x;
    float f[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) f[k] = i < n ? feat[i * kIn + k] : 0.f;
    float o[kOut];
    mlp_net(W, f, rows + threadIdx.x * kRow, o);
    if (i < n)
#pragma unroll
        for (int c = 0; c < kOut; ++c) out[i * kOut + c] = o[c];
}

// the corrected tick's last step: f_i += m_i * s_i * net(F_i)
__global__ __launch_bounds__(kFwdBlock) void mlp_apply_kernel(
    size_t n, const float* __restrict__ W, const float4* __restrict__ pos4,
    const uint32_t* __restrict__ leaf_of, const float4* __restrict__ geom,
    const float4* __restrict__ mp, const float4* __restrict__ loc,
    const uint32_t* __restrict__ parent, float* __restrict__ fx, float* __restrict__ fy,
    float* __restrict__ fz) {
    __shared__ float rows[kFwdBlock * kRow];
    const size_t i = (size_t)blockIdx.x * kFwdBlock + threadIdx.x;
    const size_t ii = i < n ? i : (n ? n - 1 : 0);
    float f[kIn];
    const float s = mlp_features(ii, pos4, leaf_of, geom, mp, loc, parent, f);
    float o[kOut];
    mlp_net(W, f, rows + threadIdx.x * kRow, o);
    if (i >= n || s == 0.f) return;
    const float w = pos4[i].w * s;
    fx[i] = fmaf(w, o[0], fx[i]);
    fy[i] = fmaf(w, o[1], fy[i]);
    fz[i] = fmaf(w, o[2], fz[i]);
}

// ---------------------------------------------------------------------------
// training: gradient of loss = (1 / n_total) sum |net(F) - t|^2 over a block
// of kS samples, written as per-block partial sums (deterministic)
// ---------------------------------------------------------------------------
constexpr int kS = 32;        // samples per block
constexpr int kTrain = 64;    // threads per block
constexpr int kFRow = kIn + 1;

__global__ __launch_bounds__(kTrain) void mlp_grad_kernel(const float* __restrict__ W,
                                                          const float* __restrict__ feat,
                                                          const float* __restrict__ target,
                                                          size_t n, float inv_total,
                                                          float* __restrict__ part,
                                                          float* __restrict__ loss_part) {
    __shared__ float F[kS * kFRow];
    __shared__ float Z1[kS * kRow], Z2[kS * kRow], Z3[kS * kRow];
    __shared__ float DA[kS * kRow], DB[kS * kRow];
    __shared__ float O[kS * 4];
    __shared__ float lsum[kS];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * kS;
    float* P = part + (size_t)blockIdx.x * PS_MLP_PARAMS;
    // a: forward, storing pre-activations; delta4 = (2 / n_total)(out - t)
    if (t < kS) {
        const size_t i = base + t;
        const bool ok = i < n;
        float f[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) {
            f[k] = ok ? feat[i * kIn + k] : 0.f;
            F[t * kFRow + k] = f[k];
        }
        float h[kH], g[kH], scratch[kH];
        (void)scratch;
        dense_gelu<kIn>(W + oW1, W + oB1, f, h, DA + t * kRow, Z1 + t * kRow);
        dense_gelu<kH>(W + oW2, W + oB2, h, g, DA + t * kRow, Z2 + t * kRow);
        dense_gelu<kH>(W + oW3, W + oB3, g, h, DA + t * kRow, Z3 + t * kRow);
        float ls = 0.f;
#pragma unroll
        for (int c = 0; c < kOut; ++c) {
            float acc = W[oB4 + c];
#pragma unroll
            for (int k = 0; k < kH; ++k) acc = fmaf(W[oW4 + c * kH + k], h[k], acc);
            const float d = ok ? acc - target[i * kOut + c] : 0.f;
            O[t * 4 + c] = 2.0f * inv_total * d;
            ls += d * d;
        }
        lsum[t] = ls;
    }
    __syncthreads();
    if (t == 0) {
        float ls = 0.f;
        for (int s = 0; s < kS; ++s) ls += lsum[s];
        loss_part[blockIdx.x] = ls;
    }
    // b: layer 4 gradients (thread t <-> input k = t), and delta3
    {
        const int k = t;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int s = 0; s < kS; ++s) {
            const float hk = gelu(Z3[s * kRow + k]);
            a0 = fmaf(O[s * 4 + 0], hk, a0);
            a1 = fmaf(O[s * 4 + 1], hk, a1);
            a2 = fmaf(O[s * 4 + 2], hk, a2);
        }
        P[oW4 + 0 * kH + k] = a0;
        P[oW4 + 1 * kH + k] = a1;
        P[oW4 + 2 * kH + k] = a2;
        if (t < kOut) {
            float b = 0.f;
            for (int s = 0; s < kS; ++s) b += O[s * 4 + t];
            P[oB4 + t] = b;
        }
    }
    if (t < kS) {
        const float d0 = O[t * 4 + 0], d1 = O[t * 4 + 1], d2 = O[t * 4 + 2];
        for (int j = 0; j < kH; ++j) {
            const float v = W[oW4 + j] * d0 + W[oW4 + kH + j] * d1 + W[oW4 + 2 * kH + j] * d2;
            DA[t * kRow + j] = v * gelu_d(Z3[t * kRow + j]);
        }
    }
    __syncthreads();
    // c: delta2 = (W3^T delta3) * gelu'(z2)
    if (t < kS) {
        float d[kH];
#pragma unroll
        for (int k = 0; k < kH; ++k) d[k] = DA[t * kRow + k];
        for (int j = 0; j < kH; ++j) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < kH; ++k) v = fmaf(W[oW3 + k * kH + j], d[k], v);
            DB[t * kRow + j] = v * gelu_d(Z2[t * kRow + j]);
        }
    }
    __syncthreads();
    // d: h2 in place of z2, then layer 3 gradients (thread t <-> output j = t)
    for (int e = t; e < kS * kH; e += kTrain) {
        const int s = e / kH, k = e % kH;
        Z2[s * kRow + k] = gelu(Z2[s * kRow + k]);
    }
    __syncthreads();
    {
        const int j = t;
        float acc[kH];
#pragma unroll
        for (int k = 0; k < kH; ++k) acc[k] = 0.f;
        float bsum = 0.f;
        for (int s = 0; s < kS; ++s) {
            const float dj = DA[s * kRow + j];
            bsum += dj;
#pragma unroll
            for (int k = 0; k < kH; ++k) acc[k] = fmaf(dj, Z2[s * kRow + k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < kH; ++k) P[oW3 + j * kH + k] = acc[k];
        P[oB3 + j] = bsum;
    }
    __syncthreads();
    // e: delta1 = (W2^T delta2) * gelu'(z1), into DA
    if (t < kS) {
        float d[kH];
#pragma unroll
        for (int k = 0; k < kH; ++k) d[k] = DB[t * kRow + k];
        for (int j = 0; j < kH; ++j) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < kH; ++k) v = fmaf(W[oW2 + k * kH + j], d[k], v);
            DA[t * kRow + j] = v * gelu_d(Z1[t * kRow + j]);
        }
    }
    __syncthreads();
    // f: h1 in place of z1, then layer 2 gradients
    for (int e = t; e < kS * kH; e += kTrain) {
        const int s = e / kH, k = e % kH;
        Z1[s * kRow + k] = gelu(Z1[s * kRow + k]);
    }
    __syncthreads();
    {
        const int j = t;
        float acc[kH];
#pragma unroll
        for (int k = 0; k < kH; ++k) acc[k] = 0.f;
        float bsum = 0.f;
        for (int s = 0; s < kS; ++s) {
            const float dj = DB[s * kRow + j];
            bsum += dj;
#pragma unroll
            for (int k = 0; k < kH; ++k) acc[k] = fmaf(dj, Z1[s * kRow + k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < kH; ++k) P[oW2 + j * kH + k] = acc[k];
        P[oB2 + j] = bsum;
    }
    // g: layer 1 gradients from delta1 (DA) and the inputs
    {
        const int j = t;
        float acc[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) acc[k] = 0.f;
        float bsum = 0.f;
        for (int s = 0; s < kS; ++s) {
            const float dj = DA[s * kRow + j];
            bsum += dj;
#pragma unroll
            for (int k = 0; k < kIn; ++k) acc[k] = fmaf(dj, F[s * kFRow + k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < kIn; ++k) P[oW1 + j * kIn + k] = acc[k];
        P[oB1 + j] = bsum;
    }
}

// grad[p] = sum over blocks of part[b][p]; loss = sum of loss_part
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ part,
                                                         uint32_t nb, float* __restrict__ grad,
                                                         const float* __restrict__ loss_part,
                                                         float* __restrict__ loss) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p < PS_MLP_PARAMS) {
        float s = 0.f;
        for (uint32_t b = 0; b < nb; ++b) s += part[(size_t)b * PS_MLP_PARAMS + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        float s = 0.f;
        for (uint32_t b = 0; b < nb; ++b) s += loss_part[b];
        *loss = s;
    }
}

// Adam (Kingma & Ba 2015), bias-corrected
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* __restrict__ W,
                                                       const float* __restrict__ g,
                                                       float* __restrict__ m1,
                                                       float* __restrict__ m2, float lr,
                                                       float b1, float b2, float eps, float c1,
                                                       float c2) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= PS_MLP_PARAMS) return;
    const float gp = g[p];
    const float a = b1 * m1[p] + (1.0f - b1) * gp;
    const float v = b2 * m2[p] + (1.0f - b2) * gp * gp;
    m1[p] = a;
    m2[p] = v;
    W[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

void ps_launch_mlp_features(size_t n, const uint32_t* idx, size_t rows, const float4* pos4,
                            const uint32_t* leaf_of, const float4* geom, const float4* mp,
                            const float4* loc, const uint32_t* parent, float* feat, float* scale,
                            hipStream_t s) {
    if (rows == 0) return;
    mlp_features_kernel<<<(unsigned)((rows + kFwdBlock - 1) / kFwdBlock), kFwdBlock, 0, s>>>(
        n, idx, rows, pos4, leaf_of, geom, mp, loc, parent, feat, scale);
}

void ps_launch_mlp_forward(const float* W, const float* feat, size_t n, float* out,
                           hipStream_t s) {
    if (n == 0) return;
    mlp_forward_kernel<<<(unsigned)((n + kFwdBlock - 1) / kFwdBlock), kFwdBlock, 0, s>>>(
        W, feat, n, out);
}

void ps_launch_mlp_apply(size_t n, const float* W, const float4* pos4, const uint32_t* leaf_of,
                         const float4* geom, const float4* mp, const float4* loc,
                         const uint32_t* parent, float* fx, float* fy, float* fz,
                         hipStream_t s) {
    if (n == 0) return;
    mlp_apply_kernel<<<(unsigned)((n + kFwdBlock - 1) / kFwdBlock), kFwdBlock, 0, s>>>(
        n, W, pos4, leaf_of, geom, mp, loc, parent, fx, fy, fz);
}

size_t ps_mlp_grad_tmp_floats(size_t n) {
    const size_t nb = (n + kS - 1) / kS;
    return nb * (PS_MLP_PARAMS + 1);
}

void ps_launch_mlp_grad(const float* W, const float* feat, const float* target, size_t n,
                        float* grad, float* loss, float* tmp, hipStream_t s) {
    const uint32_t nb = (uint32_t)((n + kS - 1) / kS);
    if (nb == 0) return;
    float* part = tmp;
    float* loss_part = tmp + (size_t)nb * PS_MLP_PARAMS;
    mlp_grad_kernel<<<nb, kTrain, 0, s>>>(W, feat, target, n, 1.0f / (float)n, part, loss_part);
    mlp_reduce_kernel<<<(PS_MLP_PARAMS + 255) / 256, 256, 0, s>>>(part, nb, grad, loss_part,
                                                                  loss);
}

void ps_launch_mlp_adam(float* W, const float* grad, float* m1, float* m2, float lr, float b1,
                        float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, (float)step);
    const float c2 = 1.0f - powf(b2, (float)step);
    mlp_adam_kernel<<<(PS_MLP_PARAMS + 255) / 256, 256, 0, s>>>(W, grad, m1, m2, lr, b1, b2, eps,
                                                                c1, c2);
}
