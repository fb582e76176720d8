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
//   net: 24 -> 64 -> 64 -> 64 -> 3, GELU (tanh form) on the hidden layers,
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

// GELU in its tanh form (Hendrycks & Gimpel 2016):
//   gelu(x) = x/2 (1 + tanh(u)) = x sigma(2u),  u = c (x + a x^3),
//   c = sqrt(2/pi), a = 0.044715,
//   gelu'(x) = sigma + x sigma (1 - sigma) 2c (1 + 3a x^2).
// sigma(2u) = 1 / (1 + 2^(-2u log2 e)) with v_exp_f32 and v_rcp_f32: a large
// negative x gives 2^(+big) = inf and sigma = 0, a large positive one sigma = 1.
constexpr float kGeluA = 0.044715f;
constexpr float kGelu2C = 1.5957691216057308f;           // 2 sqrt(2/pi)
constexpr float kGeluE0 = -2.3022081983492596f;          // -2 sqrt(2/pi) log2(e)
constexpr float kGeluE1 = kGeluE0 * kGeluA;

__device__ __forceinline__ float gelu_sigma(float x) {
    const float z = x * fmaf(x * x, kGeluE1, kGeluE0);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sigma(x); }
__device__ __forceinline__ float gelu_d(float x) {
    const float sg = gelu_sigma(x);
    return fmaf(x * sg * (1.0f - sg), kGelu2C * fmaf(3.0f * kGeluA, x * x, 1.0f), sg);
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
    const bool hasP = P != PS_NONE;
    const uint32_t Pi = hasP ? P : L;  // always a valid node: no branches around the loads
    const uint32_t G = hasP ? parent[Pi] : PS_NONE;
    const bool hasG = G != PS_NONE;
    const uint32_t Gi = hasG ? G : Pi;
    const float4 gL = geom[L], lL = loc[L], mL = mp[L];
    const float4 gP = geom[Pi], gG = geom[Gi];
    float4 lP = loc[Pi], lG = loc[Gi];
    if (!hasP) lP = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!hasG) lG = make_float4(0.f, 0.f, 0.f, 0.f);
    // v_sqrt_f32 / v_rcp_f32 (within 1 ulp): the same function computes the
    // training features, so inference sees exactly what training saw
    const float s = __builtin_amdgcn_sqrtf(lL.y * lL.y + lL.z * lL.z + lL.w * lL.w);
    const float inv = s > 0.f ? __builtin_amdgcn_rcpf(s) : 0.f;
    const float hl = __builtin_amdgcn_rcpf(gL.w), hp = __builtin_amdgcn_rcpf(gP.w),
                hg = __builtin_amdgcn_rcpf(gG.w);
    f[0] = (p.x - gL.x) * hl;
    f[1] = (p.y - gL.y) * hl;
    f[2] = (p.z - gL.z) * hl;
    f[3] = (lL.y - lP.y) * inv;
    f[4] = (lL.z - lP.z) * inv;
    f[5] = (lL.w - lP.w) * inv;
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
    const float cm = mL.x != 0.f ? hl * __builtin_amdgcn_rcpf(mL.x) : 0.f;
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

// ---------------------------------------------------------------------------
// The corrected tick's last step, f_i += m_i * s_i * net(F_i), on the matrix
// cores.  Each layer is D[out][particle] = W[out][in] x H[in][particle] with
// v_mfma_f32_16x16x32_bf16: a wavefront takes 16 particles (the 16 columns),
// 64 outputs are 4 row tiles, an input dimension of 64 two K blocks of 32
// (the 24 features: one block, zero padded).  FP32 weights and activations
// enter as bf16 pairs hi + lo (hi = bf16(v), lo = bf16(v - hi): 16 mantissa
// bits); each product is hi*hi + hi*lo + lo*hi (three MFMAs, the dropped
// lo*lo is below 2^-16 relative), accumulated in FP32.  The bias is the
// accumulator's initial value.
//
// Lane l holds row i = l % 16 of A and column j = l % 16 of B, both with the
// eight K elements of group g = l / 16; D element e of lane l is row 4g + e,
// column j.  Element e of K block kb in lane group g stands for hidden unit
// 16 (2 kb + e / 4) + 4g + e % 4 -- exactly the D element (tile 2 kb + e / 4,
// element e % 4) the same lane holds after the previous layer -- so the
// activations pass from one layer's accumulators to the next layer's B
// operand without leaving the lane; the packed weights follow the same
// order.  For the first layer element e of group g is feature 8g + e.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFrags = 22;      // (layer, row tile, K block): 4 + 8 + 8 + 2
constexpr int kBiasTiles = 13;  // (layer, row tile): 4 + 4 + 4 + 1
constexpr int kApplyBlock = 1024;
static_assert(PS_MLP_PACKED_BYTES == kFrags * 2 * 64 * 16 + kBiasTiles * 64 * 16, "packed size");

// hidden unit of element e of K block kb in lane group g (layers 2 to 4)
__device__ __forceinline__ int mlp_hidden_k(int kb, int g, int e) {
    return 16 * (2 * kb + (e >> 2)) + 4 * g + (e & 3);
}

// packed layout: bf16 fragments [frag][hi, lo][lane][8], then float4 biases
// [tile][lane]; one thread per element
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ W,
                                                       uint32_t* __restrict__ pk) {
    const int tid = blockIdx.x * 256 + threadIdx.x;
    constexpr int kFragElems = kFrags * 64 * 8;
    if (tid < kFragElems) {
        const int e = tid & 7, l = (tid >> 3) & 63, f = tid >> 9;
        const int i = l & 15, g = l >> 4;
        float w = 0.f;
        if (f < 4) {  // layer 1: row tile f, features 8g + e
            const int o = 16 * f + i, k = 8 * g + e;
            if (k < kIn) w = W[oW1 + o * kIn + k];
        } else if (f < 20) {  // layers 2 and 3: row tile t, K block kb
            const int q = f < 12 ? f - 4 : f - 12;
            const int o = 16 * (q >> 1) + i, k = mlp_hidden_k(q & 1, g, e);
            w = W[(f < 12 ? oW2 : oW3) + o * kH + k];
        } else {  // layer 4: rows 0..2 of one tile
            if (i < kOut) w = W[oW4 + i * kH + mlp_hidden_k(f - 20, g, e)];
        }
        const __bf16 hi = (__bf16)w;
        const __bf16 lo = (__bf16)(w - (float)hi);
        __bf16* b = reinterpret_cast<__bf16*>(pk);
        b[((f * 2 + 0) * 64 + l) * 8 + e] = hi;
        b[((f * 2 + 1) * 64 + l) * 8 + e] = lo;
    } else if (tid < kFragElems + kBiasTiles * 64 * 4) {
        const int r = tid - kFragElems;
        const int e = r & 3, l = (r >> 2) & 63, T = r >> 8;
        const int o = 16 * (T & 3) + 4 * (l >> 4) + e;
        float b = 0.f;
        if (T < 4) b = W[oB1 + o];
        else if (T < 8) b = W[oB2 + o];
        else if (T < 12) b = W[oB3 + o];
        else if (o < kOut) b = W[oB4 + o];
        // the fragments take kFragElems words (a hi and a lo bf16 per element)
        reinterpret_cast<float*>(pk)[kFragElems + r] = b;
    }
}

__device__ __forceinline__ void split_bf16(const float v[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)v[e];
        hi[e] = h;
        lo[e] = (__bf16)(v[e] - (float)h);
    }
}

// acc += A(frag) x (bh + bl), A = ah + al: three MFMAs
__device__ __forceinline__ f32x4 mfma3(const uint4* __restrict__ sfr, int f, int lane, bf16x8 bh,
                                       bf16x8 bl, f32x4 acc) {
    const uint4 uh = sfr[(f * 2 + 0) * 64 + lane];
    const uint4 ul = sfr[(f * 2 + 1) * 64 + lane];
    const bf16x8 ah = __builtin_bit_cast(bf16x8, uh);
    const bf16x8 al = __builtin_bit_cast(bf16x8, ul);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    return acc;
}

__device__ __forceinline__ f32x4 bias4(const float4* __restrict__ sbi, int T, int lane) {
    const float4 b = sbi[T * 64 + lane];
    f32x4 r;
    r[0] = b.x;
    r[1] = b.y;
    r[2] = b.z;
    r[3] = b.w;
    return r;
}

// GELU of the four row tiles, as the next layer's two K blocks (hi, lo)
__device__ __forceinline__ void hidden_operands(const f32x4 acc[4], bf16x8 hh[2], bf16x8 hl[2]) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = gelu(acc[2 * kb + (e >> 2)][e & 3]);
        split_bf16(v, hh[kb], hl[kb]);
    }
}

// One wavefront per 64 particles at a time: lane l computes the features of
// particle p0 + l into the wave's LDS rows (all 64 lanes busy), then the wave
// runs the net on the four 16-particle column tiles, reading each lane's B
// operand (eight features of one particle) from those rows.  The rows are
// private to the wave, so a wavefront-scope fence and barrier order the
// writes before the reads (and the reads before the next round's writes).
constexpr int kApplyWaves = kApplyBlock / 64;


__global__ __launch_bounds__(kApplyBlock) void mlp_apply_kernel(
    size_t n, const uint4* __restrict__ pk, const float4* __restrict__ pos4,
    const uint32_t* __restrict__ leaf_of, const float4* __restrict__ geom,
    const float4* __restrict__ mp, const float4* __restrict__ loc,
    const uint32_t* __restrict__ parent, float* __restrict__ fx, float* __restrict__ fy,
    float* __restrict__ fz) {
    __shared__ uint4 sfr[kFrags * 2 * 64];
    __shared__ float4 sbi[kBiasTiles * 64];
    __shared__ float4 sft[kApplyWaves][64 * (kIn / 4)];  // features, 6 float4 per particle
    __shared__ float ssc[kApplyWaves][64];                // s = |E_L| per particle
    for (int k = threadIdx.x; k < kFrags * 2 * 64; k += kApplyBlock) sfr[k] = pk[k];
    const float4* pb = reinterpret_cast<const float4*>(pk + kFrags * 2 * 64);
    for (int k = threadIdx.x; k < kBiasTiles * 64; k += kApplyBlock) sbi[k] = pb[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    float4* ft = sft[wave];
    float* sc = ssc[wave];
    const size_t rounds = (n + 63) / 64;
    const size_t nw = (size_t)gridDim.x * kApplyWaves;
    // wave-uniform loop: every lane is active at each MFMA
    for (size_t r = (size_t)blockIdx.x * kApplyWaves + wave; r < rounds; r += nw) {
        const size_t p0 = r * 64;
        {
            const size_t i = p0 + lane;
            float f[kIn];
            sc[lane] = mlp_features(i < n ? i : n - 1, pos4, leaf_of, geom, mp, loc, parent, f);
#pragma unroll
            for (int q = 0; q < kIn / 4; ++q)
                ft[lane * (kIn / 4) + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2],
                                                       f[4 * q + 3]);
        }
        wave_sync();
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            const int row = 16 * c + j;
            float v[8];
            if (g < 3) {
                const float4 a = ft[row * (kIn / 4) + 2 * g];
                const float4 b = ft[row * (kIn / 4) + 2 * g + 1];
                v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
                v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = 0.f;
            }
            bf16x8 bh, bl;
            split_bf16(v, bh, bl);
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma3(sfr, t, lane, bh, bl, bias4(sbi, t, lane));
#pragma unroll 1
            for (int layer = 0; layer < 2; ++layer) {  // layers 2 and 3
                bf16x8 hh[2], hl[2];
                hidden_operands(acc, hh, hl);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    f32x4 a = bias4(sbi, 4 + 4 * layer + t, lane);
                    a = mfma3(sfr, 4 + 8 * layer + 2 * t + 0, lane, hh[0], hl[0], a);
                    a = mfma3(sfr, 4 + 8 * layer + 2 * t + 1, lane, hh[1], hl[1], a);
                    acc[t] = a;
                }
            }
            bf16x8 hh[2], hl[2];
            hidden_operands(acc, hh, hl);
            f32x4 o = bias4(sbi, 12, lane);
            o = mfma3(sfr, 20, lane, hh[0], hl[0], o);
            o = mfma3(sfr, 21, lane, hh[1], hl[1], o);
            // lanes of group 0 hold the three outputs of particle p0 + row
            const size_t i = p0 + row;
            const float s = sc[row];
            if (g == 0 && i < n && s != 0.f) {
                const float w = pos4[i].w * s;
                fx[i] = fmaf(w, o[0], fx[i]);
                fy[i] = fmaf(w, o[1], fy[i]);
                fz[i] = fmaf(w, o[2], fz[i]);
            }
        }
        wave_sync();  // the rows are rewritten by the next round
    }
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

void ps_launch_mlp_pack(const float* W, void* packed, hipStream_t s) {
    constexpr int total = kFrags * 64 * 8 + kBiasTiles * 64 * 4;
    mlp_pack_kernel<<<(total + 255) / 256, 256, 0, s>>>(W, reinterpret_cast<uint32_t*>(packed));
}

void ps_launch_mlp_apply(size_t n, const void* packed, const float4* pos4, const uint32_t* leaf_of,
                         const float4* geom, const float4* mp, const float4* loc,
                         const uint32_t* parent, float* fx, float* fy, float* fz,
                         hipStream_t s) {
    if (n == 0) return;
    const size_t rounds = (n + 63) / 64;
    size_t blocks = (rounds + kApplyWaves - 1) / kApplyWaves;
    if (blocks > 512) blocks = 512;  // one block per CU (LDS), grid-stride beyond
    mlp_apply_kernel<<<(unsigned)blocks, kApplyBlock, 0, s>>>(
        n, reinterpret_cast<const uint4*>(packed), pos4, leaf_of, geom, mp, loc, parent, fx, fy,
        fz);
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
