// ps_p2p.hip -- near field (P2P) fused with L2P (replaces the leaf x leaf
// loop of H:1279-1452 and L2P of H:1158-1235).
//
// One thread per target particle, particles in sorted order: the lanes of a
// wave are spatially adjacent, and in dense leaves they share one leaf and so
// the same source list -- every source load is the same address across the
// wave.  A leaf's sources are the near (mask bit clear) entries of its final
// traversal list, read in place from the list arena: near leaves only, self
// included (the self term is zero: d = 0).  Output per particle, one write:
//   f = m * field(leaf) + m * sum_j m_j d / (|d|^2 + eps^2)^(3/2)
// FP32 with v_rsq_f32; optional Kahan-compensated accumulation.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;

template <int kMode, bool kKahan>
__global__ __launch_bounds__(kBlock) void p2p_kernel(
    size_t n, const float4* __restrict__ pos4, const uint32_t* __restrict__ leaf_of,
    const float4* __restrict__ loc, const uint32_t* __restrict__ p2p_off,
    const uint32_t* __restrict__ p2p_len, const uint32_t* __restrict__ p2p_moff,
    const uint32_t* __restrict__ arena, const unsigned long long* __restrict__ masks,
    const uint2* __restrict__ link, const uint32_t* __restrict__ nend, float* __restrict__ fx,
    float* __restrict__ fy, float* __restrict__ fz, unsigned long long* block_pairs,
    uint32_t nblocks) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    float ax = 0.f, ay = 0.f, az = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t pairs = 0, nsrc = 0;  // particle pairs; near source leaves
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t lf = 0;
    if (valid) {
        lf = leaf_of[i];
        p = pos4[i];
        const uint32_t L = (kMode != 1) ? p2p_len[lf] : 0u;
        if (L) {
            const uint32_t o = p2p_off[lf];
            const uint32_t mo = p2p_moff[lf];
            for (uint32_t k0 = 0; k0 < L; k0 += 64) {
                unsigned long long near = ~masks[mo + (k0 >> 6)];
                if (L - k0 < 64) near &= (1ull << (L - k0)) - 1ull;
                while (near) {
                    const uint32_t b = (uint32_t)__ffsll((long long)near) - 1u;
                    near &= near - 1ull;
                    ++nsrc;
                    const uint32_t s = arena[o + k0 + b];
                    const uint32_t j0 = link[s].x;
                    const uint32_t j1e = nend[s];
                    const uint32_t j1 = j1e < (uint32_t)n ? j1e : (uint32_t)n;
                    pairs += j1e - j0;
                    for (uint32_t j = j0; j < j1; ++j) {
                        const float4 q = pos4[j];
                        if (kKahan) {
                            float tx = 0.f, ty = 0.f, tz = 0.f;
                            ps_pair(p, q, tx, ty, tz);
                            ps_kahan_add(ax, cx, tx);
                            ps_kahan_add(ay, cy, ty);
                            ps_kahan_add(az, cz, tz);
                        } else {
                            ps_pair(p, q, ax, ay, az);
                        }
                    }
                }
            }
        }
    }
    if (block_pairs) {  // per-block partial sums, reduced by another kernel
        // leaf pairs: the first particle of each leaf counts its near sources
        if (valid && i != link[lf].x) nsrc = 0;
        __shared__ unsigned long long wsum[2][kBlock / 64];
        unsigned long long v = pairs, u = nsrc;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            v += __shfl_down(v, d, 64);
            u += __shfl_down(u, d, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            wsum[0][threadIdx.x >> 6] = v;
            wsum[1][threadIdx.x >> 6] = u;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long b = 0, c = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                b += wsum[0][w];
                c += wsum[1][w];
            }
            block_pairs[blockIdx.x] = b;
            block_pairs[nblocks + blockIdx.x] = c;
        }
    }
    if (!valid) return;
    {
#pragma clang fp contract(off)
        if (kMode == 0) {  // L2P (first write, H:1230-1232) then the P2P sum
            const float4 fl = loc[lf];
            fx[i] = p.w * fl.y + p.w * ax;
            fy[i] = p.w * fl.z + p.w * ay;
            fz[i] = p.w * fl.w + p.w * az;
        } else if (kMode == 1) {
            const float4 fl = loc[lf];
            fx[i] = p.w * fl.y;
            fy[i] = p.w * fl.z;
            fz[i] = p.w * fl.w;
        } else {
            fx[i] = fx[i] + p.w * ax;
            fy[i] = fy[i] + p.w * ay;
            fz[i] = fz[i] + p.w * az;
        }
    }
}

template <int kMode>
void launch_mode(size_t n, const float4* pos4, const uint32_t* leaf_of, const float4* loc,
                 const uint32_t* p2p_off, const uint32_t* p2p_len, const uint32_t* p2p_moff,
                 const uint32_t* arena, const unsigned long long* masks, const uint2* link,
                 const uint32_t* nend, float* fx, float* fy, float* fz, int kahan,
                 unsigned long long* bp, hipStream_t s) {
    unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    if (kahan)
        p2p_kernel<kMode, true><<<blocks, kBlock, 0, s>>>(n, pos4, leaf_of, loc, p2p_off,
                                                          p2p_len, p2p_moff, arena, masks, link,
                                                          nend, fx, fy, fz, bp, blocks);
    else
        p2p_kernel<kMode, false><<<blocks, kBlock, 0, s>>>(n, pos4, leaf_of, loc, p2p_off,
                                                           p2p_len, p2p_moff, arena, masks, link,
                                                           nend, fx, fy, fz, bp, blocks);
}

}  // namespace

uint32_t ps_p2p_blocks(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

void ps_launch_p2p(size_t n, const float4* pos4, const uint32_t* leaf_of, const float4* loc,
                   const uint32_t* p2p_off, const uint32_t* p2p_len, const uint32_t* p2p_moff,
                   const uint32_t* arena, const unsigned long long* masks, const uint2* link,
                   const uint32_t* nend, float* fx, float* fy, float* fz, int mode, int kahan,
                   unsigned long long* block_pairs, unsigned long long* pair_count,
                   hipStream_t s) {
    if (n == 0) return;
    unsigned long long* bp = pair_count ? block_pairs : nullptr;
    if (mode == 0)
        launch_mode<0>(n, pos4, leaf_of, loc, p2p_off, p2p_len, p2p_moff, arena, masks, link,
                       nend, fx, fy, fz, kahan, bp, s);
    else if (mode == 1)
        launch_mode<1>(n, pos4, leaf_of, loc, p2p_off, p2p_len, p2p_moff, arena, masks, link,
                       nend, fx, fy, fz, kahan, bp, s);
    else
        launch_mode<2>(n, pos4, leaf_of, loc, p2p_off, p2p_len, p2p_moff, arena, masks, link,
                       nend, fx, fy, fz, kahan, bp, s);
    if (bp) {
        const uint32_t nb = ps_p2p_blocks(n);
        ps_reduce_u64(bp, nb, pair_count, s);
        ps_reduce_u64(bp + nb, nb, pair_count + 1, s);
    }
}
