// ps_p2p.hip -- near field (P2P) fused with L2P (replaces the leaf x leaf
// loop of H:1279-1452 and L2P of H:1158-1235).
//
// Two kernels split the target leaves.  A leaf of >= tmin particles gets a
// wavefront: its 64-target rounds share the leaf's source list, so the
// loop runs on scalar control flow and every source particle is a scalar load
// broadcast to the wave.  Every other particle gets a thread; particles are in
// sorted order, so the lanes of a leaf walk the same list in step and their
// source loads coalesce.  A leaf's sources are the near (mask bit clear)
// entries of its final traversal list, read in place from the list arena:
// near leaves only, self included (the self term is zero: d = 0).
// Output per particle, one write:
//   f = m * field(leaf) + m * sum_j m_j d / (|d|^2 + eps^2)^(3/2)
// FP32 with v_rsq_f32; optional Kahan-compensated accumulation.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;

// one pair into target p's sum (plain or compensated)
template <bool kKahan>
__device__ __forceinline__ void p2p_acc(float4 p, float4 q, float& ax, float& ay, float& az,
                                        float& cx, float& cy, float& cz) {
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

// accumulate T targets per lane (p[0..T)) with the particles [j0, j1) of one
// source leaf; j0, j1 wave-uniform, so each source particle is a scalar load
// shared by the wave and feeds all T; each target sums in particle order
template <bool kKahan, int T>
__device__ __forceinline__ void p2p_uniform_range_t(const float4* p,
                                                    const float4* __restrict__ pos4, uint32_t j0,
                                                    uint32_t j1, float* a, float* c) {
    // unrolled so that about four source loads are in flight per round
    constexpr int kUnroll = T >= 4 ? 1 : 4 / T;
#pragma unroll kUnroll
    for (uint32_t j = j0; j < j1; ++j) {
        const float4 q = pos4[j];
#pragma unroll
        for (int u = 0; u < T; ++u) {
            if (kKahan) {
                float tx = 0.f, ty = 0.f, tz = 0.f;
                ps_pair(p[u], q, tx, ty, tz);
                ps_kahan_add(a[3 * u + 0], c[3 * u + 0], tx);
                ps_kahan_add(a[3 * u + 1], c[3 * u + 1], ty);
                ps_kahan_add(a[3 * u + 2], c[3 * u + 2], tz);
            } else {
                ps_pair(p[u], q, a[3 * u + 0], a[3 * u + 1], a[3 * u + 2]);
            }
        }
    }
}

// the big-leaf kernel's targets: leaves of >= tmin particles (tmin > 0)
__device__ __forceinline__ bool p2p_is_big(uint32_t lf, uint32_t tmin,
                                           const uint2* __restrict__ link,
                                           const uint32_t* __restrict__ nend) {
    return tmin && nend[lf] - link[lf].x >= tmin;
}

// accumulate T targets per lane against the near sources of one final list
// (wave-uniform: scalar loads); pairs += particle pairs, nsrc += source leaves
template <bool kKahan, int T>
__device__ __forceinline__ void p2p_uniform_list_t(const float4* p, size_t n, uint32_t L,
                                                   uint32_t o, uint32_t mo,
                                                   const float4* __restrict__ pos4,
                                                   const uint32_t* __restrict__ arena,
                                                   const unsigned long long* __restrict__ masks,
                                                   const uint2* __restrict__ link,
                                                   const uint32_t* __restrict__ nend, float* a,
                                                   float* c, uint32_t& pairs, uint32_t& nsrc) {
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
            p2p_uniform_range_t<kKahan, T>(p, pos4, j0, j1, a, c);
        }
    }
}

// one round of T x 64 targets from c0 (targets c0 + 64 u + lane, u < T): the
// leaf's whole near field, then the stores
template <int kMode, bool kKahan, int T>
__device__ __forceinline__ void p2p_big_round(uint32_t c0, uint32_t te, uint32_t lane, size_t n,
                                              uint32_t L, uint32_t o, uint32_t mo, float4 fl,
                                              const float4* __restrict__ pos4,
                                              const uint32_t* __restrict__ arena,
                                              const unsigned long long* __restrict__ masks,
                                              const uint2* __restrict__ link,
                                              const uint32_t* __restrict__ nend,
                                              float* __restrict__ fx, float* __restrict__ fy,
                                              float* __restrict__ fz, uint32_t& pairs,
                                              uint32_t& ws);

// f for particle i from its leaf's field and its P2P sum (mode 0: both;
// 1: field only; 2: adds the sum to f)
template <int kMode>
__device__ __forceinline__ void p2p_store(size_t i, float4 p, float4 fl, float ax, float ay,
                                          float az, float* __restrict__ fx,
                                          float* __restrict__ fy, float* __restrict__ fz) {
#pragma clang fp contract(off)
    if (kMode == 0) {  // L2P (first write, H:1230-1232) then the P2P sum
        fx[i] = p.w * fl.y + p.w * ax;
        fy[i] = p.w * fl.z + p.w * ay;
        fz[i] = p.w * fl.w + p.w * az;
    } else if (kMode == 1) {
        fx[i] = p.w * fl.y;
        fy[i] = p.w * fl.z;
        fz[i] = p.w * fl.w;
    } else {
        fx[i] = fx[i] + p.w * ax;
        fy[i] = fy[i] + p.w * ay;
        fz[i] = fz[i] + p.w * az;
    }
}

// per-block partial sums of the stats, reduced by another kernel
__device__ __forceinline__ void p2p_block_stats(uint32_t pairs, uint32_t nsrc,
                                                unsigned long long* block_pairs, uint32_t slot,
                                                uint32_t stride) {
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
        block_pairs[slot] = b;
        block_pairs[stride + slot] = c;
    }
}

// One thread per target particle of a small leaf (particles of big leaves
// are left to p2p_big_kernel).
template <int kMode, bool kKahan>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void p2p_kernel(
    size_t n, const float4* __restrict__ pos4, const uint32_t* __restrict__ leaf_of,
    const float4* __restrict__ loc, const uint32_t* __restrict__ p2p_off,
    const uint32_t* __restrict__ p2p_len, const uint32_t* __restrict__ p2p_moff,
    const uint32_t* __restrict__ arena, const unsigned long long* __restrict__ masks,
    const uint2* __restrict__ link, const uint32_t* __restrict__ nend, float* __restrict__ fx,
    float* __restrict__ fy, float* __restrict__ fz, uint32_t tmin,
    unsigned long long* block_pairs, uint32_t stride) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    float ax = 0.f, ay = 0.f, az = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t pairs = 0, nsrc = 0;  // particle pairs; near source leaves
    uint32_t lf = 0;
    bool valid = i < n;
    if (valid) {
        lf = leaf_of[i];
        valid = !p2p_is_big(lf, tmin, link, nend);
    }
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        p = pos4[i];
        const uint32_t L = (kMode != 1) ? p2p_len[lf] : 0u;
        if (L) {
            const uint32_t o = p2p_off[lf];
            const uint32_t mo = p2p_moff[lf];
            for (uint32_t k0 = 0; k0 < L; k0 += 64) {
                unsigned long long near = ~masks[mo + (k0 >> 6)];
                if (L - k0 < 64) near &= (1ull << (L - k0)) - 1ull;
                // near sources eight at a time: the eight list words, then the
                // eight ranges, are requested together (one round trip each
                // instead of eight); the sums still run source by source in
                // list order
                while (near) {
                    constexpr int kB = 8;
                    uint32_t sb[kB];
                    int cnt = 0;
#pragma unroll
                    for (int u = 0; u < kB; ++u) {
                        sb[u] = 0u;
                        if (near) {
                            sb[u] = (uint32_t)__ffsll((long long)near) - 1u;
                            near &= near - 1ull;
                            ++cnt;
                        }
                    }
                    uint32_t src[kB];
#pragma unroll
                    for (int u = 0; u < kB; ++u) src[u] = u < cnt ? arena[o + k0 + sb[u]] : 0u;
                    uint32_t r0[kB], r1[kB];
#pragma unroll
                    for (int u = 0; u < kB; ++u) {
                        r0[u] = u < cnt ? link[src[u]].x : 0u;
                        r1[u] = u < cnt ? nend[src[u]] : 0u;
                    }
                    // the first two particles of each source are loaded while
                    // the previous source is summed (one source ahead); the
                    // loads are unconditional (index 0 past cnt, the first
                    // particle again for a source of one), their use is not
                    auto second = [&](int u) {
                        const uint32_t e = r1[u] < (uint32_t)n ? r1[u] : (uint32_t)n;
                        return r0[u] + 1u < e ? r0[u] + 1u : r0[u];
                    };
                    float4 qn = pos4[r0[0]];
                    float4 qn2 = pos4[second(0)];
#pragma unroll
                    for (int u = 0; u < kB; ++u) {
                        if (u >= cnt) break;
                        ++nsrc;
                        const uint32_t j0 = r0[u];
                        const uint32_t j1e = r1[u];
                        const uint32_t j1 = j1e < (uint32_t)n ? j1e : (uint32_t)n;
                        pairs += j1e - j0;
                        const float4 qa = qn, qb = qn2;
                        if (u + 1 < kB) {
                            qn = pos4[r0[u + 1]];
                            qn2 = pos4[second(u + 1)];
                        }
                        if (j0 < j1) p2p_acc<kKahan>(p, qa, ax, ay, az, cx, cy, cz);
                        if (j0 + 1u < j1) p2p_acc<kKahan>(p, qb, ax, ay, az, cx, cy, cz);
                        uint32_t j = j0 + 2;
                        for (; j + 2 <= j1; j += 2) {  // two loads in flight
                            const float4 q0 = pos4[j], q1 = pos4[j + 1];
                            p2p_acc<kKahan>(p, q0, ax, ay, az, cx, cy, cz);
                            p2p_acc<kKahan>(p, q1, ax, ay, az, cx, cy, cz);
                        }
                        for (; j < j1; ++j) p2p_acc<kKahan>(p, pos4[j], ax, ay, az, cx, cy, cz);
                    }
                }
            }
        }
        // leaf pairs: the first particle of each leaf counts its near sources
        if (i != link[lf].x) nsrc = 0;
    }
    if (block_pairs) p2p_block_stats(pairs, nsrc, block_pairs, blockIdx.x, stride);
    if (!valid) return;
    if (kKahan) {
        ax -= cx;
        ay -= cy;
        az -= cz;
    }
    p2p_store<kMode>(i, p, (kMode != 2) ? loc[lf] : make_float4(0.f, 0.f, 0.f, 0.f), ax, ay, az,
                     fx, fy, fz);
}

// List of the big leaves: a block scan over the depth-10 nodes, one atomic
// add per block for its slice of the list (big[0] = count, ids from big[1]).
// The order of the slices does not matter: every leaf is computed alone.
constexpr int kCollect = 1024;

__global__ __launch_bounds__(kCollect) void p2p_collect_kernel(const uint32_t* __restrict__ ls,
                                                               uint32_t tmin,
                                                               const uint2* __restrict__ link,
                                                               const uint32_t* __restrict__ nend,
                                                               uint32_t* __restrict__ big) {
    __shared__ uint32_t wcnt[kCollect / 64];
    __shared__ uint32_t base;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t lend = ls[PS_MAX_DEPTH + 1];  // node count
    for (uint32_t n0 = blockIdx.x * kCollect; n0 < lend; n0 += gridDim.x * kCollect) {
        const uint32_t node = n0 + threadIdx.x;
        const bool take = node < lend && (link[node].y & PS_LEAF) &&
                          nend[node] - link[node].x >= tmin;
        const unsigned long long b = __ballot(take);
        if (lane == 0) wcnt[wave] = (uint32_t)__popcll(b);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < kCollect / 64; ++w) {
                const uint32_t c = wcnt[w];
                wcnt[w] = t;
                t += c;
            }
            base = t ? atomicAdd(&big[0], t) : 0u;
        }
        __syncthreads();
        if (take) big[2 + base + wcnt[wave] + (uint32_t)__popcll(b & ps_lanemask_lt())] = node;
        __syncthreads();  // wcnt and base are rewritten by the next round
    }
}

template <int kMode, bool kKahan, int T>
__device__ __forceinline__ void p2p_big_round(uint32_t c0, uint32_t te, uint32_t lane, size_t n,
                                              uint32_t L, uint32_t o, uint32_t mo, float4 fl,
                                              const float4* __restrict__ pos4,
                                              const uint32_t* __restrict__ arena,
                                              const unsigned long long* __restrict__ masks,
                                              const uint2* __restrict__ link,
                                              const uint32_t* __restrict__ nend,
                                              float* __restrict__ fx, float* __restrict__ fy,
                                              float* __restrict__ fz, uint32_t& pairs,
                                              uint32_t& ws) {
    float4 p[T];
    bool valid[T];
#pragma unroll
    for (int u = 0; u < T; ++u) {
        const uint32_t i = c0 + 64 * u + lane;
        valid[u] = i < te;
        p[u] = valid[u] ? pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    float a[3 * T], c[3 * T];
#pragma unroll
    for (int k = 0; k < 3 * T; ++k) {
        a[k] = 0.f;
        c[k] = 0.f;
    }
    uint32_t wp = 0;
    p2p_uniform_list_t<kKahan, T>(p, n, L, o, mo, pos4, arena, masks, link, nend, a, c, wp, ws);
#pragma unroll
    for (int u = 0; u < T; ++u) {
        if (!valid[u]) continue;
        pairs += wp;
        float ax = a[3 * u], ay = a[3 * u + 1], az = a[3 * u + 2];
        if (kKahan) {
            ax -= c[3 * u];
            ay -= c[3 * u + 1];
            az -= c[3 * u + 2];
        }
        p2p_store<kMode>(c0 + 64 * u + lane, p[u], fl, ax, ay, az, fx, fy, fz);
    }
}

// The leaf's last r < 64 targets, k = 2^lk lanes each (the largest power of
// two with r k <= 64): lane t k + q takes target c0 + t and every k-th
// source particle from the q-th; a butterfly over the k lanes adds the
// partial sums.  More lanes are busy than with one target per lane (r / 64
// of them), at the price of a vector load per source step (k addresses).
template <int kMode, bool kKahan>
__device__ __forceinline__ void p2p_split_round(uint32_t c0, uint32_t te, uint32_t lane, size_t n,
                                                uint32_t L, uint32_t o, uint32_t mo, float4 fl,
                                                const float4* __restrict__ pos4,
                                                const uint32_t* __restrict__ arena,
                                                const unsigned long long* __restrict__ masks,
                                                const uint2* __restrict__ link,
                                                const uint32_t* __restrict__ nend,
                                                float* __restrict__ fx, float* __restrict__ fy,
                                                float* __restrict__ fz, uint32_t& pairs,
                                                uint32_t& ws, uint32_t lk) {
    const uint32_t k = 1u << lk;
    const uint32_t q = lane & (k - 1u);
    const uint32_t i = c0 + (lane >> lk);
    const bool valid = i < te;
    const float4 p = valid ? pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float ax = 0.f, ay = 0.f, az = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;  // compensations (Kahan)
    uint32_t wp = 0;
    for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        unsigned long long near = ~masks[mo + (k0 >> 6)];
        if (L - k0 < 64) near &= (1ull << (L - k0)) - 1ull;
        while (near) {
            const uint32_t b = (uint32_t)__ffsll((long long)near) - 1u;
            near &= near - 1ull;
            ++ws;
            const uint32_t s = arena[o + k0 + b];
            const uint32_t j0 = link[s].x;
            const uint32_t j1e = nend[s];
            const uint32_t j1 = j1e < (uint32_t)n ? j1e : (uint32_t)n;
            wp += j1e - j0;
            uint32_t j = j0 + q;
            for (; j + 3u * k < j1; j += 4u * k) {  // four loads in flight per lane
                const float4 q0 = pos4[j], q1 = pos4[j + k], q2 = pos4[j + 2u * k],
                             q3 = pos4[j + 3u * k];
                p2p_acc<kKahan>(p, q0, ax, ay, az, cx, cy, cz);
                p2p_acc<kKahan>(p, q1, ax, ay, az, cx, cy, cz);
                p2p_acc<kKahan>(p, q2, ax, ay, az, cx, cy, cz);
                p2p_acc<kKahan>(p, q3, ax, ay, az, cx, cy, cz);
            }
            for (; j < j1; j += k) p2p_acc<kKahan>(p, pos4[j], ax, ay, az, cx, cy, cz);
        }
    }
    for (uint32_t d = 1; d < k; d <<= 1) {
        if (kKahan) {  // compensated sums merged with their compensations
            ps_kahan_merge(ax, cx, __shfl_xor(ax, (int)d, 64), __shfl_xor(cx, (int)d, 64));
            ps_kahan_merge(ay, cy, __shfl_xor(ay, (int)d, 64), __shfl_xor(cy, (int)d, 64));
            ps_kahan_merge(az, cz, __shfl_xor(az, (int)d, 64), __shfl_xor(cz, (int)d, 64));
        } else {
            ax += __shfl_xor(ax, (int)d, 64);
            ay += __shfl_xor(ay, (int)d, 64);
            az += __shfl_xor(az, (int)d, 64);
        }
    }
    if (kKahan) {
        ax -= cx;
        ay -= cy;
        az -= cz;
    }
    if (valid && q == 0) {
        pairs += wp;
        p2p_store<kMode>(i, p, fl, ax, ay, az, fx, fy, fz);
    }
}

// One wavefront per big leaf (>= tmin particles), 64 targets per
// round: the whole wave shares the leaf's source list, so the loop runs on
// scalar control flow and each source particle is one scalar load broadcast
// to the 64 lanes.  The grid is fixed (kBigBlocks); wave k starts with
// leaf k of the list and then takes the next one from a counter (big[1])
// when it is done, so the uneven leaves balance without launching a wave
// per possible leaf.  The next leaf is asked for when a leaf starts, so the
// atomic's round trip overlaps the leaf's own loads.
constexpr int kBigWaves = kBlock / 64;
constexpr uint32_t kBigBlocks = 4096;
constexpr uint32_t kCollectBlocks = 2048;

template <int kMode, bool kKahan>
__global__ __launch_bounds__(kBlock) void p2p_big_kernel(
    size_t n, const float4* __restrict__ pos4, const float4* __restrict__ loc,
    const uint32_t* __restrict__ p2p_off, const uint32_t* __restrict__ p2p_len,
    const uint32_t* __restrict__ p2p_moff, const uint32_t* __restrict__ arena,
    const unsigned long long* __restrict__ masks, const uint2* __restrict__ link,
    const uint32_t* __restrict__ nend, float* __restrict__ fx, float* __restrict__ fy,
    float* __restrict__ fz, uint32_t* big, unsigned long long* block_pairs,
    uint32_t slot0, uint32_t stride) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nbig = big[0];
    uint32_t pairs = 0, nsrc = 0;
    // wave k starts with leaf k; later leaves come from the counter
    const uint32_t nw = gridDim.x * kBigWaves;
    uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * kBigWaves + (threadIdx.x >> 6));
    while (w < nbig) {
        uint32_t wn = 0;
        if (lane == 0) wn = nw + atomicAdd(&big[1], 1u);
        const uint32_t node = big[2 + w];
        const uint32_t t0 = link[node].x;
        const uint32_t te = nend[node] < (uint32_t)n ? nend[node] : (uint32_t)n;
        const uint32_t L = (kMode != 1) ? p2p_len[node] : 0u;
        const uint32_t o = L ? p2p_off[node] : 0u;
        const uint32_t mo = L ? p2p_moff[node] : 0u;
        const float4 fl = (kMode != 2) ? loc[node] : make_float4(0.f, 0.f, 0.f, 0.f);
        // rounds of 64 T targets, T = 1..4 from what remains: each source
        // load then feeds T pairs per lane
        bool first = true;  // the leaf pairs are counted once per leaf
        for (uint32_t c0 = t0; c0 < te;) {
            const uint32_t rem = te - c0;
            uint32_t ws = 0;
            if (rem <= 32) {
                // the last targets, several lanes each
                const uint32_t lk = 31u - (uint32_t)__clz(64u / rem);
                p2p_split_round<kMode, kKahan>(c0, te, lane, n, L, o, mo, fl, pos4, arena, masks,
                                               link, nend, fx, fy, fz, pairs, ws, lk);
                c0 = te;
            } else if (rem >= 64 && rem % 64u != 0u && rem < 256) {
                // whole rounds of 64 targets per lane group, the rest split
                const uint32_t T = rem >> 6;
                if (T >= 3)
                    p2p_big_round<kMode, kKahan, 3>(c0, te, lane, n, L, o, mo, fl, pos4, arena,
                                                    masks, link, nend, fx, fy, fz, pairs, ws);
                else if (T == 2)
                    p2p_big_round<kMode, kKahan, 2>(c0, te, lane, n, L, o, mo, fl, pos4, arena,
                                                    masks, link, nend, fx, fy, fz, pairs, ws);
                else
                    p2p_big_round<kMode, kKahan, 1>(c0, te, lane, n, L, o, mo, fl, pos4, arena,
                                                    masks, link, nend, fx, fy, fz, pairs, ws);
                c0 += 64u * T;
            } else if (rem > 192) {
                p2p_big_round<kMode, kKahan, 4>(c0, te, lane, n, L, o, mo, fl, pos4, arena, masks,
                                                link, nend, fx, fy, fz, pairs, ws);
                c0 += 256;
            } else if (rem > 128) {
                p2p_big_round<kMode, kKahan, 3>(c0, te, lane, n, L, o, mo, fl, pos4, arena, masks,
                                                link, nend, fx, fy, fz, pairs, ws);
                c0 += 192;
            } else if (rem > 64) {
                p2p_big_round<kMode, kKahan, 2>(c0, te, lane, n, L, o, mo, fl, pos4, arena, masks,
                                                link, nend, fx, fy, fz, pairs, ws);
                c0 += 128;
            } else {
                p2p_big_round<kMode, kKahan, 1>(c0, te, lane, n, L, o, mo, fl, pos4, arena, masks,
                                                link, nend, fx, fy, fz, pairs, ws);
                c0 += 64;
            }
            if (first && lane == 0) nsrc += ws;
            first = false;
        }
        w = __builtin_amdgcn_readfirstlane(wn);
    }
    if (block_pairs) p2p_block_stats(pairs, nsrc, block_pairs, slot0 + blockIdx.x, stride);
}

uint32_t big_blocks(uint32_t bound) {
    const uint32_t b = (bound + kBigWaves - 1) / kBigWaves;
    return b < kBigBlocks ? b : kBigBlocks;
}

template <int kMode, bool kKahan>
void launch_mode(size_t n, const float4* pos4, const uint32_t* leaf_of, const float4* loc,
                 const uint32_t* p2p_off, const uint32_t* p2p_len, const uint32_t* p2p_moff,
                 const uint32_t* arena, const unsigned long long* masks, const uint2* link,
                 const uint32_t* nend, float* fx, float* fy, float* fz, uint32_t bound,
                 uint32_t tmin, uint32_t* big, unsigned long long* bp,
                 hipStream_t s, hipStream_t s_big, hipEvent_t fork, hipEvent_t join) {
    const uint32_t nb1 = (uint32_t)((n + kBlock - 1) / kBlock);
    const uint32_t nb2 = tmin ? big_blocks(bound) : 0u;
    // the two kernels write disjoint particles: the big-leaf one runs on its
    // own stream beside the other (fork / join through events)
    const bool two = nb2 && s_big && fork && join;
    if (nb2) {
        if (two) {
            (void)hipEventRecord(fork, s);
            (void)hipStreamWaitEvent(s_big, fork, 0);
        }
        p2p_big_kernel<kMode, kKahan><<<nb2, kBlock, 0, two ? s_big : s>>>(
            n, pos4, loc, p2p_off, p2p_len, p2p_moff, arena, masks, link, nend, fx, fy, fz, big,
            bp, nb1, nb1 + nb2);
    }
    p2p_kernel<kMode, kKahan><<<nb1, kBlock, 0, s>>>(n, pos4, leaf_of, loc, p2p_off, p2p_len,
                                                     p2p_moff, arena, masks, link, nend, fx, fy,
                                                     fz, nb2 ? tmin : 0u, bp, nb1 + nb2);
    if (two) {
        (void)hipEventRecord(join, s_big);
        (void)hipStreamWaitEvent(s, join, 0);
    }
}

}  // namespace

void ps_launch_p2p_collect(const uint32_t* ls, uint32_t bound, uint32_t tmin, const uint2* link,
                           const uint32_t* nend, uint32_t* big, hipStream_t s) {
    if (!big || bound == 0 || tmin == 0) return;
    ps_fill_u32(big, 0u, 2, s);  // count, queue
    const uint32_t cb = (bound + kCollect - 1) / kCollect;
    p2p_collect_kernel<<<cb < kCollectBlocks ? cb : kCollectBlocks, kCollect, 0, s>>>(
        ls, tmin, link, nend, big);
}

uint32_t ps_p2p_blocks(size_t n, uint32_t bound) {
    return (uint32_t)((n + kBlock - 1) / kBlock) + big_blocks(bound);
}

void ps_launch_p2p(size_t n, const float4* pos4, const uint32_t* leaf_of, const float4* loc,
                   const uint32_t* p2p_off, const uint32_t* p2p_len, const uint32_t* p2p_moff,
                   const uint32_t* arena, const unsigned long long* masks, const uint2* link,
                   const uint32_t* nend, float* fx, float* fy, float* fz, int mode, int kahan,
                   const uint32_t* ls, uint32_t bound, uint32_t tmin, uint32_t* big,
                   unsigned long long* block_pairs, unsigned long long* pair_count,
                   hipStream_t s, hipStream_t s_big, hipEvent_t fork, hipEvent_t join,
                   bool collected) {
    if (n == 0) return;
    if (!big || bound == 0) tmin = 0;  // no big-leaf kernel
    unsigned long long* bp = pair_count ? block_pairs : nullptr;
    if (tmin && !collected) ps_launch_p2p_collect(ls, bound, tmin, link, nend, big, s);
#define PS_P2P(M, K)                                                                           \
    launch_mode<M, K>(n, pos4, leaf_of, loc, p2p_off, p2p_len, p2p_moff, arena, masks, link, \
                      nend, fx, fy, fz, bound, tmin, big, bp, s, s_big, fork, join)
    switch (mode * 2 + (kahan ? 1 : 0)) {
        case 0: PS_P2P(0, false); break;
        case 1: PS_P2P(0, true); break;
        case 2: PS_P2P(1, false); break;
        case 3: PS_P2P(1, true); break;
        case 4: PS_P2P(2, false); break;
        default: PS_P2P(2, true); break;
    }
#undef PS_P2P
    if (bp) {
        const uint32_t nb = ps_p2p_blocks(n, tmin ? bound : 0u);
        ps_reduce_u64(bp, nb, pair_count, s);
        ps_reduce_u64(bp + nb, nb, pair_count + 1, s);
    }
}
