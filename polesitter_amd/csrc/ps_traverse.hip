// ps_traverse.hip -- the MAC / dual-tree traversal with M2L (replaces the
// recursions of H:1050-1146 and H:1258-1494 and the job roots of
// H:1772-1801 / H:1888,1905).
//
// The reference recursion on a pair (t, s): far -> M2L(t <- s); both leaves
// -> P2P; t leaf -> (t, children of s); s leaf -> (children of t, s);
// otherwise (children of t) x (children of s).  Grouped by target, this is a
// top-down propagation of per-target source lists:
//   L(t) is an antichain of source nodes.  For each s in L(t): far -> M2L;
//   near leaf -> kept; near internal -> replaced by its children.  The
//   result C(t) is the list of EVERY child of an internal t (they all see the
//   same sources), or the next list of a leaf t itself while C(t) still holds
//   internal nodes.  A leaf t whose list has no near internal node is final:
//   the near entries of that list are its P2P sources.
// This visits exactly the reference's pairs (the pair set, not the order),
// and every target owns its accumulations: no atomics on the results.
//
// One iteration = count -> plan -> write, a group of G lanes (8 to 64;
// several groups per wavefront) per entry:
//   count  MAC for each listed source (the group takes the list in strides),
//          M2L for far ones, a 64-bit far mask per 64 list entries, the
//          output sizes and their per-block sums; a final leaf records
//          (list, masks) for P2P.
//   plan   one block scans the block sums and does the bookkeeping (next
//          entry count, arena bases, capacity check).
//   write  for continuing entries only: the mask picks the near sources,
//          near leaves are copied and near internal nodes expanded to their
//          children, compacted with a wave prefix sum into the list arena.
// Lists are appended to one arena that lives for the whole tick, so the P2P
// kernel reads each leaf's final list in place.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

__global__ __launch_bounds__(kBlock) void trav_init_kernel(uint32_t ls1, uint32_t ls2,
                                                           uint32_t ls3,
                                                           const uint2* __restrict__ link,
                                                           PsTravEntries ent, PsTravState* st,
                                                           uint32_t* arena, uint32_t arena_cap,
                                                           uint32_t mask_cap) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n = ls1 + k;
    if (k == 0) {
        PsTravState z = {};
        z.n_ent[0] = ls3 - ls1;
        z.arena_used = 1;  // arena word 0 holds the list {root}
        z.mask_used = ls3 - ls1;
        z.arena_cap = arena_cap;
        z.mask_cap = mask_cap;
        *st = z;
        arena[0] = 0;
    }
    if (n >= ls3) return;
    ent.tgt[k] = n;
    ent.off[k] = 0;
    // jobs: every node at depth 2 and every leaf above it (H:1780)
    ent.len[k] = (n >= ls2 || (link[n].y & PS_LEAF)) ? 1u : 0u;
    ent.moff[k] = k;  // one mask word per entry
}

// One block: exclusive scan of the count kernel's per-block sums (in place),
// then the iteration's bookkeeping -- next entry count, the arena bases of
// the write, and the capacity check.  Only the blocks that held entries are
// scanned (groups = entries per count block); each thread takes kPlanV
// consecutive sums per round, loaded together.
constexpr int kPlanThreads = 1024;
constexpr int kPlanV = 8;

__global__ __launch_bounds__(kPlanThreads) void trav_plan_kernel(PsTravState* st, int k,
                                                                 uint4* __restrict__ bsum,
                                                                 uint32_t groups) {
    __shared__ uint4 wtot[kPlanThreads / 64];
    __shared__ unsigned long long wt[4][kPlanThreads / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n_e = st->n_ent[k];
    const uint32_t nb = st->overflow ? 0u : (uint32_t)(((unsigned long long)n_e + groups - 1) / groups);
    uint4 carry = make_uint4(0u, 0u, 0u, 0u);  // same value in every thread
    // exact 64-bit totals (the 32-bit running offsets can wrap on huge inputs;
    // the capacity check then fails on the totals)
    unsigned long long t_list = 0, t_ent = 0, t_chunk = 0, t_far = 0;
    for (uint32_t base = 0; base < nb; base += kPlanThreads * kPlanV) {
        const uint32_t i0 = base + tid * kPlanV;
        uint4 v[kPlanV];
#pragma unroll
        for (int j = 0; j < kPlanV; ++j)
            v[j] = (i0 + j < nb) ? bsum[i0 + j] : make_uint4(0u, 0u, 0u, 0u);
        uint4 sum = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int j = 0; j < kPlanV; ++j) {
            sum.x += v[j].x; sum.y += v[j].y; sum.z += v[j].z; sum.w += v[j].w;
            t_list += v[j].x; t_ent += v[j].y; t_chunk += v[j].z; t_far += v[j].w;
        }
        uint4 inc = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint4 o;
            o.x = __shfl_up(inc.x, d, 64);
            o.y = __shfl_up(inc.y, d, 64);
            o.z = __shfl_up(inc.z, d, 64);
            o.w = __shfl_up(inc.w, d, 64);
            if (lane >= (uint32_t)d) {
                inc.x += o.x; inc.y += o.y; inc.z += o.z; inc.w += o.w;
            }
        }
        if (lane == 63) wtot[wave] = inc;
        __syncthreads();
        uint4 run = carry, all = carry;
        for (uint32_t w = 0; w < kPlanThreads / 64; ++w) {
            const uint4 t = wtot[w];
            if (w < wave) {
                run.x += t.x; run.y += t.y; run.z += t.z; run.w += t.w;
            }
            all.x += t.x; all.y += t.y; all.z += t.z; all.w += t.w;
        }
        run.x += inc.x - sum.x; run.y += inc.y - sum.y;
        run.z += inc.z - sum.z; run.w += inc.w - sum.w;
#pragma unroll
        for (int j = 0; j < kPlanV; ++j) {
            if (i0 + j < nb) bsum[i0 + j] = run;
            run.x += v[j].x; run.y += v[j].y; run.z += v[j].z; run.w += v[j].w;
        }
        carry = all;
        __syncthreads();  // wtot is rewritten by the next round
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        t_list += __shfl_down(t_list, d, 64);
        t_ent += __shfl_down(t_ent, d, 64);
        t_chunk += __shfl_down(t_chunk, d, 64);
        t_far += __shfl_down(t_far, d, 64);
    }
    if (lane == 0) {
        wt[0][wave] = t_list;
        wt[1][wave] = t_ent;
        wt[2][wave] = t_chunk;
        wt[3][wave] = t_far;
    }
    __syncthreads();
    if (tid != 0) return;
    unsigned long long list_n = 0, ent_n = 0, chunk_n = 0, far_total = 0;
    for (int w = 0; w < kPlanThreads / 64; ++w) {
        list_n += wt[0][w];
        ent_n += wt[1][w];
        chunk_n += wt[2][w];
        far_total += wt[3][w];
    }
    PsTravState& S = *st;
    S.m2l_pairs += far_total;
    S.list_entries += list_n;
    if (S.overflow) {
        S.n_ent[k + 1] = 0;
        return;
    }
    const unsigned long long na = (unsigned long long)S.arena_used + list_n;
    const unsigned long long nm = (unsigned long long)S.mask_used + chunk_n;
    if (na > S.arena_cap || nm > S.mask_cap || ent_n > 0xFFFFFFFFull) {
        S.overflow = 1;
        S.arena_need = na;
        S.mask_need = nm;
        S.n_ent[k + 1] = 0;
        return;
    }
    S.list_base = S.arena_used;
    S.mask_base = S.mask_used;
    S.arena_used = (uint32_t)na;
    S.mask_used = (uint32_t)nm;
    S.n_ent[k + 1] = (uint32_t)ent_n;
}

__device__ __forceinline__ uint32_t chunks_of(uint32_t len) { return (len + 63u) >> 6; }

// sum over the G lanes of one group (G a power of two <= 64)
template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
    return v;
}

template <int G>
__device__ __forceinline__ unsigned long long group_bits(bool pred, uint32_t gshift) {
    const unsigned long long b = __ballot(pred) >> gshift;
    return G == 64 ? b : (b & ((1ull << G) - 1ull));
}

// G lanes per entry (8, 16, 32 or 64): the group takes the entry's list in
// strides of G; source geometry, multipole and link are gathered together so
// one dependent round trip covers the MAC and both outcomes; counts come from
// ballots.
template <bool kKahan, int G>
__global__ __launch_bounds__(kBlock) void trav_count_kernel(
    PsTravTree t, float theta2, PsTravEntries ent, const uint32_t* __restrict__ arena,
    unsigned long long* __restrict__ masks, const PsTravState* __restrict__ st, int it,
    uint4* __restrict__ cnt, uint4* __restrict__ bsum) {
    constexpr int kGroups = kBlock / G;
    __shared__ uint4 gcnt[kGroups];
    const uint32_t grp = threadIdx.x / G;
    const uint32_t e = blockIdx.x * kGroups + grp;
    const uint32_t gl = threadIdx.x & (G - 1);
    const uint32_t gshift = (threadIdx.x & 63) & ~(uint32_t)(G - 1);
    const bool active = e < st->n_ent[it] && !st->overflow;  // uniform per group
    uint4 c = make_uint4(0u, 0u, 0u, 0u);
    if (active) {
        const uint32_t mask_cap = st->mask_cap;
        const uint32_t tn = ent.tgt[e];
        const uint32_t o = ent.off[e];
        const uint32_t L = ent.len[e];
        const uint32_t mo = ent.moff[e];
        const float4 tg = t.geom[tn];
        const uint2 tl = t.link[tn];
        const float4 tloc = t.loc[tn];
        float ex = 0.f, ey = 0.f, ez = 0.f;
        float cx = 0.f, cy = 0.f, cz = 0.f;
        uint32_t nfar = 0, nleaf = 0, nchild = 0;
        unsigned long long word = 0;
        for (uint32_t k0 = 0; k0 < L; k0 += G) {
            const uint32_t k = k0 + gl;
            bool far = false, near_leaf = false;
            uint32_t cc = 0;
            if (k < L) {
                const uint32_t s = arena[o + k];
                if (s < t.n_nodes) {  // always true when counts and writes agree
                    const float4 sg = t.geom[s];
                    const float4 smp = t.mp[s];
                    const uint2 sl = t.link[s];
                    float dx, dy, dz, d2;
                    far = ps_far_d(tg, sg, theta2, dx, dy, dz, d2);
                    if (far) {
                        if (kKahan) {
                            float px = 0.f, py = 0.f, pz = 0.f;
                            ps_m2l_d(dx, dy, dz, d2, smp, px, py, pz);
                            ps_kahan_add(ex, cx, px);
                            ps_kahan_add(ey, cy, py);
                            ps_kahan_add(ez, cz, pz);
                        } else {
                            ps_m2l_d(dx, dy, dz, d2, smp, ex, ey, ez);
                        }
                    } else if (sl.y & PS_LEAF) {
                        near_leaf = true;
                    } else {
                        cc = (uint32_t)__popc(sl.y & 0xFFu);
                    }
                } else {
                    far = true;  // an invalid id is skipped by every consumer
                }
            }
            const unsigned long long fb = group_bits<G>(far, gshift);
            nfar += (uint32_t)__popcll(fb);
            nleaf += (uint32_t)__popcll(group_bits<G>(near_leaf, gshift));
#pragma unroll
            for (int b = 0; b < 4; ++b)
                nchild += (uint32_t)__popcll(group_bits<G>((cc >> b) & 1u, gshift)) << b;
            word |= fb << (k0 & 63u);
            if (((k0 + G) & 63u) == 0 || k0 + G >= L) {
                if (gl == 0 && mo + (k0 >> 6) < mask_cap) masks[mo + (k0 >> 6)] = word;
                word = 0;
            }
        }
        if (kKahan) {
#pragma unroll
            for (int d = G / 2; d > 0; d >>= 1) {
                ps_kahan_merge(ex, cx, __shfl_xor(ex, d, G), __shfl_xor(cx, d, G));
                ps_kahan_merge(ey, cy, __shfl_xor(ey, d, G), __shfl_xor(cy, d, G));
                ps_kahan_merge(ez, cz, __shfl_xor(ez, d, G), __shfl_xor(cz, d, G));
            }
            ex -= cx;
            ey -= cy;
            ez -= cz;
        } else {
            ex = group_sum<G>(ex);
            ey = group_sum<G>(ey);
            ez = group_sum<G>(ez);
        }
        const uint32_t out = nleaf + nchild;
        c = make_uint4(0u, 0u, 0u, nfar);
        if (tl.y & PS_LEAF) {
            if (nchild == 0) {  // final: the near entries of this list are the P2P sources
                if (nleaf && gl == 0) {
                    t.p2p_off[tn] = o;
                    t.p2p_len[tn] = L;
                    t.p2p_moff[tn] = mo;
                }
            } else {
                c = make_uint4(out, 1u, chunks_of(out), nfar);
            }
        } else if (out) {
            const uint32_t nk = ps_child_count(tl);
            c = make_uint4(out, nk, nk * chunks_of(out), nfar);
        }
        if (gl == 0) {
            if (nfar) {
                float4 l = tloc;
                l.y += ex;
                l.z += ey;
                l.w += ez;
                t.loc[tn] = l;
            }
            cnt[e] = c;
        }
    }
    if (gl == 0) gcnt[grp] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 b = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            b.x += gcnt[g].x;
            b.y += gcnt[g].y;
            b.z += gcnt[g].z;
            b.w += gcnt[g].w;
        }
        bsum[blockIdx.x] = b;
    }
}

template <int G>
__global__ __launch_bounds__(kBlock) void trav_write_kernel(
    PsTravTree t, PsTravEntries ent, uint32_t* __restrict__ arena,
    const unsigned long long* __restrict__ masks, const PsTravState* __restrict__ st, int it,
    const uint4* __restrict__ cnt, const uint4* __restrict__ bsum, PsTravEntries next) {
    constexpr int kGroups = kBlock / G;
    __shared__ uint4 gpos[kGroups];
    const uint32_t grp = threadIdx.x / G;
    const uint32_t e = blockIdx.x * kGroups + grp;
    const uint32_t gl = threadIdx.x & (G - 1);
    const uint32_t n_ent = st->n_ent[it];
    if (st->overflow || blockIdx.x * kGroups >= n_ent) return;  // uniform per block
    const uint4 c = e < n_ent ? cnt[e] : make_uint4(0u, 0u, 0u, 0u);
    if (gl == 0) gpos[grp] = c;
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive prefix of the block's entries, from its base
        uint4 run = bsum[blockIdx.x];
        for (int g = 0; g < kGroups; ++g) {
            const uint4 cg = gpos[g];
            gpos[g] = run;
            run.x += cg.x;
            run.y += cg.y;
            run.z += cg.z;
        }
    }
    __syncthreads();
    if (e >= n_ent) return;
    if (c.x == 0) return;  // final or empty: nothing to pass on (uniform per group)
    const uint4 p = gpos[grp];
    const uint32_t list_base = st->list_base;
    const uint32_t mask_base = st->mask_base;
    const uint32_t tn = ent.tgt[e];
    const uint32_t o = ent.off[e];
    const uint32_t L = ent.len[e];
    const uint32_t mo = ent.moff[e];
    const uint2 tl = t.link[tn];
    // the count pass's masks and link reads give exactly c.x outputs; the
    // bound only makes that an invariant of the stores
    uint32_t base = list_base + p.x;
    const uint32_t limit = base + c.x;
    for (uint32_t k0 = 0; k0 < L; k0 += G) {
        const uint32_t k = k0 + gl;
        const unsigned long long fm = masks[mo + (k0 >> 6)];
        uint32_t s = 0, nout = 0;
        uint2 sl = make_uint2(0u, 0u);
        if (k < L && !((fm >> ((k0 & 63u) + gl)) & 1ull)) {
            s = arena[o + k];
            sl = t.link[s];
            nout = (sl.y & PS_LEAF) ? 1u : (uint32_t)__popc(sl.y & 0xFFu);
        }
        uint32_t inc = nout;
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            const uint32_t v = __shfl_up(inc, d, G);
            if (gl >= (uint32_t)d) inc += v;
        }
        const uint32_t total = __shfl(inc, G - 1, G);
        const uint32_t w = base + inc - nout;
        if (nout) {
            if (sl.y & PS_LEAF) {
                if (w < limit) arena[w] = s;
            } else {
                for (uint32_t r = 0; r < nout; ++r)
                    if (w + r < limit) arena[w + r] = sl.x + r;
            }
        }
        base += total;
    }
    const uint32_t nch = chunks_of(c.x);
    if (tl.y & PS_LEAF) {
        if (gl == 0) {
            next.tgt[p.y] = tn;
            next.off[p.y] = list_base + p.x;
            next.len[p.y] = c.x;
            next.moff[p.y] = mask_base + p.z;
        }
    } else {
        for (uint32_t r = gl; r < c.y; r += G) {
            next.tgt[p.y + r] = tl.x + r;
            next.off[p.y + r] = list_base + p.x;
            next.len[p.y + r] = c.x;
            next.moff[p.y + r] = mask_base + p.z + r * nch;
        }
    }
}

}  // namespace

void ps_launch_trav_init(uint32_t ls1, uint32_t ls3, uint32_t ls2, const uint2* link,
                         PsTravEntries ent, PsTravState* st, uint32_t* arena,
                         uint32_t arena_cap, uint32_t mask_cap, hipStream_t s) {
    uint32_t n = ls3 > ls1 ? ls3 - ls1 : 1u;
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    trav_init_kernel<<<blocks, kBlock, 0, s>>>(ls1, ls2, ls3, link, ent, st, arena, arena_cap,
                                               mask_cap);
}

void ps_launch_trav_plan(PsTravState* st, int k, uint4* bsum, int lanes, hipStream_t s) {
    const int g = (lanes == 8 || lanes == 32 || lanes == 64) ? lanes : 16;
    trav_plan_kernel<<<1, kPlanThreads, 0, s>>>(st, k, bsum, (uint32_t)(kBlock / g));
}

uint32_t ps_trav_blocks(uint32_t max_ent, int lanes) {
    const int g = (lanes == 8 || lanes == 32 || lanes == 64) ? lanes : 16;
    return (max_ent + kBlock / g - 1) / (kBlock / g);
}

void ps_launch_trav_count(PsTravTree t, float theta2, PsTravEntries ent, const uint32_t* arena,
                          unsigned long long* masks, const PsTravState* st, int k,
                          uint32_t max_ent, uint4* cnt, uint4* bsum, int kahan, int lanes,
                          hipStream_t s) {
    if (max_ent == 0) return;
#define PS_COUNT(KH, G)                                                                       \
    trav_count_kernel<KH, G><<<(max_ent + kBlock / G - 1) / (kBlock / G), kBlock, 0, s>>>(     \
        t, theta2, ent, arena, masks, st, k, cnt, bsum)
    if (kahan) {
        if (lanes == 8) PS_COUNT(true, 8);
        else if (lanes == 32) PS_COUNT(true, 32);
        else if (lanes == 64) PS_COUNT(true, 64);
        else PS_COUNT(true, 16);
    } else {
        if (lanes == 8) PS_COUNT(false, 8);
        else if (lanes == 32) PS_COUNT(false, 32);
        else if (lanes == 64) PS_COUNT(false, 64);
        else PS_COUNT(false, 16);
    }
#undef PS_COUNT
}

void ps_launch_trav_write(PsTravTree t, PsTravEntries ent, uint32_t* arena,
                          const unsigned long long* masks, const PsTravState* st, int k,
                          uint32_t max_ent, const uint4* cnt, const uint4* bsum,
                          PsTravEntries next, int lanes, hipStream_t s) {
    if (max_ent == 0) return;
#define PS_WRITE(G)                                                                           \
    trav_write_kernel<G><<<(max_ent + kBlock / G - 1) / (kBlock / G), kBlock, 0, s>>>(         \
        t, ent, arena, masks, st, k, cnt, bsum, next)
    if (lanes == 8) PS_WRITE(8);
    else if (lanes == 32) PS_WRITE(32);
    else if (lanes == 64) PS_WRITE(64);
    else PS_WRITE(16);
#undef PS_WRITE
}
