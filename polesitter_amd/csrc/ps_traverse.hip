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
// One iteration = count -> scan -> write, a group of G lanes (16, 32 or 64;
// several groups per wavefront) per entry:
//   count  MAC for each listed source (the group takes the list in strides),
//          M2L for far ones, a 64-bit far mask per 64 list entries,
//          and the output sizes; a final leaf records (list, masks) for P2P.
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
                                                           PsTravEntries ent) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n = ls1 + k;
    if (n >= ls3) return;
    ent.tgt[k] = n;
    ent.off[k] = 0;  // the list {root} at arena offset 0
    // jobs: every node at depth 2 and every leaf above it (H:1780)
    ent.len[k] = (n >= ls2 || (link[n].y & PS_LEAF)) ? 1u : 0u;
    ent.moff[k] = k;  // one mask word per entry
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

// G lanes per entry (16, 32 or 64): the group takes the entry's list in
// strides of G; source geometry, multipole and link are gathered together so
// one dependent round trip covers the MAC and both outcomes; counts come from
// ballots.
template <bool kKahan, int G>
__global__ __launch_bounds__(kBlock) void trav_count_kernel(
    PsTravTree t, float theta2, PsTravEntries ent, const uint32_t* __restrict__ arena,
    unsigned long long* __restrict__ masks, uint32_t n_ent, uint4* __restrict__ cnt) {
    constexpr int kGroups = kBlock / G;
    const uint32_t e = blockIdx.x * kGroups + threadIdx.x / G;
    const uint32_t gl = threadIdx.x & (G - 1);
    const uint32_t gshift = (threadIdx.x & 63) & ~(uint32_t)(G - 1);
    if (e >= n_ent) return;  // uniform per group
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
            if (gl == 0) masks[mo + (k0 >> 6)] = word;
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
    if (gl != 0) return;
    if (nfar) {
        float4 l = tloc;
        l.y += ex;
        l.z += ey;
        l.w += ez;
        t.loc[tn] = l;
    }
    const uint32_t out = nleaf + nchild;
    uint4 c = make_uint4(0u, 0u, 0u, nfar);
    uint32_t fin = 0;
    if (tl.y & PS_LEAF) {
        if (nchild == 0) {  // final: the near entries of this list are the P2P sources
            if (nleaf) {
                t.p2p_off[tn] = o;
                t.p2p_len[tn] = L;
                t.p2p_moff[tn] = mo;
            }
            fin = nleaf;
        } else {
            c = make_uint4(out, 1u, chunks_of(out), nfar);
        }
    } else if (out) {
        const uint32_t nk = ps_child_count(tl);
        c = make_uint4(out, nk, nk * chunks_of(out), nfar);
    }
    if (t.fin_leaves) t.fin_leaves[e] = fin;
    cnt[e] = c;
}

template <int G>
__global__ __launch_bounds__(kBlock) void trav_write_kernel(
    PsTravTree t, PsTravEntries ent, uint32_t* __restrict__ arena,
    const unsigned long long* __restrict__ masks, uint32_t n_ent, const uint4* __restrict__ cnt,
    const uint4* __restrict__ pos, uint32_t list_base, uint32_t mask_base, PsTravEntries next) {
    constexpr int kGroups = kBlock / G;
    const uint32_t e = blockIdx.x * kGroups + threadIdx.x / G;
    const uint32_t gl = threadIdx.x & (G - 1);
    if (e >= n_ent) return;
    const uint4 c = cnt[e];
    if (c.x == 0) return;  // final or empty: nothing to pass on (uniform per group)
    const uint4 p = pos[e];
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
                         PsTravEntries ent, hipStream_t s) {
    if (ls3 <= ls1) return;
    uint32_t blocks = (ls3 - ls1 + kBlock - 1) / kBlock;
    trav_init_kernel<<<blocks, kBlock, 0, s>>>(ls1, ls2, ls3, link, ent);
}

void ps_launch_trav_count(PsTravTree t, float theta2, PsTravEntries ent, const uint32_t* arena,
                          unsigned long long* masks, uint32_t n_ent, uint4* cnt, int kahan,
                          int lanes, hipStream_t s) {
    if (n_ent == 0) return;
#define PS_COUNT(KH, G)                                                                       \
    trav_count_kernel<KH, G><<<(n_ent + kBlock / G - 1) / (kBlock / G), kBlock, 0, s>>>(       \
        t, theta2, ent, arena, masks, n_ent, cnt)
    if (kahan) {
        if (lanes == 16) PS_COUNT(true, 16);
        else if (lanes == 64) PS_COUNT(true, 64);
        else PS_COUNT(true, 32);
    } else {
        if (lanes == 16) PS_COUNT(false, 16);
        else if (lanes == 64) PS_COUNT(false, 64);
        else PS_COUNT(false, 32);
    }
#undef PS_COUNT
}

void ps_launch_trav_write(PsTravTree t, PsTravEntries ent, uint32_t* arena,
                          const unsigned long long* masks, uint32_t n_ent, const uint4* cnt,
                          const uint4* pos, uint32_t list_base, uint32_t mask_base,
                          PsTravEntries next, int lanes, hipStream_t s) {
    if (n_ent == 0) return;
#define PS_WRITE(G)                                                                           \
    trav_write_kernel<G><<<(n_ent + kBlock / G - 1) / (kBlock / G), kBlock, 0, s>>>(           \
        t, ent, arena, masks, n_ent, cnt, pos, list_base, mask_base, next)
    if (lanes == 16) PS_WRITE(16);
    else if (lanes == 64) PS_WRITE(64);
    else PS_WRITE(32);
#undef PS_WRITE
}
