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
// One iteration = count -> scan -> write, one wavefront per entry:
//   count  MAC for each listed source (the 64 lanes take the list in
//          strides), M2L for far ones, a 64-bit far mask per 64-entry chunk,
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

template <bool kKahan>
__global__ __launch_bounds__(kBlock) void trav_count_kernel(
    PsTravTree t, float theta2, PsTravEntries ent, const uint32_t* __restrict__ arena,
    unsigned long long* __restrict__ masks, uint32_t n_ent, uint4* __restrict__ cnt) {
    const uint32_t e = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (e >= n_ent) return;  // uniform per wave
    const uint32_t tn = ent.tgt[e];
    const uint32_t o = ent.off[e];
    const uint32_t L = ent.len[e];
    const uint32_t mo = ent.moff[e];
    const float4 tg = t.geom[tn];
    const uint2 tl = t.link[tn];
    float ex = 0.f, ey = 0.f, ez = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t nfar = 0, nleaf = 0, nchild = 0;
    for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        const uint32_t k = k0 + lane;
        bool far = false;
        if (k < L) {
            const uint32_t s = arena[o + k];
            if (s < t.n_nodes) {  // always true when counts and writes agree
                const float4 sg = t.geom[s];
                float dx, dy, dz, d2;
                far = ps_far_d(tg, sg, theta2, dx, dy, dz, d2);
                if (far) {
                    if (kKahan) {
                        float px = 0.f, py = 0.f, pz = 0.f;
                        ps_m2l_d(dx, dy, dz, d2, t.mp[s], px, py, pz);
                        ps_kahan_add(ex, cx, px);
                        ps_kahan_add(ey, cy, py);
                        ps_kahan_add(ez, cz, pz);
                    } else {
                        ps_m2l_d(dx, dy, dz, d2, t.mp[s], ex, ey, ez);
                    }
                    ++nfar;
                } else {
                    const uint2 sl = t.link[s];
                    if (sl.y & PS_LEAF)
                        ++nleaf;
                    else
                        nchild += (uint32_t)__popc(sl.y & 0xFFu);
                }
            } else {
                far = true;  // an invalid id is skipped by every consumer
            }
        }
        const unsigned long long fm = __ballot(far);
        if (lane == 0) masks[mo + (k0 >> 6)] = fm;
    }
    if (kKahan) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            ps_kahan_merge(ex, cx, __shfl_xor(ex, d, 64), __shfl_xor(cx, d, 64));
            ps_kahan_merge(ey, cy, __shfl_xor(ey, d, 64), __shfl_xor(cy, d, 64));
            ps_kahan_merge(ez, cz, __shfl_xor(ez, d, 64), __shfl_xor(cz, d, 64));
        }
        ex -= cx;
        ey -= cy;
        ez -= cz;
    } else {
        ex = ps_wave_sum(ex);
        ey = ps_wave_sum(ey);
        ez = ps_wave_sum(ez);
    }
    nfar = ps_wave_sum_u32(nfar);
    nleaf = ps_wave_sum_u32(nleaf);
    nchild = ps_wave_sum_u32(nchild);
    if (lane != 0) return;
    if (nfar) {
        float4 l = t.loc[tn];
        l.y += ex;
        l.z += ey;
        l.w += ez;
        t.loc[tn] = l;
    }
    const uint32_t out = nleaf + nchild;
    uint4 c = make_uint4(0u, 0u, 0u, nfar);
    if (tl.y & PS_LEAF) {
        if (nchild == 0) {  // final: the near entries of this list are the P2P sources
            if (nleaf) {
                t.p2p_off[tn] = o;
                t.p2p_len[tn] = L;
                t.p2p_moff[tn] = mo;
            }
            if (t.fin_leaves) t.fin_leaves[e] = nleaf;
        } else {
            if (t.fin_leaves) t.fin_leaves[e] = 0u;
            c = make_uint4(out, 1u, chunks_of(out), nfar);
        }
    } else {
        if (t.fin_leaves) t.fin_leaves[e] = 0u;
        if (out) {
            const uint32_t nk = ps_child_count(tl);
            c = make_uint4(out, nk, nk * chunks_of(out), nfar);
        }
    }
    cnt[e] = c;
}

__global__ __launch_bounds__(kBlock) void trav_write_kernel(
    PsTravTree t, PsTravEntries ent, uint32_t* __restrict__ arena,
    const unsigned long long* __restrict__ masks, uint32_t n_ent, const uint4* __restrict__ cnt,
    const uint4* __restrict__ pos, uint32_t list_base, uint32_t mask_base, PsTravEntries next) {
    const uint32_t e = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (e >= n_ent) return;
    const uint4 c = cnt[e];
    if (c.x == 0) return;  // final or empty: nothing to pass on (uniform)
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
    for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        const uint32_t k = k0 + lane;
        const unsigned long long fm = masks[mo + (k0 >> 6)];
        uint32_t s = 0, nout = 0;
        uint2 sl = make_uint2(0u, 0u);
        if (k < L && !((fm >> lane) & 1ull)) {
            s = arena[o + k];
            sl = t.link[s];
            nout = (sl.y & PS_LEAF) ? 1u : (uint32_t)__popc(sl.y & 0xFFu);
        }
        uint32_t total;
        const uint32_t w = base + ps_wave_excl_scan(nout, &total);
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
        if (lane == 0) {
            next.tgt[p.y] = tn;
            next.off[p.y] = list_base + p.x;
            next.len[p.y] = c.x;
            next.moff[p.y] = mask_base + p.z;
        }
    } else if (lane < c.y) {
        next.tgt[p.y + lane] = tl.x + lane;
        next.off[p.y + lane] = list_base + p.x;
        next.len[p.y + lane] = c.x;
        next.moff[p.y + lane] = mask_base + p.z + lane * nch;
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
                          hipStream_t s) {
    if (n_ent == 0) return;
    uint32_t blocks = (n_ent + kWavesPerBlock - 1) / kWavesPerBlock;
    if (kahan)
        trav_count_kernel<true><<<blocks, kBlock, 0, s>>>(t, theta2, ent, arena, masks, n_ent,
                                                          cnt);
    else
        trav_count_kernel<false><<<blocks, kBlock, 0, s>>>(t, theta2, ent, arena, masks, n_ent,
                                                           cnt);
}

void ps_launch_trav_write(PsTravTree t, PsTravEntries ent, uint32_t* arena,
                          const unsigned long long* masks, uint32_t n_ent, const uint4* cnt,
                          const uint4* pos, uint32_t list_base, uint32_t mask_base,
                          PsTravEntries next, hipStream_t s) {
    if (n_ent == 0) return;
    uint32_t blocks = (n_ent + kWavesPerBlock - 1) / kWavesPerBlock;
    trav_write_kernel<<<blocks, kBlock, 0, s>>>(t, ent, arena, masks, n_ent, cnt, pos, list_base,
                                                mask_base, next);
}
