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
//   internal nodes; once it holds only leaves, those are t's P2P sources.
// This visits exactly the reference's pairs (the pair set, not the order),
// and every target owns its accumulations: no atomics, deterministic output.
//
// One iteration = count (MAC + M2L, per-entry output sizes) -> scan ->
// write (MAC again, emit next lists / next entries / P2P sources).  One
// thread per entry; the siblings of a family are adjacent entries reading the
// same list, so their source loads coincide.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void trav_init_kernel(uint32_t ls1, uint32_t ls2,
                                                           uint32_t ls3,
                                                           const uint2* __restrict__ link,
                                                           uint32_t* tgt, uint32_t* off,
                                                           uint32_t* len) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n = ls1 + k;
    if (n >= ls3) return;
    tgt[k] = n;
    off[k] = 0;  // the list {root} at offset 0
    // jobs: every node at depth 2 and every leaf above it (H:1780)
    len[k] = (n >= ls2 || (link[n].y & PS_LEAF)) ? 1u : 0u;
}

template <bool kKahan>
__global__ __launch_bounds__(kBlock) void trav_count_kernel(
    PsTravTree t, float theta2, const uint32_t* __restrict__ tgt,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ len,
    const uint32_t* __restrict__ list, uint32_t n_ent, uint4* __restrict__ cnt) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= n_ent) return;
    const uint32_t tn = tgt[e];
    const uint32_t o = off[e];
    const uint32_t L = len[e];
    const float4 tg = t.geom[tn];
    const uint2 tl = t.link[tn];
    float ex = 0.f, ey = 0.f, ez = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t nfar = 0, nleaf = 0, nchild = 0;
    for (uint32_t k = 0; k < L; ++k) {
        const uint32_t s = list[o + k];
        if (s >= t.n_nodes) continue;  // never taken when counts and writes agree
        const float4 sg = t.geom[s];
        if (ps_far(tg, sg, theta2)) {
            if (kKahan) {
                float px = 0.f, py = 0.f, pz = 0.f;
                ps_m2l(tg, sg, t.mp[s], px, py, pz);
                ps_kahan_add(ex, cx, px);
                ps_kahan_add(ey, cy, py);
                ps_kahan_add(ez, cz, pz);
            } else {
                ps_m2l(tg, sg, t.mp[s], ex, ey, ez);
            }
            ++nfar;
        } else {
            const uint2 sl = t.link[s];
            if (sl.y & PS_LEAF)
                ++nleaf;
            else
                nchild += (uint32_t)__popc(sl.y & 0xFFu);
        }
    }
    if (nfar) {
        float4 l = t.loc[tn];
        l.y += ex;
        l.z += ey;
        l.w += ez;
        t.loc[tn] = l;
    }
    const uint32_t out = nleaf + nchild;
    uint4 c;
    if (tl.y & PS_LEAF) {
        c = (nchild == 0) ? make_uint4(0u, 0u, nleaf, nfar) : make_uint4(out, 1u, 0u, nfar);
    } else {
        c = out ? make_uint4(out, ps_child_count(tl), 0u, nfar) : make_uint4(0u, 0u, 0u, nfar);
    }
    cnt[e] = c;
}

__global__ __launch_bounds__(kBlock) void trav_write_kernel(
    PsTravTree t, float theta2, const uint32_t* __restrict__ tgt,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ len,
    const uint32_t* __restrict__ list, uint32_t n_ent, const uint4* __restrict__ cnt,
    const uint4* __restrict__ pos, uint32_t p2p_base, uint32_t* __restrict__ ntgt,
    uint32_t* __restrict__ noff, uint32_t* __restrict__ nlen, uint32_t* __restrict__ nlist) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= n_ent) return;
    const uint4 c = cnt[e];
    if (c.x == 0 && c.z == 0) return;
    const uint4 p = pos[e];
    const uint32_t tn = tgt[e];
    const uint32_t o = off[e];
    const uint32_t L = len[e];
    const float4 tg = t.geom[tn];
    const uint2 tl = t.link[tn];
    const bool final_p2p = c.z > 0;
    // the MAC is re-evaluated with the same code as the count pass, so the
    // outputs fill exactly the counted ranges; the bounds below only make
    // that an invariant of the writes
    uint32_t w = p.x;
    const uint32_t w_end = p.x + c.x;
    uint32_t q = p2p_base + p.z;
    const uint32_t q_end = q + c.z;
    for (uint32_t k = 0; k < L; ++k) {
        const uint32_t s = list[o + k];
        if (s >= t.n_nodes) continue;
        const float4 sg = t.geom[s];
        if (ps_far(tg, sg, theta2)) continue;
        const uint2 sl = t.link[s];
        if (sl.y & PS_LEAF) {
            if (final_p2p) {
                if (q < q_end) t.p2p_src[q] = make_uint2(sl.x, t.nend[s]);
                ++q;
            } else {
                if (w < w_end) nlist[w] = s;
                ++w;
            }
        } else {
            const uint32_t nc = (uint32_t)__popc(sl.y & 0xFFu);
            for (uint32_t r = 0; r < nc; ++r, ++w)
                if (w < w_end) nlist[w] = sl.x + r;
        }
    }
    if (c.y) {
        if (tl.y & PS_LEAF) {
            ntgt[p.y] = tn;
            noff[p.y] = p.x;
            nlen[p.y] = c.x;
        } else {
            for (uint32_t r = 0; r < c.y; ++r) {
                ntgt[p.y + r] = tl.x + r;
                noff[p.y + r] = p.x;
                nlen[p.y + r] = c.x;
            }
        }
    }
    if (final_p2p) {
        t.p2p_off[tn] = p2p_base + p.z;
        t.p2p_len[tn] = c.z;
    }
}

}  // namespace

void ps_launch_trav_init(uint32_t ls1, uint32_t ls3, uint32_t ls2, const uint2* link,
                         uint32_t* tgt, uint32_t* off, uint32_t* len, hipStream_t s) {
    if (ls3 <= ls1) return;
    uint32_t blocks = (ls3 - ls1 + kBlock - 1) / kBlock;
    trav_init_kernel<<<blocks, kBlock, 0, s>>>(ls1, ls2, ls3, link, tgt, off, len);
}

void ps_launch_trav_count(PsTravTree t, float theta2, const uint32_t* tgt, const uint32_t* off,
                          const uint32_t* len, const uint32_t* list, uint32_t n_ent,
                          uint4* cnt, int kahan, hipStream_t s) {
    if (n_ent == 0) return;
    uint32_t blocks = (n_ent + kBlock - 1) / kBlock;
    if (kahan)
        trav_count_kernel<true><<<blocks, kBlock, 0, s>>>(t, theta2, tgt, off, len, list, n_ent,
                                                          cnt);
    else
        trav_count_kernel<false><<<blocks, kBlock, 0, s>>>(t, theta2, tgt, off, len, list,
                                                           n_ent, cnt);
}

void ps_launch_trav_write(PsTravTree t, float theta2, const uint32_t* tgt, const uint32_t* off,
                          const uint32_t* len, const uint32_t* list, uint32_t n_ent,
                          const uint4* cnt, const uint4* pos, uint32_t p2p_base,
                          uint32_t* ntgt, uint32_t* noff, uint32_t* nlen, uint32_t* nlist,
                          hipStream_t s) {
    if (n_ent == 0) return;
    uint32_t blocks = (n_ent + kBlock - 1) / kBlock;
    trav_write_kernel<<<blocks, kBlock, 0, s>>>(t, theta2, tgt, off, len, list, n_ent, cnt, pos,
                                                p2p_base, ntgt, noff, nlen, nlist);
}
