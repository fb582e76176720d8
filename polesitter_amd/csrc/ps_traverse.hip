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
// write (MAC again, emit next lists / next entries / P2P sources), one
// wavefront per entry in both kernels.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

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

// One wavefront per entry: the 64 lanes take the entry's source list in
// strides, so list loads are coalesced and the per-source work (MAC, M2L or
// the near-field bookkeeping) is spread over the wave; sums are reduced
// across lanes at the end.
template <bool kKahan>
__global__ __launch_bounds__(kBlock) void trav_count_kernel(
    PsTravTree t, float theta2, const uint32_t* __restrict__ tgt,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ len,
    const uint32_t* __restrict__ list, uint32_t n_ent, uint4* __restrict__ cnt) {
    const uint32_t e = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (e >= n_ent) return;  // uniform per wave
    const uint32_t tn = tgt[e];
    const uint32_t o = off[e];
    const uint32_t L = len[e];
    const float4 tg = t.geom[tn];
    const uint2 tl = t.link[tn];
    float ex = 0.f, ey = 0.f, ez = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t nfar = 0, nleaf = 0, nchild = 0;
    for (uint32_t k = lane; k < L; k += 64) {
        const uint32_t s = list[o + k];
        if (s >= t.n_nodes) continue;  // never taken when counts and writes agree
        const float4 sg = t.geom[s];
        float dx, dy, dz, d2;
        if (ps_far_d(tg, sg, theta2, dx, dy, dz, d2)) {
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
    uint4 c;
    if (tl.y & PS_LEAF) {
        c = (nchild == 0) ? make_uint4(0u, 0u, nleaf, nfar) : make_uint4(out, 1u, 0u, nfar);
    } else {
        c = out ? make_uint4(out, ps_child_count(tl), 0u, nfar) : make_uint4(0u, 0u, 0u, nfar);
    }
    cnt[e] = c;
}

// One wavefront per entry: each chunk of 64 sources re-runs the MAC, the
// lanes' output counts (1 per near leaf, the child count per near internal
// node) are prefix-summed across the wave, and the outputs land contiguously
// in source order.
__global__ __launch_bounds__(kBlock) void trav_write_kernel(
    PsTravTree t, float theta2, const uint32_t* __restrict__ tgt,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ len,
    const uint32_t* __restrict__ list, uint32_t n_ent, const uint4* __restrict__ cnt,
    const uint4* __restrict__ pos, uint32_t p2p_base, uint32_t* __restrict__ ntgt,
    uint32_t* __restrict__ noff, uint32_t* __restrict__ nlen, uint32_t* __restrict__ nlist) {
    const uint32_t e = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
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
    // outputs fill exactly the counted range; the bound only makes that an
    // invariant of the stores
    uint32_t base = final_p2p ? p2p_base + p.z : p.x;
    const uint32_t limit = base + (final_p2p ? c.z : c.x);
    for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t s = 0, nout = 0;
        uint2 sl = make_uint2(0u, 0u);
        if (k < L) {
            s = list[o + k];
            if (s < t.n_nodes && !ps_far(tg, t.geom[s], theta2)) {
                sl = t.link[s];
                nout = (sl.y & PS_LEAF) ? 1u : (uint32_t)__popc(sl.y & 0xFFu);
            }
        }
        uint32_t total;
        const uint32_t w = base + ps_wave_excl_scan(nout, &total);
        if (nout) {
            if (sl.y & PS_LEAF) {
                if (w < limit) {
                    if (final_p2p)
                        t.p2p_src[w] = make_uint2(sl.x, t.nend[s]);
                    else
                        nlist[w] = s;
                }
            } else {
                for (uint32_t r = 0; r < nout; ++r)
                    if (w + r < limit) nlist[w + r] = sl.x + r;
            }
        }
        base += total;
    }
    if (c.y) {
        if (tl.y & PS_LEAF) {
            if (lane == 0) {
                ntgt[p.y] = tn;
                noff[p.y] = p.x;
                nlen[p.y] = c.x;
            }
        } else if (lane < c.y) {
            ntgt[p.y + lane] = tl.x + lane;
            noff[p.y + lane] = p.x;
            nlen[p.y + lane] = c.x;
        }
    }
    if (final_p2p && lane == 0) {
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
    uint32_t blocks = (n_ent + kWavesPerBlock - 1) / kWavesPerBlock;
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
    uint32_t blocks = (n_ent + kWavesPerBlock - 1) / kWavesPerBlock;
    trav_write_kernel<<<blocks, kBlock, 0, s>>>(t, theta2, tgt, off, len, list, n_ent, cnt, pos,
                                                p2p_base, ntgt, noff, nlen, nlist);
}
