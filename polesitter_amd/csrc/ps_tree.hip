// ps_tree.hip -- octree build from sorted Morton codes, upward pass (P2M /
// M2M) and L2L, replacing H:882-967, H:969-1046 and H:1238-1253.
//
// Build.  The reference builds top-down by binary search (H:886-967): a node
// at depth d+1 exists for every non-empty octant of a node at depth d that
// holds >= 2 particles, d < 10; a node is a leaf at depth 10 or with <= 1
// particle.  Equivalently, with k(i, j) = the number of leading octal digits
// (3-bit groups from bit 29 down) that codes i and j share:
//   D_i      = min(10, max(k(i-1,i), k(i,i+1)) + 1)  depth of i's leaf
//   start_d  = d <= D_i and d > k(i-1,i)            i opens a node at depth d
//   end_d    = d <= D_i and d > k(i,i+1)            i closes a node at depth d
// so every particle knows locally, from its two neighbours, at which depths
// it starts or ends a node.  One scan of the per-(depth, block) start counts,
// flattened depth-major, then gives each node its index in canonical level
// order directly.  Centres follow the reference's float recurrence
// c_child = c_parent +/- fl(hw_parent * 0.5) (H:935-939) exactly.
//
// Upward and L2L run one launch per depth (bottom-up / top-down) and keep the
// reference's summation order (particles in index order, children in slot
// order), so multipoles are bit-identical to the reference's.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kDepths = PS_MAX_DEPTH + 1;

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__device__ __forceinline__ int octal_lcp(uint32_t a, uint32_t b) {
    uint32_t x = (a ^ b) & 0x3FFFFFFFu;
    if (x == 0) return PS_MAX_DEPTH;
    int hb = 31 - __clz(x);  // highest differing bit, 0..29
    return (29 - hb) / 3;
}

struct PartDepths {
    int leaf;      // D_i
    int start_lo;  // starts at depths [start_lo, leaf]
    int end_lo;    // ends at depths [end_lo, leaf]
};

// sys (optional): system of each particle; a system boundary shares no digit,
// so every system opens its own root (the forest of a batched tick).
// leaf_cap: most particles of a leaf above depth 10 (the reference: 1), or
// sys_cap[system] per system when sys_cap is set (a forest).  The
// node of depth d around i holds more than leaf_cap particles iff some
// window of leaf_cap + 1 consecutive particles containing i shares d digits,
// so i's leaf depth is 1 + the largest lcp(j, j + leaf_cap), j in
// [i - leaf_cap, i]; with leaf_cap = 1 that is the rule above.
// The particle's own code is loaded here too, with its neighbours' (and
// their systems): all requested at once, clamped to the array's ends.
__device__ __forceinline__ PartDepths part_depths(const uint32_t* codes,
                                                  const uint32_t* __restrict__ sys, size_t n,
                                                  size_t i, uint32_t& c, int leaf_cap,
                                                  const uint8_t* __restrict__ sys_cap) {
    const size_t ip = i > 0 ? i - 1 : i, in = i + 1 < n ? i + 1 : i;
    c = codes[i];
    const uint32_t cp = codes[ip], cn = codes[in];
    uint32_t si = 0u, sp = 0u, sn = 0u;
    if (sys) {
        si = sys[i];
        sp = sys[ip];
        sn = sys[in];
    }
    if (sys && sys_cap) leaf_cap = sys_cap[si];
    int kp = (i > 0 && sp == si) ? octal_lcp(cp, c) : -1;
    int kn = (i + 1 < n && sn == si) ? octal_lcp(c, cn) : -1;
    int K = kp > kn ? kp : kn;
    if (leaf_cap > 1) {
        K = -1;
        const size_t w = (size_t)leaf_cap;
        for (size_t j = i >= w ? i - w : 0; j <= i; ++j) {
            if (j + w >= n || (sys && sys[j] != sys[j + w])) continue;
            const int k = octal_lcp(codes[j], codes[j + w]);
            K = k > K ? k : K;
        }
    }
    PartDepths r;
    r.leaf = (K + 1) < PS_MAX_DEPTH ? (K + 1) : PS_MAX_DEPTH;
    r.start_lo = kp + 1;
    r.end_lo = kn + 1;
    return r;
}

__global__ __launch_bounds__(kBlock) void tree_count_kernel(const uint32_t* __restrict__ codes,
                                                            const uint32_t* __restrict__ sys,
                                                            size_t n, uint32_t* __restrict__ cnt,
                                                            uint32_t nb, uint32_t* flags,
                                                            int leaf_cap,
                                                            const uint8_t* __restrict__ sys_cap) {
    __shared__ uint32_t ws[kWaves][kDepths];
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    PartDepths pd = {-1, kDepths, kDepths};
    if (valid) {
        uint32_t c;
        pd = part_depths(codes, sys, n, i, c, leaf_cap, sys_cap);
        if (c >> 30) atomicOr(flags, 1u);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < kDepths; ++d) {
        uint64_t b = __ballot(valid && d >= pd.start_lo && d <= pd.leaf);
        if (lane == 0) ws[wave][d] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (threadIdx.x < kDepths) {
        uint32_t s = 0;
        for (int w = 0; w < kWaves; ++w) s += ws[w][threadIdx.x];
        cnt[(size_t)threadIdx.x * nb + blockIdx.x] = s;
    }
}

__global__ void tree_levels_kernel(const uint32_t* off, uint32_t nb,
                                   const unsigned long long* total, uint32_t* level_start) {
    int d = threadIdx.x;
    if (d < kDepths) level_start[d] = off[(size_t)d * nb];
    if (d == kDepths) level_start[kDepths] = (uint32_t)total[0];
}

__global__ __launch_bounds__(kBlock) void tree_write_kernel(
    const uint32_t* __restrict__ codes, const uint32_t* __restrict__ sys, size_t n,
    const uint32_t* __restrict__ off, uint32_t nb, float4 root,
    const float4* __restrict__ roots, float4* __restrict__ geom, uint2* __restrict__ link,
    uint32_t* __restrict__ nend, uint32_t* __restrict__ parent, uint32_t* __restrict__ leaf_of,
    uint8_t* __restrict__ octant, int leaf_cap, uint32_t cap,
    const uint8_t* __restrict__ sys_cap) {
#pragma clang fp contract(off)
    __shared__ uint32_t ws[kWaves][kDepths];
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    uint32_t c = 0;
    PartDepths pd = {-1, kDepths, kDepths};
    if (valid) pd = part_depths(codes, sys, n, i, c, leaf_cap, sys_cap);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t le = ps_lanemask_lt() | (1ull << lane);
    uint32_t inc[kDepths];  // this lane's inclusive start count per depth
#pragma unroll
    for (int d = 0; d < kDepths; ++d) {
        uint64_t b = __ballot(valid && d >= pd.start_lo && d <= pd.leaf);
        if (lane == 0) ws[wave][d] = (uint32_t)__popcll(b);
        inc[d] = (uint32_t)__popcll(b & le);
    }
    __syncthreads();
    if (!valid) return;

    uint32_t node[kDepths];
#pragma unroll
    for (int d = 0; d < kDepths; ++d) {
        uint32_t p = off[(size_t)d * nb + blockIdx.x] + inc[d];
        for (int w = 0; w < wave; ++w) p += ws[w][d];
        node[d] = p - 1u;  // the depth-d node that contains particle i
    }

    uint32_t* link32 = reinterpret_cast<uint32_t*>(link);
    float4 g = roots ? roots[sys[i]] : root;
#pragma unroll
    for (int d = 0; d < kDepths; ++d) {
        if (d > pd.leaf) continue;
        uint32_t oct = 0;
        if (d > 0) {
            oct = (c >> (30 - 3 * d)) & 7u;
            float hw = g.w * 0.5f;
            g.x = g.x + ((oct & 1u) ? hw : -hw);
            g.y = g.y + ((oct & 2u) ? hw : -hw);
            g.z = g.z + ((oct & 4u) ? hw : -hw);
            g.w = hw;
        }
        // nodes past the arrays' capacity are not written (a tree that big
        // is reported by the caller from the node count)
        if (d >= pd.start_lo && node[d] < cap) {  // i opens this node
            const uint32_t nd = node[d];
            geom[nd] = g;
            parent[nd] = d > 0 ? node[d > 0 ? d - 1 : 0] : PS_NONE;
            if (d == pd.leaf) {
                link[nd] = make_uint2((uint32_t)i, PS_LEAF);
            } else {
                link32[2 * (size_t)nd] = node[d + 1 < kDepths ? d + 1 : d];
            }
            if (d > 0) octant[nd] = (uint8_t)oct;  // the parent's mask: tree_mask_kernel
        }
        if (d >= pd.end_lo && node[d] < cap) nend[node[d]] = (uint32_t)i + 1u;
    }
    uint32_t lf = node[0];
#pragma unroll
    for (int d = 1; d < kDepths; ++d)
        if (d <= pd.leaf) lf = node[d];
    leaf_of[i] = lf;
}

// Child masks without atomics: a parent's children are consecutive nodes
// (one level, Morton order), so the first child of each run ORs the octants
// of its <= 8 siblings and stores the parent's mask (link.y of an internal
// node, not written before).  m = min(*count, cap) nodes, or n_nodes.
__global__ __launch_bounds__(kBlock) void tree_mask_kernel(const uint32_t* __restrict__ parent,
                                                           const uint8_t* __restrict__ octant,
                                                           uint2* __restrict__ link,
                                                           const unsigned long long* count,
                                                           size_t n_nodes,
                                                           float4* __restrict__ loc_clear,
                                                           uint32_t* __restrict__ len_clear) {
    const size_t m = count ? (size_t)(*count < n_nodes ? *count : n_nodes) : n_nodes;
    uint32_t* link32 = reinterpret_cast<uint32_t*>(link);
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < m;
         c += (size_t)gridDim.x * kBlock) {
        // what the traversal accumulates starts at zero (optional)
        if (loc_clear) loc_clear[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (len_clear) len_clear[c] = 0u;
        const uint32_t p = parent[c];
        if (p == PS_NONE || (c > 0 && parent[c - 1] == p)) continue;
        // the next seven nodes' parents and octants are loaded at once; the
        // run ends at the first node of another parent
        uint32_t pq[7], oq[7];
#pragma unroll
        for (uint32_t k = 0; k < 7; ++k) {
            const size_t q = c + 1 + k;
            pq[k] = q < m ? parent[q] : PS_NONE;
            oq[k] = q < m ? octant[q] : 0u;
        }
        uint32_t mask = 1u << octant[c];
        bool same = true;
#pragma unroll
        for (uint32_t k = 0; k < 7; ++k) {
            same = same && pq[k] == p;
            mask |= same ? 1u << oq[k] : 0u;
        }
        link32[2 * (size_t)p + 1] = mask;
    }
}

// P2M of particles [i0, e) about the cell centre g, added in particle order
// (H:978-1004); the loads run four particles ahead of the sums
__device__ __forceinline__ void p2m_leaf(const float4* __restrict__ pos4, uint32_t i0,
                                         uint32_t e, float4 g, float& m0, float& mx, float& my,
                                         float& mz) {
#pragma clang fp contract(off)
    auto add = [&](float4 p) {
#pragma clang fp contract(off)
        const float dx = p.x - g.x;
        const float dy = p.y - g.y;
        const float dz = p.z - g.z;
        m0 += p.w;
        mx += p.w * dx;
        my += p.w * dy;
        mz += p.w * dz;
    };
    uint32_t i = i0;
    for (; i + 4 <= e; i += 4) {
        const float4 p0 = pos4[i], p1 = pos4[i + 1], p2 = pos4[i + 2], p3 = pos4[i + 3];
        add(p0);
        add(p1);
        add(p2);
        add(p3);
    }
    for (; i < e; ++i) add(pos4[i]);
}

// M2M of an internal node's children about its centre g, added in slot order
// (H:1013-1039).  The children are the consecutive nodes from l.x, one per
// set bit of the mask, so all their records are loaded before the sums.
__device__ __forceinline__ void m2m_children(const float4* __restrict__ geom,
                                             const float4* __restrict__ mp, uint2 l, float4 g,
                                             float& m0, float& mx, float& my, float& mz) {
#pragma clang fp contract(off)
    const uint32_t cnt = (uint32_t)__popc(l.y & 0xFFu);
    float4 cg[8], cm[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        if (k < cnt) {
            cg[k] = geom[l.x + k];
            cm[k] = mp[l.x + k];
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        if (k < cnt) {
            const float dx = cg[k].x - g.x;
            const float dy = cg[k].y - g.y;
            const float dz = cg[k].z - g.z;
            m0 += cm[k].x;
            mx += cm[k].y + cm[k].x * dx;
            my += cm[k].z + cm[k].x * dy;
            mz += cm[k].w + cm[k].x * dz;
        }
    }
}

// the traversal's node record (ps_tnode_word)
__device__ __forceinline__ float4 trav_node(float4 g, uint2 l, int d) {
    const uint32_t w = ps_tnode_word(d, (l.y & PS_LEAF) != 0u, ps_child_count(l));
    return make_float4(g.x, g.y, g.z, __uint_as_float(w));
}

// P2M or M2M of node nidx at depth d, and its traversal record
__device__ __forceinline__ void upward_node(uint32_t nidx, int d, float4* __restrict__ tpk,
                                            const float4* __restrict__ pos4,
                                            const float4* __restrict__ geom,
                                            const uint2* __restrict__ link,
                                            const uint32_t* __restrict__ nend,
                                            float4* __restrict__ mp) {
#pragma clang fp contract(off)
    const float4 g = geom[nidx];
    const uint2 l = link[nidx];
    tpk[nidx] = trav_node(g, l, d);
    float m0 = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    if (l.y & PS_LEAF) {  // P2M about the cell centre, H:978-1004
        p2m_leaf(pos4, l.x, nend[nidx], g, m0, mx, my, mz);
    } else {  // M2M in slot order, H:1013-1039
        m2m_children(geom, mp, l, g, m0, mx, my, mz);
    }
    mp[nidx] = make_float4(m0, mx, my, mz);
}

__global__ __launch_bounds__(kBlock) void upward_level_kernel(uint32_t begin, uint32_t end, int d,
                                                              float4* __restrict__ tpk,
                                                              const float4* __restrict__ pos4,
                                                              const float4* __restrict__ geom,
                                                              const uint2* __restrict__ link,
                                                              const uint32_t* __restrict__ nend,
                                                              float4* __restrict__ mp) {
    const uint32_t nidx = begin + blockIdx.x * kBlock + threadIdx.x;
    if (nidx < end) upward_node(nidx, d, tpk, pos4, geom, link, nend, mp);
}

// Depths d_top .. 0 in one block, level by level (a barrier between levels:
// a node's children, one level down, are written before it reads them).
constexpr int kTopBlock = 1024;
struct PsLevels {
    uint32_t start[PS_MAX_DEPTH + 2];
};
__global__ __launch_bounds__(kTopBlock) void upward_top_kernel(PsLevels ls, int d_top,
                                                               float4* __restrict__ tpk,
                                                               const float4* __restrict__ pos4,
                                                               const float4* __restrict__ geom,
                                                               const uint2* __restrict__ link,
                                                               const uint32_t* __restrict__ nend,
                                                               float4* __restrict__ mp) {
    for (int d = d_top; d >= 0; --d) {
        for (uint32_t nidx = ls.start[d] + threadIdx.x; nidx < ls.start[d + 1]; nidx += kTopBlock)
            upward_node(nidx, d, tpk, pos4, geom, link, nend, mp);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void l2l_level_kernel(uint32_t begin, uint32_t end,
                                                           const uint32_t* __restrict__ parent,
                                                           float4* __restrict__ loc) {
    const uint32_t nidx = begin + blockIdx.x * kBlock + threadIdx.x;
    if (nidx >= end) return;
    float4 pl = loc[parent[nidx]];
    float4 cl = loc[nidx];
    cl.y = cl.y + pl.y;  // child_local += parent_local, H:1248-1250
    cl.z = cl.z + pl.z;
    cl.w = cl.w + pl.w;
    loc[nidx] = cl;
}

// L2L for the leaves alone, in one launch.  The level-by-level pass makes a
// node's local its own plus its parent's summed local, so a leaf's is the sum
// of the locals on its path from the root, added from the root down; this
// kernel adds them in that order (the bits agree with the level pass).  The
// path comes from the parent links (at most PS_MAX_DEPTH steps, all loads of
// the locals in flight at once); loc is only read, the leaves' sums go to
// ltot (internal nodes' entries are not written).
__global__ __launch_bounds__(kBlock) void leaf_locals_kernel(uint32_t n_nodes,
                                                             const uint2* __restrict__ link,
                                                             const uint32_t* __restrict__ parent,
                                                             const float4* __restrict__ loc,
                                                             float4* __restrict__ ltot) {
    const uint32_t nidx = blockIdx.x * kBlock + threadIdx.x;
    if (nidx >= n_nodes || !(link[nidx].y & PS_LEAF)) return;
    uint32_t a[PS_MAX_DEPTH];  // ancestors, the parent first; PS_NONE past the root
    uint32_t cur = nidx;
#pragma unroll
    for (int k = 0; k < PS_MAX_DEPTH; ++k) {
        cur = cur != PS_NONE ? parent[cur] : PS_NONE;
        a[k] = cur;
    }
    const float4 own = loc[nidx];
    float4 v[PS_MAX_DEPTH];
#pragma unroll
    for (int k = 0; k < PS_MAX_DEPTH; ++k)
        v[k] = a[k] != PS_NONE ? loc[a[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
    bool have = false;  // the root's own local starts the sum
    float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
    for (int k = PS_MAX_DEPTH - 1; k >= 0; --k) {
        if (a[k] == PS_NONE) continue;
        if (have) {
            sx = v[k].y + sx;
            sy = v[k].z + sy;
            sz = v[k].w + sz;
        } else {
            sx = v[k].y;
            sy = v[k].z;
            sz = v[k].w;
            have = true;
        }
    }
    float4 r = own;
    if (have) {
        r.y = own.y + sx;
        r.z = own.z + sy;
        r.w = own.w + sz;
    }
    ltot[nidx] = r;
}

// the same two passes with the level range read from device memory
// (ls[d], ls[d+1]) and a fixed grid: nothing depends on host-side sizes, so
// the launches can be recorded in a graph
__global__ __launch_bounds__(kBlock) void upward_level_dev_kernel(
    const uint32_t* __restrict__ ls, int d, float4* __restrict__ tpk, const float4* __restrict__ pos4,
    const float4* __restrict__ geom, const uint2* __restrict__ link,
    const uint32_t* __restrict__ nend, float4* __restrict__ mp) {
#pragma clang fp contract(off)
    const uint32_t begin = ls[d], end = ls[d + 1];
    for (uint32_t nidx = begin + blockIdx.x * kBlock + threadIdx.x; nidx < end;
         nidx += gridDim.x * kBlock) {
        const float4 g = geom[nidx];
        const uint2 l = link[nidx];
        tpk[nidx] = trav_node(g, l, d);
        float m0 = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
        if (l.y & PS_LEAF) {
            p2m_leaf(pos4, l.x, nend[nidx], g, m0, mx, my, mz);
        } else {
            m2m_children(geom, mp, l, g, m0, mx, my, mz);
        }
        mp[nidx] = make_float4(m0, mx, my, mz);
    }
}

__global__ __launch_bounds__(kBlock) void l2l_level_dev_kernel(const uint32_t* __restrict__ ls,
                                                               int d,
                                                               const uint32_t* __restrict__ parent,
                                                               float4* __restrict__ loc) {
    const uint32_t begin = ls[d], end = ls[d + 1];
    for (uint32_t nidx = begin + blockIdx.x * kBlock + threadIdx.x; nidx < end;
         nidx += gridDim.x * kBlock) {
        float4 pl = loc[parent[nidx]];
        float4 cl = loc[nidx];
        cl.y = cl.y + pl.y;
        cl.z = cl.z + pl.z;
        cl.w = cl.w + pl.w;
        loc[nidx] = cl;
    }
}

constexpr uint32_t kDevGrid = 2048;  // blocks of the device-range launches

}  // namespace

void ps_launch_upward_dev(const uint32_t* ls, float4* tpk, const float4* pos4, const float4* geom,
                          const uint2* link, const uint32_t* nend, float4* mp, hipStream_t s) {
    for (int d = PS_MAX_DEPTH; d >= 0; --d)
        upward_level_dev_kernel<<<kDevGrid, kBlock, 0, s>>>(ls, d, tpk, pos4, geom, link, nend,
                                                            mp);
}

void ps_launch_l2l_dev(const uint32_t* ls, const uint32_t* parent, float4* loc, hipStream_t s) {
    for (int d = 1; d <= PS_MAX_DEPTH; ++d)
        l2l_level_dev_kernel<<<kDevGrid, kBlock, 0, s>>>(ls, d, parent, loc);
}

size_t ps_tree_tmp_bytes(size_t n) {
    size_t nb = (n + kBlock - 1) / kBlock;
    if (nb == 0) nb = 1;
    size_t t = (size_t)kDepths * nb;
    return align256(t * sizeof(uint32_t)) + ps_scan_tmp_bytes_u32(t);
}

hipError_t ps_tree_count(const uint32_t* codes, const uint32_t* sys, size_t n,
                         uint32_t* level_start, uint32_t* flags, unsigned long long* total64,
                         void* tmp, int leaf_cap, hipStream_t s, const uint8_t* sys_cap) {
    const uint32_t nb = (uint32_t)((n + kBlock - 1) / kBlock);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(tmp);
    const size_t t = (size_t)kDepths * nb;
    void* scan_tmp = reinterpret_cast<char*>(tmp) + align256(t * sizeof(uint32_t));
    ps_fill_u32(flags, 0u, 1, s);
    tree_count_kernel<<<nb, kBlock, 0, s>>>(codes, sys, n, cnt, nb, flags, leaf_cap, sys_cap);
    hipError_t e = ps_scan_u32(cnt, cnt, t, total64, scan_tmp, s);
    if (e != hipSuccess) return e;
    tree_levels_kernel<<<1, 64, 0, s>>>(cnt, nb, total64, level_start);
    return hipGetLastError();
}

hipError_t ps_tree_write(const uint32_t* codes, const uint32_t* sys, size_t n, float4 root,
                         const float4* roots, void* tmp, float4* geom, uint2* link,
                         uint32_t* nend, uint32_t* parent, uint32_t* leaf_of, uint8_t* octant,
                         size_t n_nodes, const unsigned long long* n_nodes_dev, int leaf_cap,
                         hipStream_t s, const uint8_t* sys_cap, float4* loc_clear,
                         uint32_t* len_clear) {
    const uint32_t nb = (uint32_t)((n + kBlock - 1) / kBlock);
    const uint32_t* off = reinterpret_cast<const uint32_t*>(tmp);
    // no clearing of link: the write stores a leaf's whole word and an
    // internal node's first child, the mask pass its child mask
    const uint32_t cap = n_nodes < 0xFFFFFFFFull ? (uint32_t)n_nodes : 0xFFFFFFFFu;
    tree_write_kernel<<<nb, kBlock, 0, s>>>(codes, sys, n, off, nb, root, roots, geom, link,
                                            nend, parent, leaf_of, octant, leaf_cap, cap,
                                            sys_cap);
    if (n_nodes) {
        size_t mb = (n_nodes + kBlock - 1) / kBlock;
        if (mb > 4096) mb = 4096;
        tree_mask_kernel<<<(unsigned)mb, kBlock, 0, s>>>(parent, octant, link, n_nodes_dev,
                                                         (size_t)cap, loc_clear, len_clear);
    }
    return hipGetLastError();
}

void ps_launch_upward_top(const uint32_t* ls, int d_top, float4* tpk, const float4* pos4,
                          const float4* geom, const uint2* link, const uint32_t* nend, float4* mp,
                          hipStream_t s) {
    if (d_top < 0) return;
    PsLevels lv;
    for (int d = 0; d <= PS_MAX_DEPTH + 1; ++d) lv.start[d] = ls[d];
    upward_top_kernel<<<1, kTopBlock, 0, s>>>(lv, d_top, tpk, pos4, geom, link, nend, mp);
}

void ps_launch_upward_level(uint32_t begin, uint32_t end, int d, float4* tpk, const float4* pos4,
                            const float4* geom, const uint2* link, const uint32_t* nend,
                            float4* mp, hipStream_t s) {
    if (end <= begin) return;
    uint32_t blocks = (end - begin + kBlock - 1) / kBlock;
    upward_level_kernel<<<blocks, kBlock, 0, s>>>(begin, end, d, tpk, pos4, geom, link, nend, mp);
}

void ps_launch_leaf_locals(uint32_t n_nodes, const uint2* link, const uint32_t* parent,
                           const float4* loc, float4* ltot, hipStream_t s) {
    if (n_nodes == 0) return;
    const uint32_t blocks = (n_nodes + kBlock - 1) / kBlock;
    leaf_locals_kernel<<<blocks, kBlock, 0, s>>>(n_nodes, link, parent, loc, ltot);
}

void ps_launch_l2l_level(uint32_t begin, uint32_t end, const uint32_t* parent, float4* loc,
                         hipStream_t s) {
    if (end <= begin) return;
    uint32_t blocks = (end - begin + kBlock - 1) / kBlock;
    l2l_level_kernel<<<blocks, kBlock, 0, s>>>(begin, end, parent, loc);
}
