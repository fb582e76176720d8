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
// One iteration = one kernel, a group of G lanes (8 to 64; several groups per
// wavefront) per entry.  The grid is fixed (one block per CU by default); the
// entries are cut into parts, one per XCD (eight, or fewer for a small
// iteration), and each block takes segments (kStep / G to kSeg entries, sized
// on the device from the entry count) from its XCD's part -- the first one
// fixed by the block index, the others through an atomic counter -- then
// from the other parts once its own is used up.  Per segment, in LDS:
//   count  MAC for each listed source (the group takes the list in strides),
//          M2L for far ones, a 64-bit far mask per 64 list entries, the
//          entry's output sizes into LDS; a final leaf records (list, masks)
//          for P2P.
//   place  block prefix sum of the sizes; one atomic add per counter (list
//          words, next entries, mask words) reserves the segment's output,
//          checked against the arena capacities.
//   write  for continuing entries only: the near records the count left in
//          LDS (a near leaf's id, a near internal node's id and child count)
//          become runs of consecutive ids stored straight to the reserved
//          list space, at positions from a group prefix sum; an entry whose
//          records did not fit the LDS pool reads its list again (the far
//          mask picks the near sources, their links give the children): the
//          same words either way.  Next entries from the word the count
//          phase kept per entry (the target itself, or its first child and
//          child count).
// Where a segment's output lands depends on the order of the atomics; what
// each entry computes does not (its list order is fixed by its producer), so
// the results are deterministic.
// Lists are appended to one arena that lives for the whole tick, so the P2P
// kernel reads each leaf's final list in place.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;            // init kernel
constexpr int kStep = 1024;            // threads of an iteration block
constexpr uint32_t kSeg = 2048;        // entries per LDS segment (one reservation each)
constexpr uint32_t kOrdBins = 32;      // list-length bins of the segment order

// Segment length of an iteration of n_e entries over nblk blocks: about spb
// segments per block, a multiple of the groups per block up to kSeg (small
// iterations spread over many blocks: one entry per group).
__device__ __forceinline__ uint32_t trav_seg(uint32_t n_e, uint32_t nblk, uint32_t spb,
                                             uint32_t groups) {
    const unsigned long long d = (unsigned long long)spb * nblk;
    uint32_t s = (uint32_t)(((unsigned long long)n_e + d - 1) / d);
    s = (s + groups - 1) / groups * groups;
    return s < groups ? groups : (s > kSeg ? kSeg : s);
}

// First entry of part x of n_e entries cut into np parts (x = 0..np).
__device__ __forceinline__ uint32_t part_start(uint32_t n_e, uint32_t np, uint32_t x) {
    return (uint32_t)((unsigned long long)n_e * x / np);
}

// Segments past the first nbx of each part are handed out by its counter:
// segment c of part x starts at part_start(x) + (nbx + c) S.  Called by the
// lanes of one wavefront.  The block's own part is tried first with an
// atomic add alone (own_done once it gave nothing); then the lanes read the
// np counters at once and take a segment from the first part, in order from
// the block's own, that has one left (when another block took the last one
// first, that part reads as used up next time round, so the loop ends).
// Every lane gets the first entry (PS_NONE: no work left) and the length.
__device__ uint32_t trav_grab(PsTravState* st, int it, uint32_t n_e, uint32_t np, uint32_t S,
                              uint32_t nbx, uint32_t xcd, uint32_t lane, bool& own_done,
                              uint32_t& len) {
    const uint32_t x = (xcd + lane) % np;
    const uint32_t lo = part_start(n_e, np, x), hi = part_start(n_e, np, x + 1u);
    const unsigned long long first = (unsigned long long)lo + (unsigned long long)nbx * S;
    uint32_t got = PS_NONE, n = 0u;
    if (!own_done) {
        if (lane == 0 && first < hi) {
            const unsigned long long start = first + atomicAdd(&st->seg_ctr[it][x], S);
            if (start < hi) {
                got = (uint32_t)start;
                n = hi - got < S ? hi - got : S;
            }
        }
        got = __shfl(got, 0, 64);
        if (got != PS_NONE) {
            len = __shfl(n, 0, 64);
            return got;
        }
        own_done = true;
    }
    for (;;) {
        bool left = false;
        if (lane < np)
            left = first + __atomic_load_n(&st->seg_ctr[it][x], __ATOMIC_RELAXED) < hi;
        const uint64_t m = __ballot(left);
        if (m == 0) {
            len = 0u;
            return PS_NONE;
        }
        const uint32_t pick = (uint32_t)__ffsll((long long)m) - 1u;
        got = PS_NONE;
        if (lane == pick) {
            const unsigned long long start = first + atomicAdd(&st->seg_ctr[it][x], S);
            if (start < hi) {
                got = (uint32_t)start;
                n = hi - got < S ? hi - got : S;
            }
        }
        got = __shfl(got, (int)pick, 64);
        if (got != PS_NONE) {
            len = __shfl(n, (int)pick, 64);
            return got;
        }
    }
}

// Iteration-0 entries.  The level starts come from device memory.  MT (the
// reference's threaded build, H:1772-1801): entry k is node k for every node
// of depth <= 2, with the list {root of its system} when the node is at depth 2
// or a leaf above it, else empty.  ST: entry k is root k with the list {k}.
// Roots are nodes [0, n_sys) (one per system, in system order); arena word s
// holds root s.
__global__ __launch_bounds__(kBlock) void trav_init_kernel(const uint32_t* __restrict__ ls,
                                                           int mt, uint32_t n_sys,
                                                           const uint32_t* __restrict__ sys,
                                                           const uint32_t* __restrict__ nend,
                                                           const uint2* __restrict__ link,
                                                           PsTravEntries ent, PsTravState* st,
                                                           uint32_t* arena, uint32_t arena_cap,
                                                           uint32_t mask_cap) {
    const uint32_t n_e = mt ? ls[3] : n_sys;
    const uint32_t ls2 = ls[2];
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n_e || k < n_sys;
         k += gridDim.x * kBlock) {
        if (k == 0) {
            PsTravState z = {};
            z.n_ent[0] = n_e;
            z.arena_ctr = (n_sys + 3u) & ~3u;  // the root lists; every list starts at a multiple of 4
            z.mask_ctr = n_e;     // one mask word per entry
            z.arena_cap = arena_cap;
            z.mask_cap = mask_cap;
            *st = z;
        }
        if (k < n_sys) arena[k] = k;
        if (k >= n_e) continue;
        ent.tgt[k] = k;
        ent.moff[k] = k;
        if (mt) {
            ent.off[k] = sys ? sys[nend[k] - 1u] : 0u;
            ent.len[k] = (k >= ls2 || (link[k].y & PS_LEAF)) ? 1u : 0u;
        } else {
            ent.off[k] = k;
            ent.len[k] = 1u;
        }
    }
}

__device__ __forceinline__ uint32_t chunks_of(uint32_t len) { return (len + 63u) >> 6; }

// What a continuing target passes to the next iteration, packed by the
// count phase: id << 4 | k -- itself (id = the target, k = 0) for a leaf,
// its k children (id = the first child) for an internal node.
__device__ __forceinline__ uint32_t next_word(uint32_t tn, uint2 tl) {
    return (tl.y & PS_LEAF) ? tn << 4 : (tl.x << 4) | ps_child_count(tl);
}

// Near records: what a near source passes on to the next list, kept in LDS
// between the count and the write phase -- the source's id (bits 0..27), a
// flag for an internal node (bit 28) and the count - 1 (bits 29..31) of the
// consecutive ids it stands for (a leaf: itself; an internal node: its
// children, from its first child, which the write phase looks up).  Node
// ids stay below PS_TRAV_MAX_NODES <= 2^28.
// s_waitcnt immediate for vmcnt(0) alone (gfx9 fields: vmcnt 15:14 and 3:0,
// expcnt 6:4, lgkmcnt 11:8 -- the last two at their no-wait maximum)
constexpr int kWaitVmem0 = 0x0F70;
constexpr uint32_t kRecShift = 29;
constexpr uint32_t kRecInternal = 1u << 28;
constexpr uint32_t kRecMask = kRecInternal - 1u;
// An entry's records: (start << 16 | count) in its group's pool space, or
// kNoPiece when they did not fit (the write phase then reads the list again).
constexpr uint32_t kNoPiece = 0xFFFFFFFFu;
constexpr uint32_t kPoolWords = 24064;  // the block's record pool (94 KiB)

// Node record i from a node array: the byte offset is a 32-bit product, so
// the load takes the array's address from scalar registers and one 32-bit
// offset register (the host keeps a tick's nodes at most PS_TRAV_MAX_NODES).
template <class T>
__device__ __forceinline__ T node_ld(const T* __restrict__ base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       (size_t)(i * (uint32_t)sizeof(T)));
}

// Count of one entry by its group of G lanes (8, 16, 32 or 64).  The list is
// taken in chunks of 64 sources, one far-mask word each; lane gl holds the
// U = 64 / G sources k0 + j G + gl, whose list words and then node records
// (traversal record and multipole) are all requested before any is used, so a
// chunk costs two dependent memory round trips.  me = the entry (target,
// list offset, length, mask offset), staged in LDS by the segment's order
// pass.  Returns (list words, next entries, mask words, far pairs) of the
// entry's output, valid in every lane.
template <bool kKahan, int G>
__device__ __forceinline__ uint4 count_entry(const PsTravTree& t, float theta2, uint4 me,
                                             const uint32_t* __restrict__ arena,
                                             unsigned long long* __restrict__ masks,
                                             uint32_t mask_cap, uint32_t gl,
                                             uint32_t gshift, uint32_t* pool, uint32_t& bump,
                                             uint32_t pool_end, uint32_t& piece, uint32_t& nx) {
    constexpr int U = 64 / G;
    const uint32_t tn = me.x;
    const uint32_t o = me.y;
    const uint32_t L = me.z;
    const uint32_t mo = me.w;
    const float4 tg = t.geom[tn];
    const uint2 tl = t.link[tn];
    // the root's half width: the target's times 2^depth, exactly
    const float root_hw = __builtin_amdgcn_ldexpf(tg.w, -ps_tnode_nexp(__float_as_uint(t.tpk[tn].w)));
    float ex = 0.f, ey = 0.f, ez = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    uint32_t nfar = 0;
    uint32_t acc = 0;  // this lane's near sources: output ids | near internal nodes << 16
    for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        // slot j holds list entry k0 + j G + gl: in the list iff j G < rem
        const int rem = (int)(L - k0) - (int)gl;
        // slot j runs when some lane of the wavefront has a source in it
        // (a scalar test: the other groups' lists may be longer).  Every
        // running slot loads: past the list's end the word comes from the
        // next list or the arena's padding (PS_LIST_PAD words past its
        // capacity) and the slot then evaluates the target itself, which is
        // discarded -- no branches around the loads.
        bool need[U];
        unsigned long long inm[U];  // lanes with slot j in their list
#pragma unroll
        for (int j = 0; j < U; ++j) {
            inm[j] = __ballot(j * G < rem);
            // slot 0 always runs (a group still in the loop has a source in
            // it): its loads are then issued first and waited for alone
            need[j] = j == 0 || inm[j] != 0ull;
        }
        uint32_t s[U];
        const uint32_t* lp = arena + (size_t)(o + k0 + gl);
#pragma unroll
        for (int j = 0; j < U; ++j) s[j] = need[j] ? lp[j * G] : tn;
        // every list word is in before the node loads start (one wait here
        // instead of one per slot, ahead of each slot's loads)
        __builtin_amdgcn_s_waitcnt(kWaitVmem0);
        float4 sg[U], smp[U];  // traversal record (centre, depth and flags), multipole
#pragma unroll
        for (int j = 0; j < U; ++j) {
            if (!need[j]) continue;
            const uint32_t sv = j * G < rem ? s[j] : tn;
            sg[j] = node_ld(t.tpk, sv);
            smp[j] = node_ld(t.mp, sv);
        }
        unsigned long long word = 0;
        uint32_t nrec = 0;  // records of the chunk so far (uniform in the group)
#pragma unroll
        for (int j = 0; j < U; ++j) {
            if (!need[j]) continue;
            const bool in = j * G < rem;
            // the source's half width: the root's times 2^-depth, exactly
            const uint32_t pw = __float_as_uint(sg[j].w);
            const float4 sgj = make_float4(sg[j].x, sg[j].y, sg[j].z,
                                           __builtin_amdgcn_ldexpf(root_hw, ps_tnode_nexp(pw)));
            float dx, dy, dz, d2;
            const bool mac = ps_far_d(tg, sgj, theta2, dx, dy, dz, d2);
            const bool far = in && mac;
            // the wavefront's far and near masks straight from the compares
            // (a ballot of the combined flags would rebuild them lane by lane)
            const unsigned long long mmac = __ballot(mac);
            const unsigned long long mfar = inm[j] & mmac, mnear = inm[j] & ~mmac;
            if (kKahan) {
                if (far) {
                    float px = 0.f, py = 0.f, pz = 0.f;
                    ps_m2l_d(dx, dy, dz, d2, smp[j], px, py, pz);
                    ps_kahan_add(ex, cx, px);
                    ps_kahan_add(ey, cy, py);
                    ps_kahan_add(ez, cz, pz);
                }
            } else {  // no branch: most slots are far
                ps_m2l_dm(dx, dy, dz, d2, smp[j], far, ex, ey, ez);
            }
            const bool near = in && !mac;
            acc += near ? (pw >> PS_TNODE_OUT_SHIFT) : 0u;  // summed over the group below
            word |= group_bits_of<G>(mfar, gshift) << (j * G);
            // the record of a near source, in list order after the chunk's
            // earlier ones: a near leaf passes on itself, a near internal
            // node its children (consecutive ids from its first child)
            const unsigned long long nb = group_bits_of<G>(mnear, gshift);
            const uint32_t pos = bump + nrec + (uint32_t)__popcll(nb & ((1ull << gl) - 1ull));
            if (near && pos < pool_end) pool[pos] = s[j] | ps_tnode_rec_bits(pw);
            nrec += (uint32_t)__popcll(nb);
        }
        if (gl == 0 && mo + (k0 >> 6) < mask_cap) masks[mo + (k0 >> 6)] = word;
        nfar += (uint32_t)__popcll(word);
        if (piece != kNoPiece) {
            if (bump + nrec <= pool_end) {
                bump += nrec;
            } else {  // no room: the write phase reads the list again
                bump = piece;  // the entry's start
                piece = kNoPiece;
            }
        }
    }
    acc = group_sum_u32<G>(acc);
    const uint32_t out = acc & 0xFFFFu;  // next list words
    const uint32_t nint = acc >> 16;      // near internal sources
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
    uint4 c = make_uint4(0u, 0u, 0u, nfar);
    nx = next_word(tn, tl);
    if (tl.y & PS_LEAF) {
        if (nint == 0) {  // final: the near entries of this list are the P2P sources
            if (piece != kNoPiece) bump = piece;  // its records are not needed
            if (out && gl == 0) {
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
    if (gl == 0 && nfar) {
        float* lp3 = reinterpret_cast<float*>(t.loc + tn) + 1;  // (Ex, Ey, Ez)
        const float3 lt = make_float3(lp3[0], lp3[1], lp3[2]);
        lp3[0] = lt.x + ex;
        lp3[1] = lt.y + ey;
        lp3[2] = lt.z + ez;
    }
    return c;
}

// The next iteration's entries of a continuing target (its next_word nx),
// all with the next list at p.x of length out.
template <int G>
__device__ __forceinline__ void write_next(const PsTravEntries& next, uint32_t nx, uint32_t out,
                                           uint4 p, uint32_t gl) {
    const uint32_t nch = chunks_of(out);
    const uint32_t id = nx >> 4, nk = nx & 15u;
    if (nk == 0) {
        if (gl == 0) {
            next.tgt[p.y] = id;
            next.off[p.y] = p.x;
            next.len[p.y] = out;
            next.moff[p.y] = p.z;
        }
    } else {
        for (uint32_t r = gl; r < nk; r += G) {
            next.tgt[p.y + r] = id + r;
            next.off[p.y + r] = p.x;
            next.len[p.y + r] = out;
            next.moff[p.y + r] = p.z + r * nch;
        }
    }
}

// Write of one continuing entry by its group: out = its next list length
// (> 0), p = (list offset, next entry index, mask word offset) in the arenas.
// Per chunk of 64 sources: the mask word and the U list words of each lane
// in one round trip, the links of the near ones in a second; outputs keep
// the list order (a prefix sum over the group per j).
template <int G>
__device__ __forceinline__ void write_entry(const PsTravTree& t, const PsTravEntries& ent,
                                            uint32_t* __restrict__ arena,
                                            const unsigned long long* __restrict__ masks,
                                            uint32_t e, uint32_t nx, uint32_t out, uint4 p,
                                            const PsTravEntries& next, uint32_t gl) {
    constexpr int U = 64 / G;
    const uint32_t o = ent.off[e];
    const uint32_t L = ent.len[e];
    const uint32_t mo = ent.moff[e];
    // the count pass's masks and link reads give exactly `out` outputs; the
    // bound only makes that an invariant of the stores
    uint32_t base = p.x;
    const uint32_t limit = base + out;
    for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        const unsigned long long fm = masks[mo + (k0 >> 6)];
        bool need[U];  // as in count_entry
#pragma unroll
        for (int j = 0; j < U; ++j) need[j] = __ballot(k0 + j * G + gl < L) != 0ull;
        uint32_t s[U];
        const uint32_t* lp = arena + (size_t)(o + k0 + gl);
#pragma unroll
        for (int j = 0; j < U; ++j) s[j] = need[j] ? lp[j * G] : PS_NONE;
        uint2 sl[U];
        bool nr[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            // near (a clear far bit) implies a valid id: invalid ones carry a far bit
            nr[j] = k0 + j * G + gl < L && !((fm >> (j * G + gl)) & 1ull) && s[j] < t.n_nodes;
            sl[j] = make_uint2(0u, 0u);
            if (nr[j]) sl[j] = node_ld(t.link, s[j]);  // the near sources only
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint32_t nout =
                !nr[j] ? 0u : ((sl[j].y & PS_LEAF) ? 1u : (uint32_t)__popc(sl[j].y & 0xFFu));
            uint32_t total;
            const uint32_t inc = group_scan_u32<G>(nout, gl, total);
            const uint32_t w = base + inc - nout;
            // a near leaf is one word (itself), a near internal node its
            // nout <= 8 children, consecutive ids from its first child
            if (nout && w + nout <= limit) {
                uint32_t* dst = arena + w;
                const uint32_t v0 = (sl[j].y & PS_LEAF) ? s[j] : sl[j].x;
#pragma unroll
                for (uint32_t r = 0; r < 8; ++r)
                    if (r < nout) dst[r] = v0 + r;
            }
            base += total;
        }
    }
    write_next<G>(next, nx, out, p, gl);
}

// Stores of a run of consecutive ids v0 .. v0 + n - 1 (n <= 8) from word
// dst: 16-, 8- and 4-byte stores (dst is only 4-byte aligned)
struct __attribute__((aligned(4))) PsRun4 {
    uint32_t a, b, c, d;
};
struct __attribute__((aligned(4))) PsRun2 {
    uint32_t a, b;
};
__device__ __forceinline__ void store_run(uint32_t* dst, uint32_t v0, uint32_t n) {
    uint32_t r = 0;
    if (n >= 4u) {
        *reinterpret_cast<PsRun4*>(dst) = PsRun4{v0, v0 + 1u, v0 + 2u, v0 + 3u};
        r = 4u;
    }
    if (n >= 8u) {
        *reinterpret_cast<PsRun4*>(dst + 4) = PsRun4{v0 + 4u, v0 + 5u, v0 + 6u, v0 + 7u};
        r = 8u;
    }
    if (n - r >= 2u) {
        *reinterpret_cast<PsRun2*>(dst + r) = PsRun2{v0 + r, v0 + r + 1u};
        r += 2u;
    }
    if (n > r) dst[r] = v0 + r;
}

// Write of one continuing entry from its near records in LDS (count of
// them: n), the same list words in the same order as write_entry: per round
// each lane takes two records (their first ids: a leaf itself, an internal
// node's first child, looked up together) and stores their ids straight to
// the list, its positions from a group prefix sum -- no staging in LDS.
template <int G>
__device__ __forceinline__ void write_entry_rec(const PsTravTree& t, uint32_t* __restrict__ arena,
                                                uint32_t* recs, uint32_t n, uint32_t nx,
                                                uint32_t out, uint4 p, const PsTravEntries& next,
                                                uint32_t gl) {
    uint32_t base = p.x;
    for (uint32_t r0 = 0; r0 < n; r0 += 2u * G) {
        // two records per lane: ids and counts, internal ones resolved
        uint32_t rv[2], v0[2], nout[2];
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) {
            const uint32_t i = r0 + q * G + gl;
            rv[q] = i < n ? recs[i] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) {
            const uint32_t i = r0 + q * G + gl;
            nout[q] = i < n ? (rv[q] >> kRecShift) + 1u : 0u;
            const uint32_t id = rv[q] & kRecMask;
            v0[q] = (rv[q] & kRecInternal) ? node_ld(t.link, id).x : id;
        }
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) {
            uint32_t total;
            const uint32_t inc = group_scan_u32<G>(nout[q], gl, total);
            if (nout[q]) store_run(arena + base + inc - nout[q], v0[q], nout[q]);
            base += total;
        }
    }
    write_next<G>(next, nx, out, p, gl);
}

// next item of a block-local work queue, the same value in every lane of the group
template <int G>
__device__ __forceinline__ uint32_t group_take(uint32_t* q, uint32_t gl) {
    uint32_t v = 0;
    if (gl == 0) v = atomicAdd(q, 1u);
    return __shfl(v, 0, G);
}

template <bool kKahan, int G>
__global__ __launch_bounds__(kStep) void trav_step_kernel(
    PsTravTree t, float theta2, PsTravEntries ent, uint32_t* __restrict__ arena,
    unsigned long long* __restrict__ masks, PsTravState* __restrict__ st, int it,
    PsTravEntries next, unsigned long long* __restrict__ far_part, uint32_t spb,
    uint32_t rec_cap) {
    constexpr uint32_t kGroups = kStep / G;
    __shared__ uint4 scnt[kSeg];  // the segment's entries, then their sizes, then their offsets
    __shared__ uint4 wtot[kStep / 64];
    __shared__ uint32_t sbase[4];  // list, entry and mask bases; 1 = reserved
    __shared__ uint32_t queue[2];  // work queues of the count and write phases
    __shared__ unsigned long long sfar[kGroups];
    __shared__ uint32_t shist[kOrdBins];         // entries per length bin, then bin offsets
    __shared__ uint16_t sord[kSeg];              // the segment's entries, longest lists first
    __shared__ uint32_t sseg[2];                 // the next segment: first entry, length
    __shared__ uint32_t spool[kPoolWords];       // near records, kPoolWords / kGroups per group
    __shared__ uint32_t spc[kSeg];               // per entry: its records (or kNoPiece)
    __shared__ uint32_t snx[kSeg];               // per entry: what it passes on (next_word)
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t grp = tid / G;
    const uint32_t gl = tid & (G - 1);
    const uint32_t gshift = lane & ~(uint32_t)(G - 1);
    // Blocks go round-robin over the 8 XCDs: the blocks of XCD x take
    // segments of the x-th part of the entries first, so neighbouring
    // targets share an L2, then help with the other parts (the work per
    // entry varies with the density of the region).  An iteration of few
    // segments keeps to few XCDs: it is a chain of memory round trips, and
    // its input was just written through those L2s.
    const uint32_t nblk = gridDim.x;
    const uint32_t xcd = blockIdx.x & 7u;
    // another block may raise the flag meanwhile: the traversal is then redone
    const uint32_t ovf = __atomic_load_n(&st->overflow, __ATOMIC_RELAXED);
    const uint32_t n_e = ovf ? 0u : st->n_ent[it];
    const uint32_t arena_cap = st->arena_cap;
    const uint32_t mask_cap = st->mask_cap;
    const uint32_t S = trav_seg(n_e, nblk, spb, kGroups);
    const uint32_t n_seg = (uint32_t)(((unsigned long long)n_e + S - 1) / S);
    const uint32_t per_x = nblk >= 8u ? nblk >> 3 : 1u;  // blocks per XCD
    uint32_t np = (n_seg + per_x - 1) / per_x;
    np = np < 1u ? 1u : (np > 8u ? 8u : np);
    if (np > nblk) np = nblk;  // every part has the block of its first segment
    uint32_t nbx = (n_seg + np - 1) / np;  // segments of a part fixed by the block index
    nbx = nbx < per_x ? nbx : per_x;
    bool own_done = false;  // of wave 0: the own part has no segment left
    unsigned long long far_acc = 0;
    if (wave == 0 && xcd < np && (blockIdx.x >> 3) < nbx) {
        const uint32_t lo = part_start(n_e, np, xcd), hi = part_start(n_e, np, xcd + 1u);
        const unsigned long long start =
            (unsigned long long)lo + (unsigned long long)(blockIdx.x >> 3) * S;
        uint32_t first = PS_NONE, len = 0u;
        if (start < hi) {
            first = (uint32_t)start;
            len = hi - first < S ? hi - first : S;
        } else {
            first = trav_grab(st, it, n_e, np, S, nbx, xcd, lane, own_done, len);
        }
        if (lane == 0) {
            sseg[0] = first;
            sseg[1] = len;
        }
    } else if (tid == 0) {
        sseg[0] = PS_NONE;
    }
    if (tid == 0) {
        queue[0] = 0u;
        queue[1] = 0u;
    }
    if (tid < kOrdBins) shist[tid] = 0u;
    __syncthreads();
    for (;;) {
        const uint32_t seg = sseg[0], segn = sseg[1];
        if (seg == PS_NONE) break;  // uniform per block
        if (tid == 0) queue[1] = 0u;  // the write phase of the last segment is over
        // order: a wavefront runs as many slot rows (G sources each, one per
        // lane of a group) as the longest list among its groups needs, so
        // the queues hand out the entries longest first, in bins of G
        // sources (the first bin: longer still).  Which group takes an
        // entry, and in which order a bin's entries come, changes nothing in
        // what an entry computes.
        if (segn <= kGroups) {
            // at most one entry per group: they all run at once, so the
            // entries keep their order (no histogram, two barriers fewer)
            if (tid < segn) {
                scnt[tid] = make_uint4(ent.tgt[seg + tid], ent.off[seg + tid],
                                       ent.len[seg + tid], ent.moff[seg + tid]);
                sord[tid] = (uint16_t)tid;
            }
            __syncthreads();
        } else {
            uint32_t rk[kSeg / kStep];  // bin << 16 | rank in the bin
#pragma unroll
            for (uint32_t r = 0; r < kSeg / kStep; ++r) {
                const uint32_t j = r * kStep + tid;
                rk[r] = PS_NONE;
                if (j < segn) {
                    // the entry itself goes to LDS for the count phase
                    // (coalesced here, where the count phase would gather it
                    // per entry)
                    const uint4 me = make_uint4(ent.tgt[seg + j], ent.off[seg + j],
                                                ent.len[seg + j], ent.moff[seg + j]);
                    scnt[j] = me;
                    const uint32_t rows = (me.z + (uint32_t)G - 1u) / (uint32_t)G;
                    const uint32_t b =
                        kOrdBins - 1u - (rows < kOrdBins - 1u ? rows : kOrdBins - 1u);
                    rk[r] = (b << 16) | atomicAdd(&shist[b], 1u);
                }
            }
            __syncthreads();
            if (wave == 0) {  // exclusive prefix over the bins
                const uint32_t v = lane < kOrdBins ? shist[lane] : 0u;
                uint32_t tot;
                const uint32_t ex = ps_wave_excl_scan(v, &tot);
                if (lane < kOrdBins) shist[lane] = ex;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t r = 0; r < kSeg / kStep; ++r)
                if (rk[r] != PS_NONE)
                    sord[shist[rk[r] >> 16] + (rk[r] & 0xFFFFu)] = (uint16_t)(r * kStep + tid);
            __syncthreads();
            if (tid < kOrdBins) shist[tid] = 0u;  // for the next segment (barriers before its use)
        }
        // count
        constexpr uint32_t kPoolPer = kPoolWords / kGroups;
        // the group's pool space: kPoolPer words from grp * kPoolPer (rec_cap
        // < kPoolPer only to test the write from the lists)
        uint32_t bump = grp * kPoolPer;
        const uint32_t pool_end = bump + (rec_cap < kPoolPer ? rec_cap : kPoolPer);
        for (uint32_t i = group_take<G>(&queue[0], gl); i < segn; i = group_take<G>(&queue[0], gl)) {
            const uint32_t k = sord[i];
            uint32_t piece = bump, nx = 0;
            const uint4 c = count_entry<kKahan, G>(t, theta2, scnt[k], arena, masks, mask_cap, gl,
                                                   gshift, spool, bump, pool_end, piece, nx);
            if (gl == 0) {
                scnt[k] = c;
                spc[k] = piece == kNoPiece ? kNoPiece : (piece | ((bump - piece) << 16));
                snx[k] = nx;
            }
            far_acc += c.w;
        }
        __syncthreads();
        // place: exclusive prefix of (list words, entries, mask words) in
        // place, the entry's list length kept in .w
        uint4 carry = make_uint4(0u, 0u, 0u, 0u);  // same value in every thread
        for (uint32_t j0 = 0; j0 < segn; j0 += kStep) {
            const uint32_t j = j0 + tid;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (j < segn) v = scnt[j];
            const uint32_t vx = (v.x + 3u) & ~3u;  // every list starts 16-byte aligned
            const uint4 inc =
                make_uint4(ps_wave_incl_scan(vx), ps_wave_incl_scan(v.y), ps_wave_incl_scan(v.z), 0u);
            if (lane == 63) wtot[wave] = inc;
            __syncthreads();
            uint4 run = carry, all = carry;
#pragma unroll
            for (uint32_t w = 0; w < kStep / 64; ++w) {
                const uint4 tw = wtot[w];
                if (w < wave) {
                    run.x += tw.x; run.y += tw.y; run.z += tw.z;
                }
                all.x += tw.x; all.y += tw.y; all.z += tw.z;
            }
            if (j < segn)
                scnt[j] = make_uint4(run.x + inc.x - vx, run.y + inc.y - v.y,
                                     run.z + inc.z - v.z, v.x);
            carry = all;
            __syncthreads();  // wtot is rewritten by the next round
        }
        if (wave == 0) {  // reserve the segment's output: three atomics in flight at once
            unsigned long long b = 0;
            if (carry.y) {
                if (lane == 0) b = atomicAdd(&st->arena_ctr, (unsigned long long)carry.x);
                if (lane == 1) b = atomicAdd(&st->mask_ctr, (unsigned long long)carry.z);
                if (lane == 2) b = atomicAdd(&st->n_ent[it + 1], carry.y);
            }
            uint32_t nlen = 0u;  // the next segment, asked for beside the reservation
            const uint32_t nseg = trav_grab(st, it, n_e, np, S, nbx, xcd, lane, own_done, nlen);
            const unsigned long long ba = __shfl(b, 0, 64);
            const unsigned long long bm = __shfl(b, 1, 64);
            const unsigned long long be = __shfl(b, 2, 64);
            if (lane == 0) {  // every thread has read sseg (barriers since)
                sseg[0] = nseg;
                sseg[1] = nlen;
            }
            if (lane == 0) {
                const uint32_t ok =
                    (carry.y && ba + carry.x <= arena_cap && bm + carry.z <= mask_cap) ? 1u : 0u;
                if (carry.y && !ok) __atomic_store_n(&st->overflow, 1u, __ATOMIC_RELAXED);
                sbase[0] = (uint32_t)ba;
                sbase[1] = (uint32_t)be;
                sbase[2] = (uint32_t)bm;
                sbase[3] = ok;
                queue[0] = 0u;  // the count phase is over
            }
        }
        __syncthreads();
        // write
        if (sbase[3]) {
            const uint32_t la = sbase[0], le = sbase[1], lm = sbase[2];
            for (uint32_t i = group_take<G>(&queue[1], gl); i < segn;
                 i = group_take<G>(&queue[1], gl)) {
                const uint32_t k = sord[i];
                const uint4 q = scnt[k];
                if (q.w == 0) continue;  // final or empty: nothing to pass on
                const uint4 pl = make_uint4(la + q.x, le + q.y, lm + q.z, 0u);
                const uint32_t pc = spc[k];
                const uint32_t nx = snx[k];
                if (pc != kNoPiece)
                    write_entry_rec<G>(t, arena, spool + (pc & 0xFFFFu), pc >> 16, nx, q.w, pl,
                                       next, gl);
                else
                    write_entry<G>(t, ent, arena, masks, seg + k, nx, q.w, pl, next, gl);
            }
        }
        __syncthreads();  // scnt and sbase are rewritten by the next segment
    }
    if (gl == 0) sfar[grp] = far_acc;
    __syncthreads();
    if (tid == 0) {
        unsigned long long f = 0;
#pragma unroll
        for (uint32_t g = 0; g < kGroups; ++g) f += sfar[g];
        far_part[blockIdx.x] = f;
    }
}

}  // namespace

void ps_launch_trav_init(const uint32_t* ls, int mt, uint32_t n_sys, const uint32_t* sys,
                         const uint32_t* nend, const uint2* link, PsTravEntries ent,
                         PsTravState* st, uint32_t* arena, uint32_t arena_cap, uint32_t mask_cap,
                         uint32_t max_ent, hipStream_t s) {
    uint32_t n = max_ent > n_sys ? max_ent : n_sys;
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    trav_init_kernel<<<blocks, kBlock, 0, s>>>(ls, mt, n_sys, sys, nend, link, ent, st, arena,
                                               arena_cap, mask_cap);
}

namespace {
int lane_group(int lanes) { return (lanes == 8 || lanes == 32 || lanes == 64) ? lanes : 16; }
}  // namespace

uint32_t ps_trav_blocks(uint32_t max_ent, int lanes, uint32_t max_blocks) {
    const uint32_t groups = kStep / lane_group(lanes);
    const uint32_t b = (uint32_t)(((unsigned long long)max_ent + groups - 1) / groups);
    return b < 1u ? 1u : (b < max_blocks ? b : max_blocks);
}

void ps_launch_trav_step(PsTravTree t, float theta2, PsTravEntries ent, uint32_t* arena,
                         unsigned long long* masks, PsTravState* st, int k, uint32_t max_ent,
                         PsTravEntries next, unsigned long long* far_part, int kahan, int lanes,
                         uint32_t max_blocks, uint32_t spb, uint32_t rec_cap, hipStream_t s) {
    const uint32_t nb = ps_trav_blocks(max_ent, lanes, max_blocks);
#define PS_STEP(KH, G)                                                                    \
    trav_step_kernel<KH, G><<<nb, kStep, 0, s>>>(t, theta2, ent, arena, masks, st, k, next, \
                                                  far_part, spb, rec_cap)
    switch (lane_group(lanes) + (kahan ? 1 : 0)) {
        case 8: PS_STEP(false, 8); break;
        case 9: PS_STEP(true, 8); break;
        case 32: PS_STEP(false, 32); break;
        case 33: PS_STEP(true, 32); break;
        case 64: PS_STEP(false, 64); break;
        case 65: PS_STEP(true, 64); break;
        case 17: PS_STEP(true, 16); break;
        default: PS_STEP(false, 16); break;
    }
#undef PS_STEP
}
