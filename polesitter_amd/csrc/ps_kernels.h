// ps_kernels.h -- host-side launchers of the gfx950 kernels (one .hip file
// per stage).  All launchers are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// ---- ps_scan.hip: device-wide exclusive prefix sums ------------------------
// out[i] = in[0] + ... + in[i-1] (component-wise for uint4); in == out is
// allowed.  When total64 is given it is overwritten with the exact 64-bit sum
// of every component (1 value for uint32, 4 for uint4).
size_t ps_scan_tmp_bytes_u32(size_t n);
size_t ps_scan_tmp_bytes_u4(size_t n);
hipError_t ps_scan_u32(const uint32_t* in, uint32_t* out, size_t n,
                       unsigned long long* total64, void* tmp, hipStream_t s);
hipError_t ps_scan_u4(const uint4* in, uint4* out, size_t n, unsigned long long* total64,
                      void* tmp, hipStream_t s);
// the same with the element count read from device memory at run time
// (n_max: host-side upper bound that sizes the launch and tmp)
size_t ps_scan_dev_tmp_bytes_u4(size_t n_max);
size_t ps_scan_dev_tmp_bytes_u32(size_t n_max);
hipError_t ps_scan_dev_u4(const uint4* in, uint4* out, const uint32_t* n_dev, size_t n_max,
                          unsigned long long* total64, void* tmp, hipStream_t s);
hipError_t ps_scan_dev_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev,
                           size_t n_max, unsigned long long* total64, void* tmp, hipStream_t s);
// graph-safe fills (no memset nodes)
void ps_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// p[0 .. per * min(*count, cap)) = v with the count in device memory
void ps_fill_u32_count(uint32_t* p, uint32_t v, const unsigned long long* count, uint32_t per,
                       size_t cap, hipStream_t s);
// p[0 .. per * m) = v and p2[0 .. per2 * m) = v, m = min(*count, cap): one launch
void ps_fill2_u32_count(uint32_t* p, uint32_t per, uint32_t* p2, uint32_t per2, uint32_t v,
                        const unsigned long long* count, size_t cap, hipStream_t s);
// *out = sum of in[0..n) (one block: stored; more: cleared by a kernel, then
// one atomic add per block)
void ps_reduce_u64(const unsigned long long* in, size_t n, unsigned long long* out,
                   hipStream_t s);

// ---- ps_prepare.hip: bounding cube + Morton codes, H:1911-2170 --------------
// prm (device, 4 floats): min, max, range, 1023/range.  part: >= 1024 float2.
// The whole prepare pass in one block for a small n (prm and codes as
// ps_launch_bbox + ps_launch_morton); false (nothing launched) above it.
bool ps_launch_prepare_small(const float* x, const float* y, const float* z, size_t n,
                             float* prm, uint32_t* codes, int mode, hipStream_t s);
void ps_launch_bbox(const float* x, const float* y, const float* z, size_t n, float2* part,
                    float* prm, hipStream_t s);
// mode 0: scalar truncation for every particle (H:2132-2157);
// mode 1: AVX2 rounding for the first 8*floor(n/8), scalar tail (H:2040-2072)
void ps_launch_morton(const float* x, const float* y, const float* z, size_t n,
                      const float* prm, uint32_t* codes, int mode, hipStream_t s);
// The same two passes with no allocation: the scratch is a static device
// array of PS_PREP_SCRATCH_BYTES in ps_prepare.hip (its per-device address
// from ps_prepare_scratch).  x, y, z, codes: device arrays; out3 (host) =
// min, max, range.  Synchronises s.  Calls that overlap on one device share
// the array and race.
#define PS_PREP_SCRATCH_BYTES 8448
hipError_t ps_prepare_scratch(void** out);
hipError_t ps_prepare_noalloc(const float* x, const float* y, const float* z, size_t n,
                              uint32_t* codes, int mode, float* out3, hipStream_t s);
// Many systems in one set of arrays, system k = particles [off[k], off[k+1])
// (device array, every system non-empty): per system its cube prm[k] = (min,
// max, range, 1023/range) and root cell root[k] = (min + range/2 (x3),
// range/2); sys[i] = system of particle i; codes with each system's cube.
void ps_launch_prepare_seg(const float* x, const float* y, const float* z, size_t n,
                           const uint32_t* off, uint32_t n_sys, float4* prm, float4* root,
                           uint32_t* sys, uint32_t* codes, int mode, hipStream_t s);

// ---- ps_sort.hip: stable LSD radix sort (3 x 10-bit digits), H:1588-1693 ----
struct PsSortArrays {
    uint32_t* key;
    uint32_t* id;
    float* x;
    float* y;
    float* z;
    float* m;
    uint32_t* aux;  // optional second key (system id, < 2^10), null when unused
};
size_t ps_sort_tmp_bytes(size_t n);
// Sorts `io` in place by key bits 0..29 (keys must be < 2^30), moving the
// payload; t0/t1 are two scratch sets of n elements.  pos4 (optional) gets the
// sorted (x, y, z, m).  With aux set (in all three sets) a fourth pass sorts
// stably by aux: the order is (aux, key).
// kcopy (optional): a second copy of the sorted keys.  Payload arrays x
// (and with it y, z, m) may be null in all three sets: keys and id only.
// Without aux, n <= 4096 is one block (the same permutation; t0, t1 and tmp
// untouched).
hipError_t ps_radix_sort(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                         size_t n, void* tmp, hipStream_t s, uint32_t* kcopy = nullptr);
// Many systems at once, each a contiguous segment [off[s], off[s + 1]) of
// the arrays (s < n_sys, off[0] = 0, off[n_sys] = n): each segment sorted by
// key bits 0..29 in place, the same permutation as ps_radix_sort with the
// segment index as a second key, in three passes.  The tiles are cut at the
// segments' ends: ps_sort_seg_tiles fills (host arrays) vstart[0..T] (tile
// starts, then n), vseg[0..T) (a tile's segment) and seg_tile[0..n_sys]
// (a segment's first tile, then T), and returns T (with null arrays: only T).
struct PsSortSegs {
    const uint32_t* off;       // device, n_sys + 1
    uint32_t n_sys;
    const uint32_t* vstart;    // device, n_tiles + 1
    const uint32_t* vseg;      // device, n_tiles
    const uint32_t* seg_tile;  // device, n_sys + 1
    uint32_t n_tiles;
};
size_t ps_sort_seg_tiles(const uint32_t* off, uint32_t n_sys, uint32_t* vstart, uint32_t* vseg,
                         uint32_t* seg_tile);
size_t ps_sort_seg_tmp_bytes(size_t n_tiles, uint32_t n_sys);
hipError_t ps_radix_sort_seg(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                             size_t n, PsSortSegs segs, void* tmp, hipStream_t s);
// The incremental sort.  prev: the keys the last sort produced, in place
// (sorted).  Step 1 (count) flags the positions whose key (low 30 bits)
// differs from prev, writes their keys and positions in order to ckey and
// cpos and their number to *count (device).  Step 2 (finish, nc = that
// number) sorts them by key, merges them with the unchanged ones by (key,
// position) -- the stable order, the same permutation as ps_radix_sort --
// into t0 and copies the result back to io; pos4 and kcopy as in
// ps_radix_sort.  With s_copy and the two events set, the copy back runs on
// s_copy after ev_sorted (recorded on s) and records ev_copied: work after
// it on s may read t0 (the merged keys, pos4) but not io until it waits for
// ev_copied.  tmp: ps_sort_tmp_bytes(n), untouched between the steps;
// ckey and cpos must not be any of the t0, t1 key or id arrays, nor t1.z
// and t1.m (the packed changed keys, 8 bytes each, from t1.z on).
size_t ps_sort_delta_tmp_bytes(size_t n);  // the flag words' share of ps_sort_tmp_bytes
hipError_t ps_sort_delta_count(const uint32_t* key, const uint32_t* prev, size_t n, void* tmp,
                               uint32_t* ckey, uint32_t* cpos, unsigned long long* count,
                               hipStream_t s);
hipError_t ps_sort_delta_finish(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                                size_t n, void* tmp, const uint32_t* prev, uint32_t* ckey,
                                uint32_t* cpos, uint32_t nc, uint32_t* kcopy, hipStream_t s,
                                hipStream_t s_copy = nullptr, hipEvent_t ev_sorted = nullptr,
                                hipEvent_t ev_copied = nullptr);

// ---- ps_tree.hip: octree build, upward and L2L passes ------------------------
size_t ps_tree_tmp_bytes(size_t n);
// Phase 1: per-depth node counts.  Writes level_start[0..11] (device) and
// flags[0] |= 1 if any key has bits 30/31 set.  total64[0] = node count.
// sys (optional): the system of each particle (sorted by (system, code));
// each system then gets its own root, in system order at depth 0.
// leaf_cap: a node above depth 10 is a leaf iff it holds <= leaf_cap
// particles (the reference's rule is leaf_cap = 1); with sys and sys_cap
// set, system k uses sys_cap[k] instead (device array, 1..64 per system).
hipError_t ps_tree_count(const uint32_t* codes, const uint32_t* sys, size_t n,
                         uint32_t* level_start, uint32_t* flags, unsigned long long* total64,
                         void* tmp, int leaf_cap, hipStream_t s,
                         const uint8_t* sys_cap = nullptr);
// Phase 2: writes every node (needs phase 1's tmp untouched).  The root cell
// is `root`, or roots[sys[i]] per system when roots is set.
// n_nodes_dev (optional): the node count in device memory (then n_nodes is
// only the capacity).  Nodes at or past n_nodes are not written.
// octant: one byte per node (the node's slot in its parent), scratch for the
// child masks, which are stored without atomics
hipError_t ps_tree_write(const uint32_t* codes, const uint32_t* sys, size_t n, float4 root,
                         const float4* roots, void* tmp, float4* geom, uint2* link,
                         uint32_t* nend, uint32_t* parent, uint32_t* leaf_of, uint8_t* octant,
                         size_t n_nodes, const unsigned long long* n_nodes_dev, int leaf_cap,
                         hipStream_t s, const uint8_t* sys_cap = nullptr, float4* loc_clear = nullptr,
                         uint32_t* len_clear = nullptr);
// Upward (depths 10..0) and L2L (depths 1..10) with the level starts read
// from device memory ls[0..11]; fixed grids, no host-side sizes.  The upward
// pass also writes each node's traversal record tpk (ps_tnode_word).
void ps_launch_upward_dev(const uint32_t* ls, float4* tpk, const float4* pos4, const float4* geom,
                          const uint2* link, const uint32_t* nend, float4* mp, hipStream_t s);
void ps_launch_l2l_dev(const uint32_t* ls, const uint32_t* parent, float4* loc, hipStream_t s);
// Depths d_top .. 0 in one launch of one block (small levels; ls: the level
// starts on the host, PS_MAX_DEPTH + 2 entries).
void ps_launch_upward_top(const uint32_t* ls, int d_top, float4* tpk, const float4* pos4,
                          const float4* geom, const uint2* link, const uint32_t* nend, float4* mp,
                          hipStream_t s);
void ps_launch_upward_level(uint32_t begin, uint32_t end, int d, float4* tpk, const float4* pos4,
                            const float4* geom, const uint2* link, const uint32_t* nend,
                            float4* mp, hipStream_t s);
void ps_launch_l2l_level(uint32_t begin, uint32_t end, const uint32_t* parent, float4* loc,
                         hipStream_t s);
// The summed local of every leaf (what the level passes leave there) into
// ltot, in one launch; loc is not modified.
void ps_launch_leaf_locals(uint32_t n_nodes, const uint2* link, const uint32_t* parent,
                           const float4* loc, float4* ltot, hipStream_t s);

// ---- ps_traverse.hip: dual-tree traversal, M2L, P2P lists -------------------
struct PsTravTree {
    uint32_t n_nodes;  // list entries are checked against it
    const float4* geom;
    const float4* tpk;  // traversal node records (from the upward pass)
    const float4* mp;
    const uint2* link;
    const uint32_t* nend;
    float4* loc;
    uint32_t* p2p_off;   // per leaf: arena offset of its final source list
    uint32_t* p2p_len;   // per leaf: length of that list (0: no near field)
    uint32_t* p2p_moff;  // per leaf: offset of the list's far masks
};
// One traversal iteration's entries (target, list offset and length in the
// list arena, offset of its far masks in the mask arena).
struct PsTravEntries {
    uint32_t* tgt;
    uint32_t* off;
    uint32_t* len;
    uint32_t* moff;
};
// Device-resident bookkeeping of one traversal: every iteration is launched
// without host synchronisation; sizes live here.
// The traversal reads whole 64-word chunks of a list, past its end too:
// the list arena keeps this many words past the capacity it reports.
constexpr uint32_t PS_LIST_PAD = 64;
// The traversal addresses node records with 32-bit byte offsets: a tick's
// nodes (or the bound its launches are given) stay at most this many.
constexpr uint32_t PS_TRAV_MAX_NODES = 1u << 28;

struct PsTravState {
    uint32_t n_ent[16];            // entries of iteration k (n_ent[0] from the init)
    uint32_t arena_cap;            // capacities of the arenas (words; see PS_LIST_PAD)
    uint32_t mask_cap;
    uint32_t overflow;             // 1: a capacity was exceeded, the traversal is redone
    uint32_t pad;
    unsigned long long arena_ctr;  // list words reserved (or asked for, at an overflow)
    unsigned long long mask_ctr;   // mask words reserved (or asked for)
    unsigned long long m2l_pairs;  // filled by a reduction of the per-block far counts
    uint32_t seg_ctr[16][8];       // iteration k: entries handed out of each eighth
};
// Iteration-0 entries from the level starts ls[0..11] in device memory.  MT:
// every node of depth <= 2 is an entry, with the list {its system's root} at
// depth 2 or for a leaf above it; ST: the roots, each with itself.  Roots are
// nodes [0, n_sys); sys (optional) = system of each particle.  Resets the
// state.  max_ent bounds the entry count (grid size only).
void ps_launch_trav_init(const uint32_t* ls, int mt, uint32_t n_sys, const uint32_t* sys,
                         const uint32_t* nend, const uint2* link, PsTravEntries ent,
                         PsTravState* st, uint32_t* arena, uint32_t arena_cap, uint32_t mask_cap,
                         uint32_t max_ent, hipStream_t s);
// Blocks of the iteration launch for max_ent entries (length of its far_part
// slice; never more than max_blocks, a multiple of 8).
uint32_t ps_trav_blocks(uint32_t max_ent, int lanes, uint32_t max_blocks);
// Iteration k: MAC + M2L for every (entry target, listed source), far masks,
// P2P lists of final leaves, next lists appended to the arena and next
// entries (count in st->n_ent[k + 1]); far pairs per block into far_part.
// lanes: lanes per entry (8, 16, 32 or 64); max_ent bounds the launch (the
// count itself is st->n_ent[k]); spb: segments per block the entries are
// cut into (more: finer balance, more per-segment overhead).
void ps_launch_trav_step(PsTravTree t, float theta2, PsTravEntries ent, uint32_t* arena,
                         unsigned long long* masks, PsTravState* st, int k, uint32_t max_ent,
                         PsTravEntries next, unsigned long long* far_part, int kahan, int lanes,
                         uint32_t max_blocks, uint32_t spb, uint32_t rec_cap, hipStream_t s);

// ---- ps_p2p.hip: near field + L2P ------------------------------------------
// mode 0: f = m*field + m*sum (fused L2P+P2P); 1: f = m*field (L2P only);
// 2: f += m*sum (P2P only).  A leaf's near sources are the non-far entries of
// its final traversal list.  Leaves of >= tmin particles go to a
// wavefront-per-leaf kernel, the rest to a thread-per-particle kernel (tmin =
// 0: all of them); ls = level starts in device memory (ls[11] = node count),
// bound >= the node count (grid sizes only), big: scratch of 2 + bound
// words for the list of the big leaves (count, queue counter, leaves).  With s_big, fork and join set, the
// big-leaf kernel runs on s_big beside the other; all work is ordered after
// earlier work on s and before later work on s.
// pair_count (stats, may be null): per-block partial sums in block_pairs
// (>= 2 * ps_p2p_blocks(n, bound) entries), reduced into pair_count[0]
// (particle pairs) and pair_count[1] (target leaf x source leaf pairs).
uint32_t ps_p2p_blocks(size_t n, uint32_t bound);
// The list of the big leaves (>= tmin particles; big = count, queue counter,
// leaves) for the wavefront-per-leaf kernel: it reads only the tree, so a
// caller may build it early on another stream and tell ps_launch_p2p
// (collected), which otherwise builds it on s first.
void ps_launch_p2p_collect(const uint32_t* ls, uint32_t bound, uint32_t tmin, const uint2* link,
                           const uint32_t* nend, uint32_t* big, hipStream_t s);
void ps_launch_p2p(size_t n, const float4* pos4, const uint32_t* leaf_of, const float4* loc,
                   const uint32_t* p2p_off, const uint32_t* p2p_len, const uint32_t* p2p_moff,
                   const uint32_t* arena, const unsigned long long* masks, const uint2* link,
                   const uint32_t* nend, float* fx, float* fy, float* fz, int mode, int kahan,
                   const uint32_t* ls, uint32_t bound, uint32_t tmin, uint32_t* big,
                   unsigned long long* block_pairs, unsigned long long* pair_count,
                   hipStream_t s, hipStream_t s_big, hipEvent_t fork, hipEvent_t join, bool collected = false);

// ---- ps_direct.hip: FP64 direct sum (the error oracle on the device) -------
void ps_launch_direct_f64(const float4* pos4, size_t n, const uint32_t* targets, size_t nt,
                          double* out3, hipStream_t s);
void ps_launch_pack_pos4(const float* x, const float* y, const float* z, const float* m,
                         size_t n, float4* pos4, hipStream_t s);

// ---- ps_step.hip: the caller's step around the tick (README.md:128-161) ----
void ps_launch_gather_vel(const uint32_t* id, const float* vx, const float* vy,
                          const float* vz, float* ox, float* oy, float* oz, size_t n,
                          hipStream_t s);
void ps_launch_euler(float* x, float* y, float* z, float* vx, float* vy, float* vz,
                     const float* fx, const float* fy, const float* fz, const float* m,
                     float dt, size_t n, hipStream_t s);
void ps_launch_iota(uint32_t* id, size_t n, hipStream_t s);
// gather + Euler + reset ids in two kernels; tmp3: 3n floats of scratch
void ps_launch_frame(uint32_t* id, float* vx, float* vy, float* vz, float* tmp3, float* x,
                     float* y, float* z, const float* fx, const float* fy, const float* fz,
                     const float* m, float dt, size_t n, hipStream_t s);

// ---- ps_ops.hip: single-operator hooks for the known-answer tests ----------
hipError_t ps_op_m2l(const float* tgeom, const float* sgeom, const float* smp, float theta,
                     float* out4, hipStream_t s);
hipError_t ps_op_morton_exhaustive(unsigned long long* out2, hipStream_t s);
// plain copy of bytes (a multiple of 16) between device buffers
hipError_t ps_op_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// streaming kernels of known traffic (the profiler's byte-counter checks):
// kind 0 copy n 16-byte words src -> dst; 1 read n words of src (dst: one
// 16-byte word per block, at most 8192); 2 write n words of dst; 3 gather
// dst[i] = src[idx[i]] over n 4-byte elements
hipError_t ps_op_bandwidth(int kind, const void* src, void* dst, const uint32_t* idx, size_t n,
                           hipStream_t s);

// ---- ps_mlp.hip: learned far-field correction ------------------------------
#define PS_MLP_IN 24
#define PS_MLP_HIDDEN 64
#define PS_MLP_PARAMS 10115  // W1[64][24] b1 W2[64][64] b2 W3[64][64] b3 W4[3][64] b4[3]
// rows of the last tick's features feat[rows][24] and leaf field magnitudes
// scale[rows]: row r is particle idx[r] (sorted order; idx null: particle r);
// an index >= n gives a zero row
void ps_launch_mlp_features(size_t n, const uint32_t* idx, size_t rows, const float4* pos4,
                            const uint32_t* leaf_of, const float4* geom, const float4* mp,
                            const float4* loc, const uint32_t* parent, float* feat, float* scale,
                            hipStream_t s);
// out[n][3] = net(feat)
void ps_launch_mlp_forward(const float* W, const float* feat, size_t n, float* out,
                           hipStream_t s);
// the parameters W as the MFMA kernel's operands (PS_MLP_PACKED_BYTES)
#define PS_MLP_PACKED_BYTES 58368
void ps_launch_mlp_pack(const float* W, void* packed, hipStream_t s);
// f += m * scale * net(features), fused, on the matrix cores (packed weights)
void ps_launch_mlp_apply(size_t n, const void* packed, const float4* pos4,
                         const uint32_t* leaf_of, const float4* geom, const float4* mp,
                         const float4* loc, const uint32_t* parent, float* fx, float* fy,
                         float* fz, hipStream_t s);
// gradient of (1/n) sum |net(feat) - target|^2 into grad[PS_MLP_PARAMS] and the
// sum of squared errors into *loss (device); tmp: ps_mlp_grad_tmp_floats(n)
size_t ps_mlp_grad_tmp_floats(size_t n);
void ps_launch_mlp_grad(const float* W, const float* feat, const float* target, size_t n,
                        float* grad, float* loss, float* tmp, hipStream_t s);
// one Adam step (step counts from 1)
void ps_launch_mlp_adam(float* W, const float* grad, float* m1, float* m2, float lr, float b1,
                        float b2, float eps, int step, hipStream_t s);
