/*
 * polesitter_hip.h -- the C-ABI of libpolesitter_hip.so, the gfx950
 * (MI355X) implementation of polesitter's FMM tick.
 *
 * Plain pointers and sizes only.  Particle arrays may be device pointers
 * (used in place, zero copy) or host pointers (staged through device memory
 * and copied back); each array is classified on its own.  Every call is
 * synchronous on return, like the reference's API.
 *
 * The public C99 API (include/polesitter.h) is thin host code over these
 * entry points.  Each declaration names the reference interface it replaces
 * (src/polesitter.h = H).  Entry points marked "no reference counterpart" are
 * extensions: device memory helpers for callers without a HIP runtime of
 * their own, the on-device FP64 error oracle, statistics, white-box stages
 * for tests, and the caller's frame step on the device.
 *
 * Return values equal ps_result_t: 0 OK, -1 out of memory, -2 invalid
 * argument, -5 device error.
 */
#ifndef POLESITTER_HIP_H
#define POLESITTER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PS_HIP_OK      0
#define PS_HIP_EOOM    (-1)
#define PS_HIP_EINVAL  (-2)
#define PS_HIP_EDEVICE (-5)

/* traversal roots: PS_MULTITHREADING's depth-2 jobs (H:1772-1801, H:1883)
 * or the single (root, root) pair (H:1888,1905) */
#define PS_HIP_TRAV_MT 0
#define PS_HIP_TRAV_ST 1

/* Morton quantisation: scalar truncation (H:2132-2157), or the AVX2 path's
 * round-to-nearest-even for the first 8*floor(n/8) particles (H:2040-2072) */
#define PS_HIP_MORTON_TRUNC 0
#define PS_HIP_MORTON_AVX2  1

/* engine options (ps_hip_engine_set_option) */
#define PS_HIP_OPT_TRAVERSAL 1 /* PS_HIP_TRAV_MT / PS_HIP_TRAV_ST */
#define PS_HIP_OPT_KAHAN     2 /* 1: compensated far-field and P2P sums */
#define PS_HIP_OPT_THETA     3 /* MAC theta, like ps_config_t.theta */
#define PS_HIP_OPT_STATS     4 /* 1: time each stage and count P2P pairs; 2: time only;
                                  3: time the traversal only (two events per tick) */
#define PS_HIP_OPT_TRAV_LANES 5 /* lanes per traversal entry: 8, 16, 32, 64, or 0 (default):
                                  per iteration, the widest group that gives every entry
                                  a group at once, 8 when none does */
#define PS_HIP_OPT_P2P_BIG   6 /* P2P wavefront-per-leaf threshold (particles; 0 = off; 24) */
#define PS_HIP_OPT_LEAF_CAP  7 /* most particles of a leaf above depth 10: 1 (the reference) .. 64 */
#define PS_HIP_OPT_ARENA_WORDS 8 /* traversal list arena at first use, words per particle
                                    (96; grown and the tick re-run when too small) */
#define PS_HIP_OPT_TRAV_BLOCKS 9 /* workgroups of one traversal iteration: a multiple of 8
                                    from 8 to 8192 (the device's CU count) */
#define PS_HIP_OPT_TRAV_SEGMENTS 10 /* traversal work pieces per workgroup and iteration,
                                       1 to 64 (2): more balance the load, each costs a few
                                       block-wide synchronisations */
#define PS_HIP_OPT_SORT_INCREMENTAL 11 /* 1 (default): when at most 1/8 of the keys differ
                                          from the last tick's sorted keys in place, sort
                                          those and merge (the same permutation); 0: always
                                          the full radix sort */
#define PS_HIP_OPT_TRAV_RECORDS 12 /* 1 (default): the traversal keeps each entry's near
                                      sources in LDS between counting and writing;
                                      0: every next list is written from its list, far
                                      mask and links (the same words) */

typedef struct ps_hip_engine ps_hip_engine_t;

typedef struct {
    uint64_t n_particles;
    uint64_t n_nodes;
    uint64_t n_leaves;
    uint32_t max_depth;
    uint32_t iterations;      /* traversal iterations */
    uint32_t level_start[12]; /* depth d nodes: [level_start[d], level_start[d+1]) */
    uint64_t m2l_pairs;
    uint64_t p2p_leaf_pairs;
    uint64_t p2p_pairs;       /* sum over leaf pairs of n_t*n_s (stats on) */
    uint64_t list_entries;    /* traversal list entries written */
    /* milliseconds per stage of the last tick, HIP events on the engine's
     * stream (stats on); ms_p2p brackets the single P2P kernel launch,
     * ms_trav_steps brackets the traversal's iteration kernels */
    float ms_upload, ms_sort, ms_build, ms_upward, ms_traverse, ms_l2l, ms_p2p,
        ms_download, ms_total;
    float ms_trav_steps;
    float ms_mlp;             /* the learned correction, when one is set */
    uint64_t sort_changed;    /* keys that differed from the last tick's (n: not compared) */
    uint32_t sort_incremental; /* 1: the last tick sorted incrementally */
    uint32_t reserved_;
} ps_hip_stats_t;

typedef struct ps_hip_batch ps_hip_batch_t;

/* statistics of the last batched tick */
typedef struct {
    uint64_t n_particles;
    uint64_t n_systems;
    uint64_t n_nodes;       /* nodes of the whole forest */
    uint64_t m2l_pairs;
    uint64_t graph_replays; /* ticks run by replaying the recorded graph so far */
    uint32_t used_graph;    /* 1: the last tick was a graph replay */
    float ms_tick;          /* device time of the last tick (HIP events) */
} ps_hip_batch_stats_t;

/* batch options (ps_hip_batch_set_option; the engine options apply too,
 * except PS_HIP_OPT_STATS) */
#define PS_HIP_BATCH_OPT_GRAPH 100 /* 1 (default): record the tick once, replay it */

/* ps_init, H:1699-1767: device state for up to max_particles particles and
 * max_nodes octree nodes on the current device. */
int ps_hip_engine_create(ps_hip_engine_t** out, size_t max_particles,
                         size_t max_nodes, float theta, int traversal);

/* ps_init with caller-owned memory (no reference counterpart beyond the
 * reference's own rule that the caller owns every buffer, H:23-24,
 * H:1699-1767).  host_mem (optional, at least ps_hip_engine_host_bytes()):
 * the engine's state and its host mirrors are placed there instead of the
 * heap.  dev_mem (optional, device memory of at least
 * ps_hip_engine_scratch_bytes(max_particles) bytes): every device buffer a
 * tick uses is carved from it, and the engine makes no device allocation,
 * ever -- a tick that would need more (a traversal list arena overflow, a
 * learned correction's training buffers) returns PS_HIP_EOOM.  The caller
 * frees both after ps_hip_engine_destroy. */
int ps_hip_engine_create_in(ps_hip_engine_t** out, size_t max_particles,
                            size_t max_nodes, float theta, int traversal,
                            void* host_mem, size_t host_bytes, void* dev_mem,
                            size_t dev_bytes);
/* sizes for ps_hip_engine_create_in: host bytes for the engine's state;
 * device bytes for every buffer of a tick of up to max_particles particles
 * (node arrays for the most nodes that many particles can give, the list
 * arena at 144 words per particle -- the capacity a self-allocating engine
 * has after its first tick) */
size_t ps_hip_engine_host_bytes(void);
size_t ps_hip_engine_scratch_bytes(size_t max_particles);
/* the most octree nodes n particles can give (one new node per particle and
 * depth, at most 8^d at depth d, depth <= 10) */
size_t ps_hip_worst_case_nodes(size_t n);
/* no reference counterpart: device and pinned host allocations the library
 * has made in this process (count and bytes; ps_hip_malloc not included) */
int ps_hip_alloc_stats(uint64_t* calls, uint64_t* bytes);

/* Test hook, no counterpart in the reference: reach the library's error
 * paths without a real fault.  PS_HIP_FAIL_ALLOC: of the device and pinned
 * host allocations from now on, the one after `after` others reports out of
 * memory (PS_HIP_EOOM from the call that made it).  PS_HIP_FAIL_WAIT: of the
 * library's host waits on its own device work, the one after `after` others
 * reports a device failure once it has waited (PS_HIP_EDEVICE).  Each arming
 * fails one operation; after < 0 disarms one kind, PS_HIP_FAIL_OFF both.
 * Process-wide; PS_HIP_EINVAL for an unknown kind. */
#define PS_HIP_FAIL_OFF   0
#define PS_HIP_FAIL_ALLOC 1
#define PS_HIP_FAIL_WAIT  2
int ps_hip_debug_fail(int what, int after);

/* ps_destroy, H:2172-2192 */
int ps_hip_engine_destroy(ps_hip_engine_t* e);

/* ps_config_t.theta (H:304-310) and compile-time PS_MULTITHREADING (H:207)
 * made run-time; PS_HIP_OPT_KAHAN / _STATS have no reference counterpart */
int ps_hip_engine_set_option(ps_hip_engine_t* e, int key, double value);

/* no reference counterpart: counters and stage times of the last tick */
int ps_hip_engine_get_stats(const ps_hip_engine_t* e, ps_hip_stats_t* out);

/* ps_prepare_particles, H:1911-2170 (stateless, like the reference) */
int ps_hip_prepare(const float* x, const float* y, const float* z, size_t n,
                   uint32_t* codes, float* out_min, float* out_max,
                   float* out_range, int morton_mode);

/* No reference counterpart: ps_hip_prepare for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  The same results as ps_hip_prepare, bit for bit.
 * It allocates nothing: its scratch (the bounding-box partials and the four
 * cube parameters, PS_HIP_PREPARE_NOALLOC_SCRATCH bytes) is a static device
 * array that each device holds until the process ends.  x, y, z and codes
 * must be device arrays (host arrays: PS_HIP_EINVAL); n at most
 * PS_HIP_PREPARE_NOALLOC_MAX (larger: PS_HIP_EINVAL).  Runs on the default
 * stream and synchronises it.  Every call on one device shares that device's
 * scratch: calls that overlap on one device race. */
#define PS_HIP_PREPARE_NOALLOC_SCRATCH 8448
#define PS_HIP_PREPARE_NOALLOC_MAX 4294967295u
int ps_hip_prepare_noalloc(const float* x, const float* y, const float* z,
                           size_t n, uint32_t* codes, float* out_min,
                           float* out_max, float* out_range, int morton_mode);

/* ps_calc_forces, H:1840-1909: sorts x, y, z, mass, id and codes in place
 * by code and writes fx, fy, fz in the sorted order. */
int ps_hip_calc_forces(ps_hip_engine_t* e, float* x, float* y, float* z,
                       float* mass, float* fx, float* fy, float* fz,
                       uint32_t* id, size_t n, uint32_t* codes, float cx,
                       float cy, float cz, float hw);

/* ps_calc_forces (H:1840-1909) enqueued on the caller's HIP stream (NULL:
 * the default stream); it returns without waiting for the device.  Device
 * arrays only (host arrays: PS_HIP_EINVAL).  Every size of the tick is read
 * from device memory, as in the batched tick: the node arrays are grown to
 * ps_hip_worst_case_nodes(n) first and all traversal iterations are
 * launched.  The results are bit for bit those of ps_hip_calc_forces at the
 * same lanes per traversal entry (PS_HIP_OPT_TRAV_LANES; with the default 0
 * the asynchronous tick chooses them from the node capacity, the synchronous
 * one from the level sizes, which can change the order of an entry's M2L
 * sums).  ps_hip_calc_forces_wait must follow before the arrays are read,
 * another tick is started on the engine, or the stream is destroyed: it waits
 * for the tick, reports what ps_hip_calc_forces would have returned (invalid
 * codes, a tree over the node capacity) and, after a traversal list arena
 * overflow, grows the arenas and runs the tick again from the traversal
 * before it returns.  Stage events and pair counts (PS_HIP_OPT_STATS) are
 * not recorded by the asynchronous tick. */
int ps_hip_calc_forces_async(ps_hip_engine_t* e, float* x, float* y, float* z,
                             float* mass, float* fx, float* fy, float* fz,
                             uint32_t* id, size_t n, uint32_t* codes, float cx,
                             float cy, float cz, float hw, void* stream);
int ps_hip_calc_forces_wait(ps_hip_engine_t* e);

/* no reference counterpart: copies the last tick's octree (canonical level
 * order) to host arrays of capacity cap: geom4 = (x,y,z,hw), mp4, loc4 (total
 * field after L2L), link2 = (first child | first particle, mask | leaf flag
 * 0x80000000), nend, parent. */
int ps_hip_engine_export_tree(ps_hip_engine_t* e, size_t cap, float* geom4,
                              float* mp4, float* loc4, uint32_t* link2,
                              uint32_t* nend, uint32_t* parent,
                              uint32_t* n_nodes);

/* White-box stages over device arrays (no reference counterpart; the
 * reference's own tests call its static passes the same way,
 * tests/main.c:125-476).  bind, then any prefix of:
 * 1 sort, 2 build, 3 upward, 4 interaction (M2L), 5 downward (L2L + L2P,
 * f = m*field), 6 p2p (f += near field); or 7 after 3: 4 to 6 as the tick
 * runs them (the leaves' summed locals in one launch, the node locals
 * summed when they are read). */
int ps_hip_debug_bind(ps_hip_engine_t* e, float* x, float* y, float* z,
                      float* mass, float* fx, float* fy, float* fz,
                      uint32_t* id, size_t n, uint32_t* codes, float cx,
                      float cy, float cz, float hw);
int ps_hip_debug_stage(ps_hip_engine_t* e, int stage);
int ps_hip_debug_set_local(ps_hip_engine_t* e, uint32_t node,
                           const float* field3);

/* no reference counterpart: MAC + M2L of one cell pair (H:1056-1095) on the
 * device; out4 = (far, Ex, Ey, Ez) */
int ps_hip_op_m2l(const float* tgeom4, const float* sgeom4, const float* smp4,
                  float theta, float* out4);

/* no reference counterpart: the Morton interleave of all 2^30 (x,y,z)
 * against a bit loop; out2 = (mismatches, sum of codes) */
int ps_hip_op_morton_exhaustive(uint64_t* out2);

/* no reference counterpart: a plain streaming copy of `bytes` (a multiple of
 * 16) between two device buffers, synchronous -- the known-size kernel that
 * calibrates the profiler's HBM byte counters */
int ps_hip_op_copy(const void* src, void* dst, size_t bytes);

/* no reference counterpart: streaming kernels of known traffic for checking
 * the profiler's HBM byte counters, synchronous.  kind 0: copy n 16-byte
 * words src -> dst; 1: read n 16-byte words of src (dst receives one 16-byte
 * word per workgroup, at most 8192 of them); 2: write n 16-byte words of dst
 * (src unused); 3: gather dst[i] = src[idx[i]] over n 4-byte elements
 * (idx: device array, every index inside src). */
int ps_hip_op_bandwidth(int kind, const void* src, void* dst, const uint32_t* idx,
                        size_t n);

/* no reference counterpart: FP64 direct sum (tests/accuracy.c:42-72) on the
 * device for nt target positions (host array) against all n sources (device
 * arrays); out3 (host) = 3 doubles per target. */
int ps_hip_direct_f64(const float* x, const float* y, const float* z,
                      const float* mass, size_t n, const uint32_t* targets,
                      size_t nt, double* out3);

/* no reference counterpart: the caller's frame step of README.md:94-161 on
 * device arrays -- reset ids, gather velocities by id, semi-implicit Euler.
 * Runs on the engine's stream and synchronises. */
int ps_hip_step_reset_ids(ps_hip_engine_t* e, uint32_t* id, size_t n);
int ps_hip_step_gather(ps_hip_engine_t* e, const uint32_t* id, float* vx,
                       float* vy, float* vz, float* tmp3, size_t n);
int ps_hip_step_euler(ps_hip_engine_t* e, float* x, float* y, float* z,
                      float* vx, float* vy, float* vz, const float* fx,
                      const float* fy, const float* fz, const float* mass,
                      float dt, size_t n);
/* the three fused, after a tick: gather the velocities by id, semi-implicit
 * Euler (massless particles keep x and the gathered v), then reset the ids
 * to 0..n-1 for the next frame (README.md:94-161 from the gather to the top
 * of the next frame).  tmp3: 3n device floats of scratch.  Equal to
 * ps_hip_step_gather + ps_hip_step_euler + ps_hip_step_reset_ids. */
int ps_hip_step_frame(ps_hip_engine_t* e, uint32_t* id, float* vx, float* vy,
                      float* vz, float* tmp3, float* x, float* y, float* z,
                      const float* fx, const float* fy, const float* fz,
                      const float* mass, float dt, size_t n);

/* no reference counterpart (batched entry point): one tick of many
 * independent systems -- for each, ps_prepare_particles (H:1911-2170) and
 * ps_calc_forces (H:1840-1909) with the root at min + range/2
 * (README.md:104-126).  The systems are consecutive segments of one set of
 * device arrays: system k is [offsets[k], offsets[k+1]) (offsets: host array
 * of n_systems + 1, offsets[0] = 0, every system non-empty).  Each segment is
 * sorted in place and gets its forces in sorted order, as separate calls
 * would produce them.  The whole tick is one stream-ordered sequence with
 * every size in device memory: the first tick of a layout runs directly and
 * is recorded as a HIP graph, later ticks with the same arrays, offsets and
 * mode replay it.  morton_mode: PS_HIP_MORTON_TRUNC / _AVX2. */
int ps_hip_batch_create(ps_hip_batch_t** out, size_t max_particles,
                        size_t max_systems, float theta, int traversal);
int ps_hip_batch_destroy(ps_hip_batch_t* b);
int ps_hip_batch_set_option(ps_hip_batch_t* b, int key, double value);
int ps_hip_batch_tick(ps_hip_batch_t* b, float* x, float* y, float* z,
                      float* mass, float* fx, float* fy, float* fz,
                      uint32_t* id, uint32_t* codes, const uint32_t* offsets,
                      size_t n_systems, int morton_mode);
int ps_hip_batch_get_stats(const ps_hip_batch_t* b, ps_hip_batch_stats_t* out);
/* no reference counterpart: a leaf capacity per system (host array of
 * n_systems values in 1..64, the reference's rule is 1) in place of
 * PS_HIP_OPT_LEAF_CAP for the batch's later ticks, which must have at most
 * n_systems systems; NULL restores the option.  System k's tree is the one a
 * single tick with PS_HIP_OPT_LEAF_CAP = caps[k] builds. */
int ps_hip_batch_set_leaf_caps(ps_hip_batch_t* b, const uint8_t* caps,
                               size_t n_systems);

/* no reference counterpart: the learned far-field correction.  A small MLP
 * (24 -> 64 -> 64 -> 64 -> 3, GELU, FP32) maps per-particle leaf-local
 * features (offsets from the leaf, parent and grandparent centres, each
 * level's own M2L field, the leaf centroid) to a correction of the
 * acceleration in units of the leaf's far field: f += m * |E_L| * net(F).
 * params: PS_HIP_MLP_PARAMS floats, packed W1[64][24] b1[64] W2[64][64]
 * b2[64] W3[64][64] b3[64] W4[3][64] b4[3].  All arrays are device arrays
 * unless noted. */
#define PS_HIP_MLP_IN     24
#define PS_HIP_MLP_PARAMS 10115
/* set (params, host or device) or clear (NULL) the correction the engine
 * applies at the end of every tick */
int ps_hip_engine_set_mlp(ps_hip_engine_t* e, const float* params);
/* features feat[n][24] and scales scale[n] of the last tick (sorted order) */
int ps_hip_mlp_features(ps_hip_engine_t* e, float* feat, float* scale);
/* the same for the sorted positions idx[rows] only: row r is particle idx[r]
 * (an index >= n gives a zero row) */
int ps_hip_mlp_features_at(ps_hip_engine_t* e, const uint32_t* idx, size_t rows,
                           float* feat, float* scale);
/* out[n][3] = net(feat) */
int ps_hip_mlp_forward(ps_hip_engine_t* e, const float* params, const float* feat,
                       size_t n, float* out3);
/* grad = d/dparams of (1/n) sum |net(feat) - target|^2; *loss_host = that
 * mean (host pointer, may be NULL) */
int ps_hip_mlp_grad(ps_hip_engine_t* e, const float* params, const float* feat,
                    const float* target, size_t n, float* grad, float* loss_host);
/* one bias-corrected Adam step of params with moments m1, m2 (step >= 1) */
int ps_hip_mlp_adam(ps_hip_engine_t* e, float* params, const float* grad,
                    float* m1, float* m2, float lr, float beta1, float beta2,
                    float eps, int step);

/* no reference counterpart: device memory for callers without a HIP runtime
 * of their own (kind: 1 host->device, 2 device->host, 3 device->device) */
void* ps_hip_malloc(size_t bytes);
int   ps_hip_free(void* p);
int   ps_hip_memcpy(void* dst, const void* src, size_t bytes, int kind);
int   ps_hip_synchronize(void);
/* HIP streams for callers without a HIP runtime of their own (the
 * asynchronous tick's stream argument); non-blocking, default priority */
int   ps_hip_stream_create(void** out);
int   ps_hip_stream_destroy(void* stream);
int   ps_hip_stream_synchronize(void* stream);
int   ps_hip_device_name(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* POLESITTER_HIP_H */
