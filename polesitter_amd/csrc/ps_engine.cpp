// ps_engine.cpp -- host orchestration of the gfx950 FMM tick and the C-ABI
// of libpolesitter_hip.so (include/polesitter_hip.h).
//
// One engine per ps_context_t: a HIP stream, particle-sized device buffers
// for max_particles (sort scratch, pos4, leaf_of, host-pointer staging),
// node arrays for max_nodes, and traversal buffers grown on demand (kept
// across ticks).  A tick is a stream-ordered kernel sequence,
//   sort -> build -> upward -> traversal (M2L + P2P lists) -> L2L -> P2P+L2P
// with two host synchronisations: the node count (checked against the node
// capacity before the tree is written) and the traversal's overflow flag.
//
// A batch (ps_hip_batch_*) runs many independent systems as one forest with
// every size in device memory, so its whole tick is recorded once in a HIP
// graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <iterator>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/polesitter_hip.h"
#include "ps_kernels.h"

namespace {

int map_err(hipError_t e) {
    if (e == hipSuccess) return PS_HIP_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PS_HIP_EOOM;
    if (e == hipErrorInvalidValue) return PS_HIP_EINVAL;
    return PS_HIP_EDEVICE;
}

#define PS_CK(expr)                                   \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return map_err(_e);     \
    } while (0)

#define PS_RC(expr)                                   \
    do {                                              \
        int _r = (expr);                              \
        if (_r != PS_HIP_OK) return _r;               \
    } while (0)

bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool fixed = false;  // carved from caller-owned memory: never reallocated or freed
};

// device and pinned host allocations the library has made (process-wide)
std::atomic<unsigned long long> g_alloc_calls{0}, g_alloc_bytes{0};

// ps_hip_debug_fail's countdowns (-1: off): the allocation, or the host
// wait, that the count reaches reports a failure, once
std::atomic<int> g_fail_alloc{-1}, g_fail_wait{-1};

bool fail_now(std::atomic<int>& c) {
    int v = c.load(std::memory_order_relaxed);
    while (v >= 0)
        if (c.compare_exchange_weak(v, v - 1, std::memory_order_relaxed)) return v == 0;
    return false;
}

hipError_t dev_malloc(void** p, size_t bytes) {
    if (fail_now(g_fail_alloc)) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        g_alloc_calls.fetch_add(1, std::memory_order_relaxed);
        g_alloc_bytes.fetch_add(bytes, std::memory_order_relaxed);
    }
    return e;
}

hipError_t host_malloc(void** p, size_t bytes) {
    if (fail_now(g_fail_alloc)) return hipErrorOutOfMemory;
    hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
        g_alloc_calls.fetch_add(1, std::memory_order_relaxed);
        g_alloc_bytes.fetch_add(bytes, std::memory_order_relaxed);
    }
    return e;
}

// the library's host waits on its own work; a failure the device reports
// surfaces here
hipError_t wait_stream(hipStream_t s) {
    const hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess && fail_now(g_fail_wait) ? hipErrorLaunchFailure : e;
}

hipError_t wait_event(hipEvent_t ev) {
    const hipError_t e = hipEventSynchronize(ev);
    return e == hipSuccess && fail_now(g_fail_wait) ? hipErrorLaunchFailure : e;
}

// grow-only device buffer; synchronises the stream before replacing memory.
// A buffer carved from caller-owned memory cannot grow: a need beyond its
// size is out of memory.
hipError_t grow(DevBuf& b, size_t need, hipStream_t s) {
    if (need <= b.bytes && b.p) return hipSuccess;
    if (b.fixed) return hipErrorOutOfMemory;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t nb = align256(std::max<size_t>(need + need / 2, 256));
    hipError_t e = dev_malloc(&b.p, nb);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return e;
    }
    b.bytes = nb;
    return hipSuccess;
}

// capacity of the list arena in words, PS_LIST_PAD words short of its size
uint32_t list_cap_words(const DevBuf& b) {
    const size_t w = b.bytes / 4;
    return (uint32_t)std::min<size_t>(w > PS_LIST_PAD ? w - PS_LIST_PAD : 0, 0xFFFFFFFFull);
}

void release(DevBuf& b) {
    if (b.p && !b.fixed) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    b.fixed = false;
}

// most nodes n particles can produce: at most one new node per particle per
// depth, and at most 8^d nodes at depth d
size_t worst_case_nodes(size_t n) {
    size_t total = 0, width = 1;
    for (int d = 0; d <= 10; ++d) {
        total += std::min(width, std::max<size_t>(n, 1));
        width = width > (SIZE_MAX / 8) ? SIZE_MAX : width * 8;
    }
    return total;
}

enum Stage {
    kSort = 1, kBuild = 2, kUpward = 3, kInteract = 4, kDownward = 5, kP2P = 6,
    kTickTail = 7  // 4..6 as the tick runs them
};

constexpr int PS_MAX_DEPTH_ITERS = 11;  // traversal iterations of a depth-10 tree

}  // namespace

struct ps_hip_engine {
    int device = 0;
    hipStream_t s = nullptr;
    hipStream_t s2 = nullptr;  // second stream: the big-leaf P2P kernel beside the other one
    hipEvent_t fork = nullptr, join = nullptr;
    // the second stream's side work of a tick: the big-leaf list once the
    // tree is written (ev_tree), the traversal's pair count and state once it
    // is done (ev_trav); joined before the tick's last event (ev_side)
    hipEvent_t ev_tree = nullptr, ev_trav = nullptr, ev_side = nullptr;
    bool collected = false;  // this tick's big-leaf list is built (on s2)
    size_t max_p = 0;
    size_t node_cap = 0;
    float theta = 1.0f;
    int trav = PS_HIP_TRAV_MT;
    int kahan = 0;
    int stats_on = 0;          // stage events (all of them)
    uint32_t ev_mask = 0;      // stage events recorded (bit k: ev[k])
    bool trav_ev = false;      // events around the traversal iterations
    int count_on = 0;   // pair counts (extra reductions)
    int trav_lanes = 0;  // lanes per traversal entry (0: chosen per iteration)
    uint32_t trav_blocks = 256;  // most workgroups of a traversal iteration (one per CU)
    uint32_t trav_spb = 2;       // traversal segments per workgroup
    uint32_t p2p_big = 24;  // leaves of >= this many particles: wavefront per leaf
    int leaf_cap = 1;       // most particles of a leaf above depth 10 (the reference: 1)
    uint32_t arena_words = 96;  // list arena at first use, words per particle
    DevBuf mlp_b;           // learned correction parameters (empty: none)
    DevBuf mlp_pk_b;        // the same as the MFMA kernel's packed operands
    bool mlp_on = false;
    // loc holds each node's own local (the tick gave the leaves their sums
    // in ltot): the level-by-level L2L is still to be applied before loc is read
    bool l2l_pending = false;
    DevBuf mlp_tmp;         // training scratch

    // host-pointer staging, one buffer per array (allocated on first use)
    DevBuf stage[9];
    // sort scratch: two sets of (key, id, x, y, z, m)
    DevBuf sort_t;
    DevBuf sort_tmp, tree_tmp;
    DevBuf pos4_b, leaf_of_b;
    // the keys the last sort produced, in place (two, alternating), for the
    // incremental sort of the next tick
    DevBuf prev_b[2];
    int prev_cur = 0;        // prev_b[prev_cur] holds them
    size_t prev_n = 0;       // their count (0: none)
    int incremental = 1;     // PS_HIP_OPT_SORT_INCREMENTAL
    int trav_records = 1;    // PS_HIP_OPT_TRAV_RECORDS
    // the incremental sort's copy back into the caller's arrays runs on s2;
    // skey is where the sorted keys are meanwhile (the tree reads them)
    bool copy_pending = false;
    hipEvent_t ev_sorted = nullptr, ev_copied = nullptr;
    uint32_t* skey = nullptr;
    // nodes
    DevBuf geom_b, tpk_b, mp_b, loc_b, link_b, nend_b, parent_b, p2p_off_b, p2p_len_b, p2p_moff_b;
    DevBuf ltot_b;  // summed locals of the leaves (the tick's L2L)
    DevBuf oct_b;   // each node's slot in its parent (tree write scratch)
    // traversal: entries (tgt, off, len, moff) ping-pong; the list arena and
    // the far-mask arena fill over the whole tick
    DevBuf ent[2][4];
    DevBuf arena_b, masks_b;
    DevBuf far_b, block_pairs_b, tstate_b, big_b;
    // small device / pinned host scalars
    DevBuf small_b;  // level_start[12], flags, pad | u64 totals[4] | u64 pairs | u64 leaf pairs
    uint32_t* h_small = nullptr;  // pinned mirror (in caller memory when placed)
    PsTravState* h_tstate = nullptr;  // pinned copy of the traversal's state
    bool placed = false;    // the engine and its host mirrors live in caller memory
    bool borrowed = false;  // every device buffer is carved from caller memory
    // an asynchronous tick launched and not yet waited for
    bool pending = false;
    size_t node_alloc = 0;  // node capacity of the asynchronous tick (worst case of n)
    hipStream_t async_s = nullptr;
    hipEvent_t ev_async = nullptr;
    hipEvent_t ev_counts = nullptr;   // the tree's counts have reached h_small

    // arrays of the current tick
    float *x = nullptr, *y = nullptr, *z = nullptr, *m = nullptr;
    float *fx = nullptr, *fy = nullptr, *fz = nullptr;
    uint32_t *id = nullptr, *code = nullptr;
    size_t n = 0;
    float4 root = make_float4(0.f, 0.f, 0.f, 0.f);
    bool bound = false, built = false;

    uint32_t ls[12] = {0};
    uint32_t n_nodes = 0, max_depth = 0;
    ps_hip_stats_t st;
    hipEvent_t ev[12] = {nullptr};
    hipEvent_t ev_it[4 * 16] = {nullptr};  // per traversal iteration: count and write

    uint32_t* small_u32() { return reinterpret_cast<uint32_t*>(small_b.p); }
    // changed keys of the incremental sort (device), and its host mirror
    unsigned long long* delta_count() {
        return reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(small_b.p) + 160);
    }
    unsigned long long host_delta_count() const {
        return *reinterpret_cast<const unsigned long long*>(
            reinterpret_cast<const char*>(h_small) + 160);
    }
    uint32_t* prev_keys(int k) { return reinterpret_cast<uint32_t*>(prev_b[k].p); }
    unsigned long long* totals() {
        return reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(small_b.p) + 64);
    }
    unsigned long long* pairs() { return totals() + 4; }
    unsigned long long* leaf_pairs() { return totals() + 5; }
    float4* pos4() { return reinterpret_cast<float4*>(pos4_b.p); }
    uint32_t* leaf_of() { return reinterpret_cast<uint32_t*>(leaf_of_b.p); }
    float4* geom() { return reinterpret_cast<float4*>(geom_b.p); }
    float4* tpk() { return reinterpret_cast<float4*>(tpk_b.p); }
    float4* mp() { return reinterpret_cast<float4*>(mp_b.p); }
    float4* loc() { return reinterpret_cast<float4*>(loc_b.p); }
    float4* ltot() { return reinterpret_cast<float4*>(ltot_b.p); }
    uint8_t* oct() { return reinterpret_cast<uint8_t*>(oct_b.p); }
    uint2* link() { return reinterpret_cast<uint2*>(link_b.p); }
    uint32_t* nend() { return reinterpret_cast<uint32_t*>(nend_b.p); }
    uint32_t* parent() { return reinterpret_cast<uint32_t*>(parent_b.p); }
    uint32_t* p2p_off() { return reinterpret_cast<uint32_t*>(p2p_off_b.p); }
    uint32_t* p2p_len() { return reinterpret_cast<uint32_t*>(p2p_len_b.p); }
    uint32_t* p2p_moff() { return reinterpret_cast<uint32_t*>(p2p_moff_b.p); }
    uint32_t* arena() { return reinterpret_cast<uint32_t*>(arena_b.p); }
    unsigned long long* masks() { return reinterpret_cast<unsigned long long*>(masks_b.p); }
    PsTravEntries entries(int k) {
        PsTravEntries r;
        r.tgt = reinterpret_cast<uint32_t*>(ent[k][0].p);
        r.off = reinterpret_cast<uint32_t*>(ent[k][1].p);
        r.len = reinterpret_cast<uint32_t*>(ent[k][2].p);
        r.moff = reinterpret_cast<uint32_t*>(ent[k][3].p);
        return r;
    }

    PsSortArrays sort_set(int k) {
        char* base = reinterpret_cast<char*>(sort_t.p) + (size_t)k * 6 * max_p * 4;
        PsSortArrays a;
        a.key = reinterpret_cast<uint32_t*>(base + 0 * max_p * 4);
        a.id = reinterpret_cast<uint32_t*>(base + 1 * max_p * 4);
        a.x = reinterpret_cast<float*>(base + 2 * max_p * 4);
        a.y = reinterpret_cast<float*>(base + 3 * max_p * 4);
        a.z = reinterpret_cast<float*>(base + 4 * max_p * 4);
        a.m = reinterpret_cast<float*>(base + 5 * max_p * 4);
        a.aux = nullptr;
        return a;
    }

    void record(int k) {
        if ((ev_mask >> k) & 1u) (void)hipEventRecord(ev[k], s);
    }
    float elapsed(int a, int b) {
        float ms = 0.f;
        if (((ev_mask >> a) & (ev_mask >> b) & 1u) &&
            hipEventElapsedTime(&ms, ev[a], ev[b]) == hipSuccess)
            return ms;
        return 0.f;
    }

    void free_all() {
        for (auto& b : stage) release(b);
        for (auto& b : prev_b) release(b);
        DevBuf* all[] = {&sort_t,  &sort_tmp,  &tree_tmp,  &pos4_b,    &leaf_of_b, &geom_b, &tpk_b,
                         &mp_b,    &loc_b,     &link_b,    &nend_b,    &parent_b,  &p2p_off_b,
                         &p2p_len_b, &p2p_moff_b, &arena_b, &masks_b,  &far_b,  &ltot_b, &oct_b,
                         &small_b,   &block_pairs_b, &tstate_b, &big_b,  &mlp_b, &mlp_pk_b, &mlp_tmp};
        for (DevBuf* b : all) release(*b);
        for (auto& es : ent)
            for (auto& b : es) release(b);
        if (!placed) {
            if (h_small) (void)hipHostFree(h_small);
            if (h_tstate) (void)hipHostFree(h_tstate);
        }
        h_small = nullptr;
        h_tstate = nullptr;
        if (ev_counts) (void)hipEventDestroy(ev_counts);
        ev_counts = nullptr;
        if (ev_async) (void)hipEventDestroy(ev_async);
        ev_async = nullptr;
        if (ev_sorted) (void)hipEventDestroy(ev_sorted);
        if (ev_copied) (void)hipEventDestroy(ev_copied);
        ev_sorted = ev_copied = nullptr;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev_it)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        fork = join = nullptr;
        for (hipEvent_t* v : {&ev_tree, &ev_trav, &ev_side}) {
            if (*v) (void)hipEventDestroy(*v);
            *v = nullptr;
        }
        if (s2) (void)hipStreamDestroy(s2);
        s2 = nullptr;
        if (s) (void)hipStreamDestroy(s);
        s = nullptr;
    }
};

namespace {

constexpr uint32_t kMaxTravBlocks = 8192;  // PS_HIP_OPT_TRAV_BLOCKS upper bound

// Every device buffer a tick can use, with its size for P particles and M
// nodes: the buffers ps_hip_engine_create sizes at once, and the ones the
// tick grows (traversal entries and arenas, P2P lists, host-array staging,
// the correction's parameters).  With caller-owned device memory all of them
// are carved from it at creation; the sizes are upper bounds of every need
// below, so nothing grows later (a larger need, e.g. a list arena overflow,
// is PS_HIP_EOOM).
template <class F>
void for_each_tick_buffer(ps_hip_engine* e, size_t P, size_t M, uint32_t arena_words, F&& f) {
    f(e->sort_t, 2 * 6 * P * 4);
    f(e->sort_tmp, ps_sort_tmp_bytes(P));
    f(e->tree_tmp, ps_tree_tmp_bytes(P));
    f(e->pos4_b, 16 * P);
    f(e->leaf_of_b, 4 * P);
    f(e->prev_b[0], 4 * P);
    f(e->prev_b[1], 4 * P);
    DevBuf* node16[] = {&e->geom_b, &e->tpk_b, &e->mp_b, &e->loc_b, &e->ltot_b};
    for (DevBuf* d : node16) f(*d, 16 * M);
    f(e->oct_b, M);
    f(e->link_b, 8 * M);
    DevBuf* node4[] = {&e->nend_b, &e->parent_b, &e->p2p_off_b, &e->p2p_len_b, &e->p2p_moff_b};
    for (DevBuf* d : node4) f(*d, 4 * M);
    f(e->small_b, 256);
    for (auto& es : e->ent)
        for (auto& b : es) f(b, 4 * M);
    f(e->far_b, 8 * (size_t)kMaxTravBlocks * PS_MAX_DEPTH_ITERS);
    f(e->tstate_b, sizeof(PsTravState));
    // the list arena with the slack a self-allocating engine's first
    // allocation has (grow adds half): 144 words per particle by default
    f(e->arena_b, 4 * ((size_t)arena_words * P * 3 / 2 + 4096));
    f(e->masks_b, 8 * (5 * M + 4112));
    f(e->big_b, 4 * (M + 2));
    f(e->block_pairs_b, 16 * (size_t)ps_p2p_blocks(P, (uint32_t)std::min<size_t>(M, 0xFFFFFFFFu)));
    for (auto& b : e->stage) f(b, 4 * P);
    f(e->mlp_b, 4 * PS_MLP_PARAMS);
    f(e->mlp_pk_b, PS_MLP_PACKED_BYTES);
}

size_t tick_scratch_bytes(size_t P, size_t M, uint32_t arena_words) {
    ps_hip_engine probe;
    size_t total = 0;
    for_each_tick_buffer(&probe, P, M, arena_words,
                         [&](DevBuf&, size_t b) { total += align256(std::max<size_t>(b, 1)); });
    return total + 256;  // room to align the base
}

// Carve every tick buffer from [base, base + bytes).  The node arrays hold
// the most nodes max_p particles can give, so the asynchronous tick (which
// cannot look at the tree's size before it launches) always fits; the
// caller's node capacity is still checked after each tick.
int carve_all(ps_hip_engine* e, void* base, size_t bytes) {
    const size_t M = worst_case_nodes(e->max_p);
    const size_t need = tick_scratch_bytes(e->max_p, M, e->arena_words);
    if (bytes < need) return PS_HIP_EOOM;
    char* p = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(base) + 255) & ~(uintptr_t)255);
    for_each_tick_buffer(e, e->max_p, M, e->arena_words, [&](DevBuf& b, size_t sz) {
        b.p = p;
        b.bytes = align256(std::max<size_t>(sz, 1));
        b.fixed = true;
        p += b.bytes;
    });
    // the one buffer outside a tick (the correction's training scratch) gets
    // no memory: with caller-owned memory it cannot grow either
    e->mlp_tmp.fixed = true;
    e->borrowed = true;
    return PS_HIP_OK;
}

hipError_t alloc_fixed(ps_hip_engine* e) {
    const size_t P = e->max_p, M = e->node_cap;
    hipError_t r;
#define PS_AL(buf, bytes)                                     \
    if ((r = grow(e->buf, (bytes), e->s)) != hipSuccess) return r;
    PS_AL(sort_t, 2 * 6 * P * 4);
    PS_AL(sort_tmp, ps_sort_tmp_bytes(P));
    PS_AL(tree_tmp, ps_tree_tmp_bytes(P));
    PS_AL(pos4_b, 16 * P);
    PS_AL(leaf_of_b, 4 * P);
    PS_AL(prev_b[0], 4 * P);
    PS_AL(prev_b[1], 4 * P);
    PS_AL(geom_b, 16 * M);
    PS_AL(tpk_b, 16 * M);
    PS_AL(mp_b, 16 * M);
    PS_AL(loc_b, 16 * M);
    PS_AL(ltot_b, 16 * M);
    PS_AL(oct_b, M);
    PS_AL(link_b, 8 * M);
    PS_AL(nend_b, 4 * M);
    PS_AL(parent_b, 4 * M);
    PS_AL(p2p_off_b, 4 * M);
    PS_AL(p2p_len_b, 4 * M);
    PS_AL(p2p_moff_b, 4 * M);
    PS_AL(small_b, 256);
#undef PS_AL
    return hipSuccess;
}

hipError_t alloc_mirrors(ps_hip_engine* e) {
    if (e->placed) return hipSuccess;  // in the caller's host memory
    hipError_t r = host_malloc(reinterpret_cast<void**>(&e->h_small), 256);
    if (r == hipSuccess)
        r = host_malloc(reinterpret_cast<void**>(&e->h_tstate), sizeof(PsTravState));
    return r;
}

// bind one caller array: device pointers are used in place, host arrays go
// through the engine's staging buffer k
template <class T>
int bind_array(ps_hip_engine* e, int k, T* p, size_t n, bool upload, T** out) {
    if (!p) return PS_HIP_EINVAL;
    if (is_device_ptr(p)) {
        *out = p;
        return PS_HIP_OK;
    }
    PS_CK(grow(e->stage[k], std::max<size_t>(e->max_p, 1) * sizeof(T), e->s));
    *out = reinterpret_cast<T*>(e->stage[k].p);
    if (upload) PS_CK(hipMemcpyAsync(*out, p, n * sizeof(T), hipMemcpyHostToDevice, e->s));
    return PS_HIP_OK;
}

constexpr size_t kIncrementalMin = 4096;  // smaller ticks sort in full (cheaper than a readback)

// The tick's sort.  After a tick of the same size the last sort's keys are
// kept in place: when few keys changed (the caller's frame loop moves the
// particles a little between ticks), only those are sorted and merged with
// the rest (ps_sort_delta_*: the same permutation as the full radix sort);
// otherwise, or for the first tick, the full radix sort.  One small
// readback (the changed count) decides.  Either way the sorted keys are
// kept for the next tick.
int stage_sort(ps_hip_engine* e) {
    PsSortArrays io = {e->code, e->id, e->x, e->y, e->z, e->m};
    const int cur = e->prev_cur, nxt = cur ^ 1;
    e->st.sort_changed = e->n;
    e->st.sort_incremental = 0;
    if (e->incremental && e->prev_n == e->n && e->n >= kIncrementalMin) {
        PsSortArrays t1 = e->sort_set(1);
        uint32_t* ckey = reinterpret_cast<uint32_t*>(t1.x);
        uint32_t* cpos = reinterpret_cast<uint32_t*>(t1.y);
        PS_CK(ps_sort_delta_count(e->code, e->prev_keys(cur), e->n, e->sort_tmp.p, ckey, cpos,
                                  e->delta_count(), e->s));
        PS_CK(hipMemcpyAsync(reinterpret_cast<char*>(e->h_small) + 160, e->delta_count(), 8,
                             hipMemcpyDeviceToHost, e->s));
        PS_CK(wait_stream(e->s));
        const unsigned long long nc = e->host_delta_count();
        e->st.sort_changed = nc;
        if (nc * 8 <= e->n) {
            PsSortArrays t0 = e->sort_set(0);
            PS_CK(ps_sort_delta_finish(io, t0, t1, e->pos4(), e->n, e->sort_tmp.p,
                                       e->prev_keys(cur), ckey, cpos, (uint32_t)nc,
                                       e->prev_keys(nxt), e->s, e->s2, e->ev_sorted,
                                       e->ev_copied));
            e->copy_pending = true;
            e->skey = t0.key;
            e->prev_cur = nxt;
            e->st.sort_incremental = 1;
            return PS_HIP_OK;
        }
    }
    PS_CK(ps_radix_sort(io, e->sort_set(0), e->sort_set(1), e->pos4(), e->n, e->sort_tmp.p,
                        e->s, e->prev_keys(nxt)));
    e->skey = e->code;
    e->prev_cur = nxt;
    e->prev_n = e->n;
    return PS_HIP_OK;
}

// Work after this on the engine stream sees the caller's arrays sorted.
int finish_copy(ps_hip_engine* e) {
    if (!e->copy_pending) return PS_HIP_OK;
    e->copy_pending = false;
    e->skey = e->code;
    PS_CK(hipStreamWaitEvent(e->s, e->ev_copied, 0));
    return PS_HIP_OK;
}

// The counts go to the host while the GPU already writes the tree: the write
// reads the node count from device memory and stores no node past the
// arrays' capacity, so the host only has to look at the counts before it
// sizes the next launches (invalid codes, a tree over the capacity: error,
// after the stream has drained in the caller).
int stage_build(ps_hip_engine* e) {
    uint32_t* small = e->small_u32();
    const uint32_t* keys = e->skey ? e->skey : e->code;  // the sorted keys
    PS_CK(ps_tree_count(keys, nullptr, e->n, small, small + 12, e->totals(), e->tree_tmp.p,
                        e->leaf_cap, e->s));
    PS_CK(hipMemcpyAsync(e->h_small, e->small_b.p, 96, hipMemcpyDeviceToHost, e->s));
    PS_CK(hipEventRecord(e->ev_counts, e->s));
    // the mask pass also clears the locals and the P2P list lengths
    PS_CK(ps_tree_write(keys, nullptr, e->n, e->root, nullptr, e->tree_tmp.p, e->geom(),
                        e->link(), e->nend(), e->parent(), e->leaf_of(), e->oct(),
                        e->node_cap, e->totals(), e->leaf_cap, e->s, nullptr, e->loc(),
                        e->p2p_len()));
    e->l2l_pending = false;
    e->collected = false;  // a new tree: its big-leaf list is still to be built
    PS_CK(wait_event(e->ev_counts));
    const uint32_t flags = e->h_small[12];
    const unsigned long long total =
        *reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(e->h_small) + 64);
    if (flags & 1u) return PS_HIP_EINVAL;  // codes must be 30-bit Morton codes
    if (total > e->node_cap) return PS_HIP_EOOM;
    std::memcpy(e->ls, e->h_small, sizeof(e->ls));
    e->n_nodes = (uint32_t)total;
    e->max_depth = 0;
    for (int d = 0; d <= 10; ++d)
        if (e->ls[d + 1] > e->ls[d]) e->max_depth = (uint32_t)d;
    e->built = true;
    return PS_HIP_OK;
}

// Levels of at most this many nodes, from the root down to the first larger
// one, go to one launch of one block (a launch per level costs more there
// than the level's work).
constexpr uint32_t kUpwardTopNodes = 2048;

int stage_upward(ps_hip_engine* e) {
    int d_top = -1;
    while (d_top < (int)e->max_depth && e->ls[d_top + 2] - e->ls[d_top + 1] <= kUpwardTopNodes)
        ++d_top;
    for (int d = (int)e->max_depth; d > d_top; --d)
        ps_launch_upward_level(e->ls[d], e->ls[d + 1], d, e->tpk(), e->pos4(), e->geom(),
                               e->link(), e->nend(), e->mp(), e->s);
    ps_launch_upward_top(e->ls, d_top, e->tpk(), e->pos4(), e->geom(), e->link(), e->nend(),
                         e->mp(), e->s);
    return map_err(hipGetLastError());
}

// Upper bound on the entries of iteration k: MT targets at iteration k lie
// at depths 0..k+2, ST targets at depths 0..k.
uint32_t entries_bound(const ps_hip_engine* e, bool mt, int k) {
    return e->ls[std::min(k + (mt ? 3 : 1), 11)];
}

// Launch the whole traversal: one kernel per iteration, sizes in device
// memory, no host synchronisation in between.  An internal source at
// iteration k sits at depth k (each iteration expands every near internal
// source), so a tree of depth D needs D + 1 iterations.
// Lanes per entry of a traversal iteration of at most ub entries: the fixed
// setting, or (0, the default) the widest group (64, 32, 16) with which every
// entry has its group at once on the iteration's grid of trav_blocks blocks of
// 1024 threads, else 8.  The small top iterations are chains of memory round
// trips: with wide groups they spread over the whole chip, and each lane has
// fewer of the entry's sources to test and expand.
// words of near records a traversal group may keep in LDS (0: none, every
// continuing entry is written from its list)
uint32_t rec_cap_of(const ps_hip_engine* e) { return e->trav_records ? 0xFFFFFFFFu : 0u; }

int trav_lanes_for(const ps_hip_engine* e, uint32_t ub) {
    if (e->trav_lanes) return e->trav_lanes;
    const unsigned long long lanes = (unsigned long long)e->trav_blocks * 1024ull;
    for (int g = 64; g > 8; g >>= 1)
        if ((unsigned long long)ub * (unsigned)g <= lanes) return g;
    return 8;
}

int launch_traverse(ps_hip_engine* e, float theta2, bool mt, int iters) {
    const uint32_t n_nodes = e->n_nodes;
    PsTravState* st = reinterpret_cast<PsTravState*>(e->tstate_b.p);
    const uint32_t arena_cap = list_cap_words(e->arena_b);
    const uint32_t mask_cap = (uint32_t)std::min<size_t>(e->masks_b.bytes / 8, 0xFFFFFFFFull);
    ps_launch_trav_init(e->small_u32(), mt ? 1 : 0, 1u, nullptr, e->nend(), e->link(),
                        e->entries(0), st, e->arena(), arena_cap, mask_cap,
                        mt ? e->ls[3] : 1u, e->s);
    PsTravTree t;
    t.n_nodes = n_nodes;
    t.geom = e->geom();
    t.tpk = e->tpk();
    t.mp = e->mp();
    t.link = e->link();
    t.nend = e->nend();
    t.loc = e->loc();
    t.p2p_off = e->p2p_off();
    t.p2p_len = e->p2p_len();
    t.p2p_moff = e->p2p_moff();
    unsigned long long* far_part = reinterpret_cast<unsigned long long*>(e->far_b.p);
    size_t far_n = 0;
    if (e->trav_ev) (void)hipEventRecord(e->ev_it[0], e->s);
    for (int k = 0; k < iters; ++k) {
        const int cur = k & 1, nxt = cur ^ 1;
        const uint32_t ub = std::min(entries_bound(e, mt, k), n_nodes);
        const int lanes = trav_lanes_for(e, ub);
        ps_launch_trav_step(t, theta2, e->entries(cur), e->arena(), e->masks(), st, k, ub,
                            e->entries(nxt), far_part + far_n, e->kahan, lanes,
                            e->trav_blocks, e->trav_spb, rec_cap_of(e), e->s);
        far_n += ps_trav_blocks(ub, lanes, e->trav_blocks);
    }
    if (e->trav_ev) (void)hipEventRecord(e->ev_it[1], e->s);
    // the pair count and the state's copy to the host run beside the rest
    // of the tick (joined before its end)
    PS_CK(hipEventRecord(e->ev_trav, e->s));
    PS_CK(hipStreamWaitEvent(e->s2, e->ev_trav, 0));
    ps_reduce_u64(far_part, far_n, &st->m2l_pairs, e->s2);
    return map_err(hipGetLastError());
}

float theta_squared(const ps_hip_engine* e) {
    volatile float t = e->theta;  // fl(theta * theta), H:1064
    return t * t;
}

// Launches the traversal; its state (overflow flag, counts) is copied to the
// host behind it and looked at by trav_finish once the stream has drained.
int stage_traverse(ps_hip_engine* e) {
    e->l2l_pending = false;
    const float theta2 = theta_squared(e);
    const bool mt = e->trav == PS_HIP_TRAV_MT;
    const int iters = (int)e->max_depth + 1;
    const size_t M = e->n_nodes;
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 4; ++k) PS_CK(grow(e->ent[b][k], 4 * M, e->s));
    PS_CK(grow(e->far_b, 8 * (size_t)e->trav_blocks * (size_t)iters, e->s));
    PS_CK(grow(e->tstate_b, sizeof(PsTravState), e->s));
    // first use: room for arena_words list words per particle (96); a larger
    // need is met by growing and running the traversal again
    const size_t n0 = std::max<size_t>(e->ls[3], 1) + 16;
    PS_CK(grow(e->arena_b, 4 * ((size_t)e->arena_words * e->n + 4096), e->s));
    PS_CK(grow(e->masks_b, 8 * std::max<size_t>(4 * M + n0 + 4096, n0), e->s));
    PS_RC(launch_traverse(e, theta2, mt, iters));
    PS_CK(hipMemcpyAsync(e->h_tstate, e->tstate_b.p, sizeof(PsTravState), hipMemcpyDeviceToHost,
                         e->s2));
    return PS_HIP_OK;
}

// After the stream has drained: the traversal's outcome.  *redo: an arena was
// too small -- it has been grown and the partial results cleared, and the
// tick has to run again from the traversal (the sort and the tree stand).
int trav_finish(ps_hip_engine* e, int attempt, bool* redo) {
    const PsTravState hst = *e->h_tstate;
    *redo = false;
    if (hst.overflow) {
        if (attempt >= 4 || hst.arena_ctr > 0xFFFFFFFFull || hst.mask_ctr > 0xFFFFFFFFull)
            return PS_HIP_EOOM;
        // grow past what was asked for, clear the partial results
        PS_CK(grow(e->arena_b, std::max<size_t>(2 * e->arena_b.bytes, 5 * hst.arena_ctr), e->s));
        PS_CK(grow(e->masks_b, std::max<size_t>(2 * e->masks_b.bytes, 10 * hst.mask_ctr), e->s));
        ps_fill_u32(reinterpret_cast<uint32_t*>(e->loc()), 0u, 4 * (size_t)e->n_nodes, e->s);
        ps_fill_u32(e->p2p_len(), 0u, e->n_nodes, e->s);
        *redo = true;
        return map_err(hipGetLastError());
    }
    e->st.m2l_pairs = hst.m2l_pairs;
    e->st.list_entries = hst.arena_ctr - 1;
    e->st.iterations = e->max_depth + 1;
    if (e->trav_ev) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e->ev_it[0], e->ev_it[1]) == hipSuccess)
            e->st.ms_trav_steps = ms;
    }
    return PS_HIP_OK;
}

int stage_l2l(ps_hip_engine* e) {
    for (uint32_t d = 1; d <= e->max_depth; ++d)
        ps_launch_l2l_level(e->ls[d], e->ls[d + 1], e->parent(), e->loc(), e->s);
    e->l2l_pending = false;
    return map_err(hipGetLastError());
}

// The tick's L2L: only the leaves' sums feed the forces, so one launch puts
// them in ltot and the level-by-level pass over loc waits until something
// reads the node locals (materialize_locals).
int stage_leaf_locals(ps_hip_engine* e) {
    ps_launch_leaf_locals(e->n_nodes, e->link(), e->parent(), e->loc(), e->ltot(), e->s);
    e->l2l_pending = true;
    return map_err(hipGetLastError());
}

// Applies the pending L2L to loc (after a tick, before the locals are read).
int materialize_locals(ps_hip_engine* e) {
    if (!e->l2l_pending) return PS_HIP_OK;
    return stage_l2l(e);
}

int stage_p2p(ps_hip_engine* e, int mode, int ev_before = -1, const float4* leaf_loc = nullptr) {
    unsigned long long* pc = e->count_on ? e->pairs() : nullptr;
    // candidates for the wavefront-per-leaf kernel: every node (leaves of
    // more than one particle sit at depth 10 unless leaf_cap > 1)
    const uint32_t bound = e->n_nodes;
    if (e->count_on)
        PS_CK(grow(e->block_pairs_b, 16 * (size_t)ps_p2p_blocks(e->n, bound), e->s));
    if (!e->arena_b.p) PS_CK(grow(e->arena_b, 256, e->s));
    if (!e->masks_b.p) PS_CK(grow(e->masks_b, 256, e->s));
    PS_CK(grow(e->big_b, 4 * ((size_t)bound + 2), e->s));
    if (ev_before >= 0) e->record(ev_before);
    ps_launch_p2p(e->n, e->pos4(), e->leaf_of(), leaf_loc ? leaf_loc : e->loc(), e->p2p_off(),
                  e->p2p_len(), e->p2p_moff(), e->arena(), e->masks(), e->link(), e->nend(),
                  e->fx, e->fy, e->fz, mode, e->kahan, e->small_u32(), bound, e->p2p_big,
                  reinterpret_cast<uint32_t*>(e->big_b.p),
                  reinterpret_cast<unsigned long long*>(e->block_pairs_b.p), pc, e->s, e->s2,
                  e->fork, e->join, e->collected);
    return map_err(hipGetLastError());
}

// Later work on the engine stream waits for the second stream's side work
// (the traversal's pair count and state copy, the big-leaf list).
hipError_t join_side(ps_hip_engine* e) {
    hipError_t r = hipEventRecord(e->ev_side, e->s2);
    return r == hipSuccess ? hipStreamWaitEvent(e->s, e->ev_side, 0) : r;
}

// The big-leaf list on the second stream, as soon as the tree is written
// (the P2P stage forks its big-leaf kernel onto the same stream after it).
int stage_collect(ps_hip_engine* e) {
    e->collected = false;
    if (!e->p2p_big || !e->n_nodes) return PS_HIP_OK;
    PS_CK(grow(e->big_b, 4 * ((size_t)e->n_nodes + 2), e->s));
    PS_CK(hipEventRecord(e->ev_tree, e->s));
    PS_CK(hipStreamWaitEvent(e->s2, e->ev_tree, 0));
    ps_launch_p2p_collect(e->small_u32(), e->n_nodes, e->p2p_big, e->link(), e->nend(),
                          reinterpret_cast<uint32_t*>(e->big_b.p), e->s2);
    PS_CK(hipGetLastError());
    e->collected = true;
    return PS_HIP_OK;
}

// traversal, L2L (or the leaves' locals) and P2P
int traverse_to_p2p(ps_hip_engine* e) {
    e->record(4);
    PS_RC(stage_traverse(e));
    e->record(5);
    if (e->mlp_on) {  // the correction's features read internal nodes' locals
        PS_RC(stage_l2l(e));
        PS_RC(stage_p2p(e, 0, 6));  // records event 6 right before the P2P launch
    } else {
        PS_RC(stage_leaf_locals(e));
        PS_RC(stage_p2p(e, 0, 6, e->ltot()));
    }
    return PS_HIP_OK;
}

// the tick from the traversal on (also its re-run after an arena overflow)
int run_from_traverse(ps_hip_engine* e) {
    const int rc = traverse_to_p2p(e);
    // the second stream's work joins the engine stream on every path, so
    // nothing of this tick still runs there once the engine stream drains
    const hipError_t rj = join_side(e);
    if (rc != PS_HIP_OK) return rc;
    PS_CK(rj);
    e->record(7);
    if (e->mlp_on) {
        ps_launch_mlp_apply(e->n, e->mlp_pk_b.p, e->pos4(),
                            e->leaf_of(), e->geom(), e->mp(), e->loc(), e->parent(), e->fx, e->fy,
                            e->fz, e->s);
        PS_CK(hipGetLastError());
    }
    // the caller's arrays are sorted before the tick's last event
    PS_RC(finish_copy(e));
    e->record(8);
    return PS_HIP_OK;
}

// Launches the whole tick; one host wait inside (the tree's counts, while
// the GPU writes the tree).
int run_tick(ps_hip_engine* e) {
    e->collected = false;
    e->record(1);
    PS_RC(stage_sort(e));
    e->record(2);
    int rc = stage_build(e);
    e->record(3);
    if (rc == PS_HIP_OK) rc = stage_collect(e);
    if (rc == PS_HIP_OK) rc = stage_upward(e);
    if (rc == PS_HIP_OK)
        rc = run_from_traverse(e);
    else
        (void)join_side(e);  // a failed stage: the side stream's work still joins
    const int rj = finish_copy(e);  // joined already unless a stage failed
    return rc != PS_HIP_OK ? rc : rj;
}

// ---------------------------------------------------------------------------
// The asynchronous tick: the same stages as run_tick with every size read
// from device memory (as the batched tick does), so nothing waits on the
// host between the launches.  The node arrays are sized for the most nodes
// n particles can give (worst_case_nodes), so the tree's real size, known
// only on the device, always fits; the traversal launches all
// PS_MAX_DEPTH_ITERS iterations (the ones past the tree's depth find no
// entries) and the L2L runs level by level over the node locals (bit for
// bit the leaves' sums of the one-launch form).
// ---------------------------------------------------------------------------

int async_ensure(ps_hip_engine* e, size_t n) {
    hipStream_t s = e->s;
    const size_t M = worst_case_nodes(n);
    DevBuf* node16[] = {&e->geom_b, &e->tpk_b, &e->mp_b, &e->loc_b};
    for (DevBuf* d : node16) PS_CK(grow(*d, 16 * M, s));
    PS_CK(grow(e->link_b, 8 * M, s));
    PS_CK(grow(e->oct_b, M, s));
    DevBuf* node4[] = {&e->nend_b, &e->parent_b, &e->p2p_off_b, &e->p2p_len_b, &e->p2p_moff_b};
    for (DevBuf* d : node4) PS_CK(grow(*d, 4 * M, s));
    for (auto& es : e->ent)
        for (auto& b : es) PS_CK(grow(b, 4 * M, s));
    PS_CK(grow(e->far_b, 8 * (size_t)e->trav_blocks * PS_MAX_DEPTH_ITERS, s));
    PS_CK(grow(e->tstate_b, sizeof(PsTravState), s));
    PS_CK(grow(e->arena_b, 4 * ((size_t)e->arena_words * n + 4096), s));
    PS_CK(grow(e->masks_b, 8 * (4 * M + 4096), s));
    PS_CK(grow(e->big_b, 4 * (M + 2), s));
    e->node_alloc = M;
    return PS_HIP_OK;
}

// sort, build (locals and P2P list lengths cleared), upward
int async_front(ps_hip_engine* e, hipStream_t cs) {
    const size_t M = e->node_alloc;
    uint32_t* ls = e->small_u32();
    PsSortArrays io = {e->code, e->id, e->x, e->y, e->z, e->m};
    // the full radix sort (the incremental one would need a readback); its
    // keys are kept for a later synchronous tick
    const int nxt = e->prev_cur ^ 1;
    PS_CK(ps_radix_sort(io, e->sort_set(0), e->sort_set(1), e->pos4(), e->n, e->sort_tmp.p, cs,
                        e->prev_keys(nxt)));
    e->prev_cur = nxt;
    e->prev_n = e->n;
    PS_CK(ps_tree_count(e->code, nullptr, e->n, ls, ls + 12, e->totals(), e->tree_tmp.p,
                        e->leaf_cap, cs));
    PS_CK(ps_tree_write(e->code, nullptr, e->n, e->root, nullptr, e->tree_tmp.p, e->geom(),
                        e->link(), e->nend(), e->parent(), e->leaf_of(), e->oct(), M,
                        e->totals(), e->leaf_cap, cs, nullptr, e->loc(), e->p2p_len()));
    ps_launch_upward_dev(ls, e->tpk(), e->pos4(), e->geom(), e->link(), e->nend(), e->mp(), cs);
    return map_err(hipGetLastError());
}

// traversal (M2L + P2P lists), L2L, P2P + L2P, the correction; then the
// counts and the traversal's state to the host mirrors, and ev_async
int async_back(ps_hip_engine* e, hipStream_t cs) {
    const uint32_t ub = (uint32_t)std::min<size_t>(e->node_alloc, 0xFFFFFFFFull);
    uint32_t* ls = e->small_u32();
    const float theta2 = theta_squared(e);
    const bool mt = e->trav == PS_HIP_TRAV_MT;
    PsTravState* st = reinterpret_cast<PsTravState*>(e->tstate_b.p);
    const uint32_t arena_cap = list_cap_words(e->arena_b);
    const uint32_t mask_cap = (uint32_t)std::min<size_t>(e->masks_b.bytes / 8, 0xFFFFFFFFull);
    ps_launch_trav_init(ls, mt ? 1 : 0, 1u, nullptr, e->nend(), e->link(), e->entries(0), st,
                        e->arena(), arena_cap, mask_cap, ub, cs);
    PsTravTree t;
    t.n_nodes = ub;
    t.geom = e->geom();
    t.tpk = e->tpk();
    t.mp = e->mp();
    t.link = e->link();
    t.nend = e->nend();
    t.loc = e->loc();
    t.p2p_off = e->p2p_off();
    t.p2p_len = e->p2p_len();
    t.p2p_moff = e->p2p_moff();
    unsigned long long* far_part = reinterpret_cast<unsigned long long*>(e->far_b.p);
    size_t far_n = 0;
    for (int k = 0; k < PS_MAX_DEPTH_ITERS; ++k) {
        const int cur = k & 1, nxt = cur ^ 1;
        const int lanes = trav_lanes_for(e, ub);
        ps_launch_trav_step(t, theta2, e->entries(cur), e->arena(), e->masks(), st, k, ub,
                            e->entries(nxt), far_part + far_n, e->kahan, lanes, e->trav_blocks,
                            e->trav_spb, rec_cap_of(e), cs);
        far_n += ps_trav_blocks(ub, lanes, e->trav_blocks);
    }
    ps_reduce_u64(far_part, far_n, &st->m2l_pairs, cs);
    ps_launch_l2l_dev(ls, e->parent(), e->loc(), cs);
    ps_launch_p2p(e->n, e->pos4(), e->leaf_of(), e->loc(), e->p2p_off(), e->p2p_len(),
                  e->p2p_moff(), e->arena(), e->masks(), e->link(), e->nend(), e->fx, e->fy,
                  e->fz, 0, e->kahan, ls, ub, e->p2p_big, reinterpret_cast<uint32_t*>(e->big_b.p),
                  nullptr, nullptr, cs, e->s2, e->fork, e->join);
    if (e->mlp_on)
        ps_launch_mlp_apply(e->n, e->mlp_pk_b.p, e->pos4(), e->leaf_of(), e->geom(), e->mp(),
                            e->loc(), e->parent(), e->fx, e->fy, e->fz, cs);
    PS_CK(hipGetLastError());
    PS_CK(hipMemcpyAsync(e->h_small, e->small_b.p, 96, hipMemcpyDeviceToHost, cs));
    PS_CK(hipMemcpyAsync(e->h_tstate, e->tstate_b.p, sizeof(PsTravState), hipMemcpyDeviceToHost,
                         cs));
    PS_CK(hipEventRecord(e->ev_async, cs));
    return PS_HIP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched ticks: many systems in one set of device arrays, one forest.
// ---------------------------------------------------------------------------
struct ps_hip_batch {
    ps_hip_engine* e = nullptr;  // streams, options and the tick buffers
    size_t max_sys = 0;
    DevBuf off_b;                // u32[n_sys + 1] particle offsets of the systems
    DevBuf sys_b;                // u32 per particle: its system (the sort's second key)
    DevBuf aux_b;                // two u32 arrays: the second key's sort scratch
    DevBuf prm_b, roots_b;       // float4 per system
    DevBuf caps_b;               // u8 per system: leaf capacities (when caps_n > 0)
    DevBuf seg_b;                // the segmented sort's tile tables (vstart, vseg, seg_tile)
    std::vector<uint32_t> seg_h;  // their host copy (the upload's source)
    PsSortSegs segs{};           // n_tiles 0: the sort takes the system id as a fourth pass
    size_t caps_n = 0;           // systems covered by caps_b (0: the engine option)
    size_t node_cap = 0;         // guaranteed bound for the current layout
    int use_graph = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // what the captured graph was recorded for
    std::vector<uint32_t> g_off;
    void* g_ptr[9] = {nullptr};
    size_t g_n = 0;
    int g_mode = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ps_hip_batch_stats_t st;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        g_mode = -1;
    }
};

namespace {

// most nodes of a forest whose system k holds off[k+1] - off[k] particles
size_t forest_node_bound(const std::vector<uint32_t>& off) {
    size_t total = 0;
    for (size_t k = 0; k + 1 < off.size(); ++k) total += worst_case_nodes(off[k + 1] - off[k]);
    return total;
}

// Systems up to this many sort segment by segment (three passes, tile tables
// and 4 KiB of digit starts per system); more take the system id as a
// fourth pass.
constexpr uint32_t kSegSortMaxSys = 4096;

// The segmented sort's tables for a layout (off: n_sys + 1 offsets), on the
// device; the sort scratch grown to fit.
int batch_segments(ps_hip_batch* b, const std::vector<uint32_t>& off, uint32_t n_sys) {
    ps_hip_engine* e = b->e;
    hipStream_t s = e->s;
    b->segs = PsSortSegs{};
    if (n_sys == 0 || n_sys > kSegSortMaxSys) return PS_HIP_OK;
    const size_t t = ps_sort_seg_tiles(off.data(), n_sys, nullptr, nullptr, nullptr);
    if (t == 0 || t > 0xFFFFFFFEull) return PS_HIP_OK;
    b->seg_h.assign((t + 1) + t + (n_sys + 1), 0u);
    uint32_t* vstart = b->seg_h.data();
    uint32_t* vseg = vstart + t + 1;
    uint32_t* seg_tile = vseg + t;
    ps_sort_seg_tiles(off.data(), n_sys, vstart, vseg, seg_tile);
    PS_CK(grow(b->seg_b, 4 * b->seg_h.size(), s));
    PS_CK(grow(e->sort_tmp, ps_sort_seg_tmp_bytes(t, n_sys), s));
    PS_CK(hipMemcpyAsync(b->seg_b.p, b->seg_h.data(), 4 * b->seg_h.size(),
                         hipMemcpyHostToDevice, s));
    // the copy reads seg_h: it stays unchanged until the next layout
    const uint32_t* d = reinterpret_cast<const uint32_t*>(b->seg_b.p);
    b->segs.off = reinterpret_cast<const uint32_t*>(b->off_b.p);
    b->segs.n_sys = n_sys;
    b->segs.vstart = d;
    b->segs.vseg = d + t + 1;
    b->segs.seg_tile = d + 2 * t + 1;
    b->segs.n_tiles = (uint32_t)t;
    return PS_HIP_OK;
}

// every buffer of a batch tick over n particles and `cap` nodes; no
// allocation may happen while a graph is being recorded
int batch_ensure(ps_hip_batch* b, size_t n, size_t cap) {
    ps_hip_engine* e = b->e;
    hipStream_t s = e->s;
    DevBuf* node16[] = {&e->geom_b, &e->tpk_b, &e->mp_b, &e->loc_b};
    for (DevBuf* d : node16) PS_CK(grow(*d, 16 * cap, s));
    PS_CK(grow(e->link_b, 8 * cap, s));
    PS_CK(grow(e->oct_b, cap, s));
    DevBuf* node4[] = {&e->nend_b, &e->parent_b, &e->p2p_off_b, &e->p2p_len_b, &e->p2p_moff_b};
    for (DevBuf* d : node4) PS_CK(grow(*d, 4 * cap, s));
    for (int g = 0; g < 2; ++g)
        for (int k = 0; k < 4; ++k) PS_CK(grow(e->ent[g][k], 4 * cap, s));
    PS_CK(grow(e->far_b, 8 * (size_t)e->trav_blocks * PS_MAX_DEPTH_ITERS, s));
    PS_CK(grow(e->tstate_b, sizeof(PsTravState), s));
    PS_CK(grow(e->arena_b, 4 * ((size_t)e->arena_words * n + 4096), s));
    PS_CK(grow(e->masks_b, 8 * (4 * cap + 4096), s));
    PS_CK(grow(e->big_b, 4 * (cap + 2), s));
    PS_CK(grow(b->sys_b, 4 * std::max<size_t>(n, 1), s));
    PS_CK(grow(b->aux_b, 8 * std::max<size_t>(e->max_p, 1), s));
    e->node_cap = std::max(e->node_cap, cap);
    return PS_HIP_OK;
}

// The batch tick up to the multipoles, launches only: per-system prepare,
// sort by (system, code), forest build (locals and P2P list lengths
// cleared), upward.  It sorts the caller's arrays in place, so it runs once
// per tick: in the AVX2 mode a particle's quantisation depends on its
// position in its system (the first 8 floor(N_s/8) are rounded), and a
// second prepare over the sorted arrays would give other codes.
int batch_run_front(ps_hip_batch* b, uint32_t n_sys, int morton_mode) {
    ps_hip_engine* e = b->e;
    hipStream_t s = e->s;
    const size_t n = e->n;
    const size_t cap = b->node_cap;
    const uint32_t* off = reinterpret_cast<const uint32_t*>(b->off_b.p);
    uint32_t* sys = reinterpret_cast<uint32_t*>(b->sys_b.p);
    float4* prm = reinterpret_cast<float4*>(b->prm_b.p);
    float4* roots = reinterpret_cast<float4*>(b->roots_b.p);
    uint32_t* ls = e->small_u32();
    ps_launch_prepare_seg(e->x, e->y, e->z, n, off, n_sys, prm, roots, sys, e->code, morton_mode,
                          s);
    PsSortArrays io = {e->code, e->id, e->x, e->y, e->z, e->m, sys};
    PsSortArrays t0 = e->sort_set(0), t1 = e->sort_set(1);
    if (b->segs.n_tiles) {  // segment by segment: every particle stays in its system
        PS_CK(ps_radix_sort_seg(io, t0, t1, e->pos4(), n, b->segs, e->sort_tmp.p, s));
    } else {
        t0.aux = reinterpret_cast<uint32_t*>(b->aux_b.p);
        t1.aux = t0.aux + e->max_p;
        PS_CK(ps_radix_sort(io, t0, t1, e->pos4(), n, e->sort_tmp.p, s));
    }
    const uint8_t* caps = b->caps_n ? reinterpret_cast<const uint8_t*>(b->caps_b.p) : nullptr;
    PS_CK(ps_tree_count(e->code, sys, n, ls, ls + 12, e->totals(), e->tree_tmp.p, e->leaf_cap,
                        s, caps));
    PS_CK(ps_tree_write(e->code, sys, n, make_float4(0.f, 0.f, 0.f, 0.f), roots, e->tree_tmp.p,
                        e->geom(), e->link(), e->nend(), e->parent(), e->leaf_of(), e->oct(),
                        cap, e->totals(), e->leaf_cap, s, caps, e->loc(), e->p2p_len()));
    ps_launch_upward_dev(ls, e->tpk(), e->pos4(), e->geom(), e->link(), e->nend(), e->mp(), s);
    return map_err(hipGetLastError());
}

// The batch tick from the traversal on (all PS_MAX_DEPTH_ITERS iterations;
// the unused ones find no entries), L2L, P2P + L2P, the correction.  Needs
// loc and p2p_len cleared (batch_run_front, or batch_clear_far before a
// re-run after an arena overflow).
int batch_run_back(ps_hip_batch* b, uint32_t n_sys) {
    ps_hip_engine* e = b->e;
    hipStream_t s = e->s;
    const size_t n = e->n;
    const size_t cap = b->node_cap;
    uint32_t* sys = reinterpret_cast<uint32_t*>(b->sys_b.p);
    uint32_t* ls = e->small_u32();
    float theta2;
    {
        volatile float tv = e->theta;  // fl(theta * theta), H:1064
        theta2 = tv * tv;
    }
    const bool mt = e->trav == PS_HIP_TRAV_MT;
    PsTravState* st = reinterpret_cast<PsTravState*>(e->tstate_b.p);
    const uint32_t arena_cap = list_cap_words(e->arena_b);
    const uint32_t mask_cap = (uint32_t)std::min<size_t>(e->masks_b.bytes / 8, 0xFFFFFFFFull);
    const uint32_t ub = (uint32_t)std::min<size_t>(cap, 0xFFFFFFFFull);
    ps_launch_trav_init(ls, mt ? 1 : 0, n_sys, sys, e->nend(), e->link(), e->entries(0), st,
                        e->arena(), arena_cap, mask_cap, ub, s);
    PsTravTree t;
    t.n_nodes = ub;
    t.geom = e->geom();
    t.tpk = e->tpk();
    t.mp = e->mp();
    t.link = e->link();
    t.nend = e->nend();
    t.loc = e->loc();
    t.p2p_off = e->p2p_off();
    t.p2p_len = e->p2p_len();
    t.p2p_moff = e->p2p_moff();
    unsigned long long* far_part = reinterpret_cast<unsigned long long*>(e->far_b.p);
    size_t far_n = 0;
    for (int k = 0; k < PS_MAX_DEPTH_ITERS; ++k) {
        const int cur = k & 1, nxt = cur ^ 1;
        const int lanes = trav_lanes_for(e, ub);
        ps_launch_trav_step(t, theta2, e->entries(cur), e->arena(), e->masks(), st, k, ub,
                            e->entries(nxt), far_part + far_n, e->kahan, lanes,
                            e->trav_blocks, e->trav_spb, rec_cap_of(e), s);
        far_n += ps_trav_blocks(ub, lanes, e->trav_blocks);
    }
    ps_reduce_u64(far_part, far_n, &st->m2l_pairs, s);
    ps_launch_l2l_dev(ls, e->parent(), e->loc(), s);
    ps_launch_p2p(n, e->pos4(), e->leaf_of(), e->loc(), e->p2p_off(), e->p2p_len(),
                  e->p2p_moff(), e->arena(), e->masks(), e->link(), e->nend(), e->fx, e->fy,
                  e->fz, 0, e->kahan, ls, ub, e->p2p_big,
                  reinterpret_cast<uint32_t*>(e->big_b.p), nullptr, nullptr, s, e->s2, e->fork,
                  e->join);
    if (e->mlp_on)
        ps_launch_mlp_apply(n, e->mlp_pk_b.p, e->pos4(),
                            e->leaf_of(), e->geom(), e->mp(), e->loc(), e->parent(), e->fx, e->fy,
                            e->fz, s);
    return map_err(hipGetLastError());
}

// The whole batch tick (launches only).
int batch_run(ps_hip_batch* b, uint32_t n_sys, int morton_mode) {
    PS_RC(batch_run_front(b, n_sys, morton_mode));
    return batch_run_back(b, n_sys);
}

// Clears what the traversal accumulates (node locals, P2P list lengths) for
// a re-run of batch_run_back over the same sorted state and forest.
void batch_clear_far(ps_hip_batch* b) {
    ps_hip_engine* e = b->e;
    ps_fill2_u32_count(reinterpret_cast<uint32_t*>(e->loc()), 4, e->p2p_len(), 1, 0u,
                       e->totals(), b->node_cap, e->s);
}

struct PrepWS {
    int dev = -1;
    hipStream_t s = nullptr;
    DevBuf part, prm, x, y, z, code;
    float* h_out = nullptr;
};

std::mutex g_prep_mu;
std::vector<PrepWS*> g_prep;

PrepWS* prep_ws(int dev, hipError_t* err) {
    for (PrepWS* w : g_prep)
        if (w->dev == dev) return w;
    PrepWS* w = new (std::nothrow) PrepWS();
    if (!w) {
        *err = hipErrorOutOfMemory;
        return nullptr;
    }
    w->dev = dev;
    if ((*err = hipStreamCreateWithFlags(&w->s, hipStreamNonBlocking)) != hipSuccess ||
        (*err = grow(w->part, 1024 * sizeof(float2), w->s)) != hipSuccess ||
        (*err = grow(w->prm, 16, w->s)) != hipSuccess ||
        (*err = host_malloc(reinterpret_cast<void**>(&w->h_out), 64)) !=
            hipSuccess) {
        delete w;
        return nullptr;
    }
    g_prep.push_back(w);
    return w;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

size_t ps_hip_engine_host_bytes(void) {
    return align256(sizeof(ps_hip_engine)) + 256 + align256(sizeof(PsTravState)) + 256;
}

size_t ps_hip_engine_scratch_bytes(size_t max_particles) {
    const size_t P = std::max<size_t>(max_particles, 1);
    return tick_scratch_bytes(P, worst_case_nodes(P), 96);
}

size_t ps_hip_worst_case_nodes(size_t n) { return worst_case_nodes(n); }

int ps_hip_engine_create_in(ps_hip_engine_t** out, size_t max_particles, size_t max_nodes,
                            float theta, int traversal, void* host_mem, size_t host_bytes,
                            void* dev_mem, size_t dev_bytes) {
    if (!out) return PS_HIP_EINVAL;
    *out = nullptr;
    if (dev_mem && !is_device_ptr(dev_mem)) return PS_HIP_EINVAL;
    ps_hip_engine* e = nullptr;
    char* host_tail = nullptr;
    if (host_mem && host_bytes >= ps_hip_engine_host_bytes()) {
        // the engine and its two host mirrors in the caller's memory
        char* h = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(host_mem) + 255) &
                                          ~(uintptr_t)255);
        e = new (h) ps_hip_engine();
        e->placed = true;
        host_tail = h + align256(sizeof(ps_hip_engine));
    } else {
        e = new (std::nothrow) ps_hip_engine();
        if (!e) return PS_HIP_EOOM;
    }
    std::memset(&e->st, 0, sizeof(e->st));
    e->max_p = std::max<size_t>(max_particles, 1);
    // the traversal's 32-bit node offsets bound every tree
    e->node_cap = std::max<size_t>(
        std::min({max_nodes, worst_case_nodes(e->max_p), (size_t)PS_TRAV_MAX_NODES}), 1);
    e->theta = theta > 0.0f ? theta : 1.0f;
    e->trav = traversal == PS_HIP_TRAV_ST ? PS_HIP_TRAV_ST : PS_HIP_TRAV_MT;
    if (host_tail) {
        e->h_small = reinterpret_cast<uint32_t*>(host_tail);
        e->h_tstate = reinterpret_cast<PsTravState*>(host_tail + 256);
    }
    hipError_t r = hipGetDevice(&e->device);
    int cus = 0;
    if (r == hipSuccess)
        r = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device);
    if (r == hipSuccess && cus >= 8) e->trav_blocks = std::min<uint32_t>((uint32_t)cus & ~7u,
                                                                         kMaxTravBlocks);
    if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking);
    if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->s2, hipStreamNonBlocking);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&e->join, hipEventDisableTiming);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&e->ev_counts, hipEventDisableTiming);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&e->ev_async, hipEventDisableTiming);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&e->ev_sorted, hipEventDisableTiming);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&e->ev_copied, hipEventDisableTiming);
    for (hipEvent_t* v : {&e->ev_tree, &e->ev_trav, &e->ev_side})
        if (r == hipSuccess) r = hipEventCreateWithFlags(v, hipEventDisableTiming);
    for (int k = 0; k < 12 && r == hipSuccess; ++k) r = hipEventCreate(&e->ev[k]);
    for (int k = 0; k < 64 && r == hipSuccess; ++k) r = hipEventCreate(&e->ev_it[k]);
    int rc = PS_HIP_OK;
    if (r == hipSuccess) r = alloc_mirrors(e);
    if (r == hipSuccess) {
        if (dev_mem)
            rc = carve_all(e, dev_mem, dev_bytes);
        else
            r = alloc_fixed(e);
    }
    if (r != hipSuccess || rc != PS_HIP_OK) {
        (void)hipGetLastError();
        e->free_all();
        if (e->placed)
            e->~ps_hip_engine();
        else
            delete e;
        if (rc != PS_HIP_OK) return rc;
        return map_err(r) == PS_HIP_EINVAL ? PS_HIP_EDEVICE : map_err(r);
    }
    *out = e;
    return PS_HIP_OK;
}

int ps_hip_engine_create(ps_hip_engine_t** out, size_t max_particles, size_t max_nodes,
                         float theta, int traversal) {
    return ps_hip_engine_create_in(out, max_particles, max_nodes, theta, traversal, nullptr, 0,
                                   nullptr, 0);
}

int ps_hip_engine_destroy(ps_hip_engine_t* e) {
    if (!e) return PS_HIP_EINVAL;
    if (e->s) (void)hipStreamSynchronize(e->s);
    if (e->pending && e->async_s) (void)hipStreamSynchronize(e->async_s);
    e->free_all();
    if (e->placed)
        e->~ps_hip_engine();
    else
        delete e;
    return PS_HIP_OK;
}

int ps_hip_debug_fail(int what, int after) {
    const int v = after < 0 ? -1 : after;
    switch (what) {
        case PS_HIP_FAIL_OFF:
            g_fail_alloc.store(-1);
            g_fail_wait.store(-1);
            return PS_HIP_OK;
        case PS_HIP_FAIL_ALLOC: g_fail_alloc.store(v); return PS_HIP_OK;
        case PS_HIP_FAIL_WAIT: g_fail_wait.store(v); return PS_HIP_OK;
        default: return PS_HIP_EINVAL;
    }
}

int ps_hip_alloc_stats(uint64_t* calls, uint64_t* bytes) {
    if (calls) *calls = g_alloc_calls.load(std::memory_order_relaxed);
    if (bytes) *bytes = g_alloc_bytes.load(std::memory_order_relaxed);
    return PS_HIP_OK;
}

int ps_hip_engine_set_option(ps_hip_engine_t* e, int key, double value) {
    if (!e) return PS_HIP_EINVAL;
    switch (key) {
        case PS_HIP_OPT_TRAVERSAL:
            if (value != PS_HIP_TRAV_MT && value != PS_HIP_TRAV_ST) return PS_HIP_EINVAL;
            e->trav = (int)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_KAHAN:
            e->kahan = value != 0.0;
            return PS_HIP_OK;
        case PS_HIP_OPT_THETA:
            e->theta = value > 0.0 ? (float)value : 1.0f;
            return PS_HIP_OK;
        case PS_HIP_OPT_STATS:
            if (value != 0.0 && value != 1.0 && value != 2.0 && value != 3.0)
                return PS_HIP_EINVAL;
            e->stats_on = value == 1.0 || value == 2.0;
            e->count_on = value == 1.0;
            e->trav_ev = value != 0.0;
            // mode 3: the traversal and the P2P stage only (four events)
            e->ev_mask = e->stats_on ? 0x3FFu : (value == 3.0 ? (1u << 6) | (1u << 7) : 0u);
            return PS_HIP_OK;
        case PS_HIP_OPT_TRAV_LANES:
            if (value != 0 && value != 8 && value != 16 && value != 32 && value != 64)
                return PS_HIP_EINVAL;
            e->trav_lanes = (int)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_LEAF_CAP:
            if (!(value >= 1.0 && value <= 64.0) || value != (double)(int)value) return PS_HIP_EINVAL;
            e->leaf_cap = (int)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_P2P_BIG:
            if (!(value >= 0.0 && value <= 4294967295.0)) return PS_HIP_EINVAL;
            e->p2p_big = (uint32_t)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_TRAV_BLOCKS:
            if (!(value >= 8.0 && value <= 8192.0) || value != (double)(int)value ||
                ((int)value & 7))
                return PS_HIP_EINVAL;
            e->trav_blocks = (uint32_t)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_TRAV_SEGMENTS:
            if (!(value >= 1.0 && value <= 64.0) || value != (double)(int)value)
                return PS_HIP_EINVAL;
            e->trav_spb = (uint32_t)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_TRAV_RECORDS:
            if (value != 0.0 && value != 1.0) return PS_HIP_EINVAL;
            e->trav_records = (int)value;
            return PS_HIP_OK;
        case PS_HIP_OPT_SORT_INCREMENTAL:
            if (value != 0.0 && value != 1.0) return PS_HIP_EINVAL;
            e->incremental = value != 0.0;
            return PS_HIP_OK;
        case PS_HIP_OPT_ARENA_WORDS:
            if (!(value >= 1.0 && value <= 65536.0) || value != (double)(int)value)
                return PS_HIP_EINVAL;
            e->arena_words = (uint32_t)value;
            return PS_HIP_OK;
        default:
            return PS_HIP_EINVAL;
    }
}

static_assert(PS_HIP_PREPARE_NOALLOC_SCRATCH == PS_PREP_SCRATCH_BYTES,
              "the documented scratch size is the array's");
static_assert(PS_HIP_MLP_PARAMS == PS_MLP_PARAMS && PS_HIP_MLP_IN == PS_MLP_IN,
              "the public and kernel layouts of the correction agree");

int ps_hip_engine_set_mlp(ps_hip_engine_t* e, const float* params) {
    if (!e) return PS_HIP_EINVAL;
    if (!params) {
        e->mlp_on = false;
        return PS_HIP_OK;
    }
    PS_CK(grow(e->mlp_b, 4 * PS_MLP_PARAMS, e->s));
    PS_CK(grow(e->mlp_pk_b, PS_MLP_PACKED_BYTES, e->s));
    PS_CK(hipMemcpyAsync(e->mlp_b.p, params, 4 * PS_MLP_PARAMS, hipMemcpyDefault, e->s));
    ps_launch_mlp_pack(reinterpret_cast<const float*>(e->mlp_b.p), e->mlp_pk_b.p, e->s);
    PS_CK(hipGetLastError());
    PS_CK(wait_stream(e->s));
    e->mlp_on = true;
    return PS_HIP_OK;
}

int ps_hip_mlp_features(ps_hip_engine_t* e, float* feat, float* scale) {
    if (!e || !feat || !scale || !e->built) return PS_HIP_EINVAL;
    if (!is_device_ptr(feat) || !is_device_ptr(scale)) return PS_HIP_EINVAL;
    PS_RC(materialize_locals(e));
    ps_launch_mlp_features(e->n, nullptr, e->n, e->pos4(), e->leaf_of(), e->geom(), e->mp(),
                           e->loc(), e->parent(), feat, scale, e->s);
    PS_CK(hipGetLastError());
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_mlp_features_at(ps_hip_engine_t* e, const uint32_t* idx, size_t rows, float* feat,
                           float* scale) {
    if (!e || !idx || !feat || !scale || !e->built) return PS_HIP_EINVAL;
    if (rows == 0) return PS_HIP_OK;
    if (!is_device_ptr(idx) || !is_device_ptr(feat) || !is_device_ptr(scale))
        return PS_HIP_EINVAL;
    PS_RC(materialize_locals(e));
    ps_launch_mlp_features(e->n, idx, rows, e->pos4(), e->leaf_of(), e->geom(), e->mp(),
                           e->loc(), e->parent(), feat, scale, e->s);
    PS_CK(hipGetLastError());
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_mlp_forward(ps_hip_engine_t* e, const float* params, const float* feat, size_t n,
                       float* out3) {
    if (!e || !params || !feat || !out3) return PS_HIP_EINVAL;
    // the kernels read and write these directly: host memory would fault
    if (!is_device_ptr(params) || !is_device_ptr(feat) || !is_device_ptr(out3))
        return PS_HIP_EINVAL;
    ps_launch_mlp_forward(params, feat, n, out3, e->s);
    PS_CK(hipGetLastError());
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_mlp_grad(ps_hip_engine_t* e, const float* params, const float* feat,
                    const float* target, size_t n, float* grad, float* loss_host) {
    if (!e || !params || !feat || !target || !grad || n == 0) return PS_HIP_EINVAL;
    if (!is_device_ptr(params) || !is_device_ptr(feat) || !is_device_ptr(target) ||
        !is_device_ptr(grad))
        return PS_HIP_EINVAL;
    PS_CK(grow(e->mlp_tmp, 4 * (ps_mlp_grad_tmp_floats(n) + 1), e->s));
    float* tmp = reinterpret_cast<float*>(e->mlp_tmp.p);
    float* loss_dev = tmp + ps_mlp_grad_tmp_floats(n);
    ps_launch_mlp_grad(params, feat, target, n, grad, loss_dev, tmp, e->s);
    PS_CK(hipGetLastError());
    float sse = 0.f;
    PS_CK(hipMemcpyAsync(&sse, loss_dev, 4, hipMemcpyDeviceToHost, e->s));
    PS_CK(wait_stream(e->s));
    if (loss_host) *loss_host = sse / (float)n;
    return PS_HIP_OK;
}

int ps_hip_mlp_adam(ps_hip_engine_t* e, float* params, const float* grad, float* m1, float* m2,
                    float lr, float beta1, float beta2, float eps, int step) {
    if (!e || !params || !grad || !m1 || !m2 || step < 1) return PS_HIP_EINVAL;
    if (!is_device_ptr(params) || !is_device_ptr(grad) || !is_device_ptr(m1) ||
        !is_device_ptr(m2))
        return PS_HIP_EINVAL;
    ps_launch_mlp_adam(params, grad, m1, m2, lr, beta1, beta2, eps, step, e->s);
    PS_CK(hipGetLastError());
    return PS_HIP_OK;
}

int ps_hip_batch_create(ps_hip_batch_t** out, size_t max_particles, size_t max_systems,
                        float theta, int traversal) {
    if (!out || max_particles == 0 || max_systems == 0 || max_systems > 1024)
        return PS_HIP_EINVAL;
    *out = nullptr;
    ps_hip_batch* b = new (std::nothrow) ps_hip_batch();
    if (!b) return PS_HIP_EOOM;
    std::memset(&b->st, 0, sizeof(b->st));
    int rc = ps_hip_engine_create(&b->e, max_particles, 1, theta, traversal);
    if (rc != PS_HIP_OK) {
        delete b;
        return rc;
    }
    b->max_sys = max_systems;
    hipStream_t s = b->e->s;
    hipError_t r = grow(b->off_b, 4 * (max_systems + 1), s);
    if (r == hipSuccess) r = grow(b->prm_b, 16 * max_systems, s);
    if (r == hipSuccess) r = grow(b->roots_b, 16 * max_systems, s);
    if (r == hipSuccess) r = hipEventCreate(&b->ev0);
    if (r == hipSuccess) r = hipEventCreate(&b->ev1);
    if (r != hipSuccess) {
        (void)ps_hip_batch_destroy(b);
        return map_err(r);
    }
    *out = b;
    return PS_HIP_OK;
}

int ps_hip_batch_destroy(ps_hip_batch_t* b) {
    if (!b) return PS_HIP_EINVAL;
    if (b->e && b->e->s) (void)hipStreamSynchronize(b->e->s);
    b->drop_graph();
    DevBuf* bufs[] = {&b->off_b, &b->sys_b, &b->aux_b, &b->prm_b, &b->roots_b, &b->caps_b,
                      &b->seg_b};
    for (DevBuf* d : bufs) release(*d);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->e) (void)ps_hip_engine_destroy(b->e);
    delete b;
    return PS_HIP_OK;
}

int ps_hip_batch_set_option(ps_hip_batch_t* b, int key, double value) {
    if (!b) return PS_HIP_EINVAL;
    if (key == PS_HIP_BATCH_OPT_GRAPH) {
        if (value != 0.0 && value != 1.0) return PS_HIP_EINVAL;
        b->use_graph = value != 0.0;
        b->drop_graph();
        return PS_HIP_OK;
    }
    if (key == PS_HIP_OPT_STATS) return PS_HIP_EINVAL;  // a batch times whole ticks only
    b->drop_graph();  // the recorded launches carry the old settings
    return ps_hip_engine_set_option(b->e, key, value);
}

int ps_hip_batch_set_leaf_caps(ps_hip_batch_t* b, const uint8_t* caps, size_t n_systems) {
    if (!b || n_systems > b->max_sys) return PS_HIP_EINVAL;
    b->drop_graph();  // the recorded build carries the old capacities
    if (!caps || n_systems == 0) {
        b->caps_n = 0;
        return PS_HIP_OK;
    }
    for (size_t k = 0; k < n_systems; ++k)
        if (caps[k] < 1 || caps[k] > 64) return PS_HIP_EINVAL;
    PS_CK(grow(b->caps_b, b->max_sys, b->e->s));
    PS_CK(hipMemcpyAsync(b->caps_b.p, caps, n_systems, hipMemcpyHostToDevice, b->e->s));
    PS_CK(wait_stream(b->e->s));
    b->caps_n = n_systems;
    return PS_HIP_OK;
}

int ps_hip_batch_tick(ps_hip_batch_t* b, float* x, float* y, float* z, float* mass, float* fx,
                      float* fy, float* fz, uint32_t* id, uint32_t* codes,
                      const uint32_t* offsets, size_t n_systems, int morton_mode) {
    if (!b || !offsets || n_systems == 0 || n_systems > b->max_sys) return PS_HIP_EINVAL;
    if (morton_mode != PS_HIP_MORTON_TRUNC && morton_mode != PS_HIP_MORTON_AVX2)
        return PS_HIP_EINVAL;
    if (offsets[0] != 0) return PS_HIP_EINVAL;
    for (size_t k = 0; k < n_systems; ++k)
        if (offsets[k + 1] <= offsets[k]) return PS_HIP_EINVAL;  // every system non-empty
    const size_t n = offsets[n_systems];
    ps_hip_engine* e = b->e;
    if (n > e->max_p) return PS_HIP_EINVAL;
    if (b->caps_n && n_systems > b->caps_n) return PS_HIP_EINVAL;  // a system without a cap
    void* arrs[9] = {x, y, z, mass, fx, fy, fz, id, codes};
    for (void* p : arrs)
        if (!is_device_ptr(p)) return PS_HIP_EINVAL;  // a batch works on device arrays
    std::vector<uint32_t> off(offsets, offsets + n_systems + 1);
    e->stats_on = 0;
    e->count_on = 0;
    e->trav_ev = false;
    e->ev_mask = 0;
    e->n = n;
    e->x = x;
    e->y = y;
    e->z = z;
    e->m = mass;
    e->fx = fx;
    e->fy = fy;
    e->fz = fz;
    e->id = id;
    e->code = codes;
    e->bound = false;  // the single-system debug stages do not apply to a forest
    hipStream_t s = e->s;
    const uint32_t nsys = (uint32_t)n_systems;
    const bool same = b->exec && b->g_mode == morton_mode && b->g_n == n && b->g_off == off &&
                      std::equal(std::begin(arrs), std::end(arrs), std::begin(b->g_ptr));
    PS_CK(hipEventRecord(b->ev0, s));
    if (same) {
        PS_CK(hipGraphLaunch(b->exec, s));
    } else {
        b->drop_graph();
        b->node_cap = forest_node_bound(off);
        PS_RC(batch_ensure(b, n, b->node_cap));
        PS_CK(hipMemcpyAsync(b->off_b.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, s));
        PS_RC(batch_segments(b, off, nsys));
        PS_RC(batch_run(b, nsys, morton_mode));
    }
    PS_CK(hipEventRecord(b->ev1, s));
    PsTravState hst;
    unsigned long long nodes = 0;
    for (int attempt = 0;; ++attempt) {
        PS_CK(hipMemcpyAsync(&hst, e->tstate_b.p, sizeof(hst), hipMemcpyDeviceToHost, s));
        PS_CK(hipMemcpyAsync(&nodes, e->totals(), 8, hipMemcpyDeviceToHost, s));
        PS_CK(wait_stream(s));
        if (nodes > PS_TRAV_MAX_NODES) return PS_HIP_EOOM;  // past the traversal's offsets
        if (!hst.overflow) break;
        // an arena was too small: grow past the need and run the tick again
        // from the traversal (the sorted arrays and the forest stand; the
        // prepare and the sort must not run twice, see batch_run_front)
        b->drop_graph();
        if (attempt >= 4 || hst.arena_ctr > 0xFFFFFFFFull || hst.mask_ctr > 0xFFFFFFFFull)
            return PS_HIP_EOOM;
        PS_CK(grow(e->arena_b, std::max<size_t>(2 * e->arena_b.bytes, 5 * hst.arena_ctr), s));
        PS_CK(grow(e->masks_b, std::max<size_t>(2 * e->masks_b.bytes, 10 * hst.mask_ctr), s));
        PS_CK(hipEventRecord(b->ev0, s));
        batch_clear_far(b);
        PS_RC(batch_run_back(b, nsys));
        PS_CK(hipEventRecord(b->ev1, s));
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->st.n_particles = n;
    b->st.n_systems = n_systems;
    b->st.n_nodes = nodes;
    b->st.m2l_pairs = hst.m2l_pairs;
    b->st.used_graph = (same && b->exec) ? 1u : 0u;
    b->st.ms_tick = ms;
    if (b->st.used_graph) ++b->st.graph_replays;
    if (b->use_graph && !b->exec) {
        // record the tick for the next calls with this layout (nothing runs)
        hipGraph_t g = nullptr;
        PS_CK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = batch_run(b, nsys, morton_mode);
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != PS_HIP_OK || ec != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return rc != PS_HIP_OK ? rc : map_err(ec != hipSuccess ? ec : hipErrorUnknown);
        }
        hipGraphExec_t x_ = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x_, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return map_err(ei);
        }
        b->graph = g;
        b->exec = x_;
        b->g_off = off;
        b->g_n = n;
        b->g_mode = morton_mode;
        std::copy(std::begin(arrs), std::end(arrs), std::begin(b->g_ptr));
    }
    return PS_HIP_OK;
}

int ps_hip_batch_get_stats(const ps_hip_batch_t* b, ps_hip_batch_stats_t* out) {
    if (!b || !out) return PS_HIP_EINVAL;
    *out = b->st;
    return PS_HIP_OK;
}


int ps_hip_engine_get_stats(const ps_hip_engine_t* e, ps_hip_stats_t* out) {
    if (!e || !out) return PS_HIP_EINVAL;
    *out = e->st;
    return PS_HIP_OK;
}

int ps_hip_prepare(const float* x, const float* y, const float* z, size_t n, uint32_t* codes,
                   float* out_min, float* out_max, float* out_range, int morton_mode) {
    if (!x || !y || !z || !codes || n == 0) return PS_HIP_EINVAL;  // H:1914
    if (morton_mode != PS_HIP_MORTON_TRUNC && morton_mode != PS_HIP_MORTON_AVX2)
        return PS_HIP_EINVAL;
    int dev = 0;
    PS_CK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_prep_mu);
    hipError_t err = hipSuccess;
    PrepWS* w = prep_ws(dev, &err);
    if (!w) return map_err(err);
    const float* in[3] = {x, y, z};
    DevBuf* stg[3] = {&w->x, &w->y, &w->z};
    const float* d[3];
    for (int a = 0; a < 3; ++a) {
        if (is_device_ptr(in[a])) {
            d[a] = in[a];
        } else {
            PS_CK(grow(*stg[a], n * 4, w->s));
            PS_CK(hipMemcpyAsync(stg[a]->p, in[a], n * 4, hipMemcpyHostToDevice, w->s));
            d[a] = reinterpret_cast<const float*>(stg[a]->p);
        }
    }
    uint32_t* dcode = codes;
    const bool host_codes = !is_device_ptr(codes);
    if (host_codes) {
        PS_CK(grow(w->code, n * 4, w->s));
        dcode = reinterpret_cast<uint32_t*>(w->code.p);
    }
    float* prm = reinterpret_cast<float*>(w->prm.p);
    if (!ps_launch_prepare_small(d[0], d[1], d[2], n, prm, dcode, morton_mode, w->s)) {
        ps_launch_bbox(d[0], d[1], d[2], n, reinterpret_cast<float2*>(w->part.p), prm, w->s);
        ps_launch_morton(d[0], d[1], d[2], n, prm, dcode, morton_mode, w->s);
    }
    PS_CK(hipMemcpyAsync(w->h_out, prm, 16, hipMemcpyDeviceToHost, w->s));
    if (host_codes) PS_CK(hipMemcpyAsync(codes, dcode, n * 4, hipMemcpyDeviceToHost, w->s));
    PS_CK(wait_stream(w->s));
    if (out_min) *out_min = w->h_out[0];
    if (out_max) *out_max = w->h_out[1];
    if (out_range) *out_range = w->h_out[2];
    return PS_HIP_OK;
}

int ps_hip_prepare_noalloc(const float* x, const float* y, const float* z, size_t n,
                           uint32_t* codes, float* out_min, float* out_max, float* out_range,
                           int morton_mode) {
    if (!x || !y || !z || !codes || n == 0) return PS_HIP_EINVAL;  // H:1914
    if (n > PS_HIP_PREPARE_NOALLOC_MAX) return PS_HIP_EINVAL;
    if (morton_mode != PS_HIP_MORTON_TRUNC && morton_mode != PS_HIP_MORTON_AVX2)
        return PS_HIP_EINVAL;
    // no staging: host arrays would need device buffers
    if (!is_device_ptr(x) || !is_device_ptr(y) || !is_device_ptr(z) || !is_device_ptr(codes))
        return PS_HIP_EINVAL;
    float out3[3] = {0.f, 0.f, 0.f};
    PS_CK(ps_prepare_noalloc(x, y, z, n, codes, morton_mode, out3, nullptr));
    if (out_min) *out_min = out3[0];
    if (out_max) *out_max = out3[1];
    if (out_range) *out_range = out3[2];
    return PS_HIP_OK;
}

static int bind_all(ps_hip_engine_t* e, float* x, float* y, float* z, float* mass, float* fx,
                    float* fy, float* fz, uint32_t* id, size_t n, uint32_t* codes, bool upload) {
    PS_RC(bind_array(e, 0, x, n, upload, &e->x));
    PS_RC(bind_array(e, 1, y, n, upload, &e->y));
    PS_RC(bind_array(e, 2, z, n, upload, &e->z));
    PS_RC(bind_array(e, 3, mass, n, upload, &e->m));
    PS_RC(bind_array(e, 4, fx, n, false, &e->fx));
    PS_RC(bind_array(e, 5, fy, n, false, &e->fy));
    PS_RC(bind_array(e, 6, fz, n, false, &e->fz));
    PS_RC(bind_array(e, 7, id, n, upload, &e->id));
    PS_RC(bind_array(e, 8, codes, n, upload, &e->code));
    return PS_HIP_OK;
}

int ps_hip_calc_forces(ps_hip_engine_t* e, float* x, float* y, float* z, float* mass, float* fx,
                       float* fy, float* fz, uint32_t* id, size_t n, uint32_t* codes, float cx,
                       float cy, float cz, float hw) {
    if (!e || e->pending) return PS_HIP_EINVAL;
    if (n == 0) return PS_HIP_OK;  // nothing to sort or build (H:1592)
    if (n > e->max_p) return PS_HIP_EINVAL;
    std::memset(&e->st, 0, sizeof(e->st));
    e->record(0);
    e->n = n;
    e->root = make_float4(cx, cy, cz, hw);
    PS_RC(bind_all(e, x, y, z, mass, fx, fy, fz, id, n, codes, true));
    e->bound = true;
    int rc = run_tick(e);
    void* host[9] = {x, y, z, mass, fx, fy, fz, id, codes};
    void* dev[9] = {e->x, e->y, e->z, e->m, e->fx, e->fy, e->fz, e->id, e->code};
    for (int attempt = 0;; ++attempt) {
        if (rc != PS_HIP_OK) {
            (void)hipStreamSynchronize(e->s);
            return rc;
        }
        // copy staged arrays back: the sorted payload and the forces
        for (int k = 0; k < 9; ++k)
            if (host[k] != dev[k])
                PS_CK(hipMemcpyAsync(host[k], dev[k], n * 4, hipMemcpyDeviceToHost, e->s));
        if (e->count_on)
            PS_CK(hipMemcpyAsync(e->h_small + 32, e->pairs(), 16, hipMemcpyDeviceToHost, e->s));
        e->record(9);
        PS_CK(wait_stream(e->s));
        bool redo = false;
        PS_RC(trav_finish(e, attempt, &redo));
        if (!redo) break;
        rc = run_from_traverse(e);
    }
    e->st.n_particles = n;
    e->st.n_nodes = e->n_nodes;
    e->st.max_depth = e->max_depth;
    std::memcpy(e->st.level_start, e->ls, sizeof(e->ls));
    if (e->count_on) {
        e->st.p2p_pairs = *reinterpret_cast<unsigned long long*>(e->h_small + 32);
        e->st.p2p_leaf_pairs = *reinterpret_cast<unsigned long long*>(e->h_small + 34);
    }
    if (e->ev_mask) {
        e->st.ms_sort = e->elapsed(1, 2);
        e->st.ms_build = e->elapsed(2, 3);
        e->st.ms_upward = e->elapsed(3, 4);
        e->st.ms_traverse = e->elapsed(4, 5);
        e->st.ms_l2l = e->elapsed(5, 6);
        e->st.ms_p2p = e->elapsed(6, 7);
        e->st.ms_mlp = e->elapsed(7, 8);
        e->st.ms_download = e->elapsed(8, 9);
        e->st.ms_upload = e->elapsed(0, 1);
        e->st.ms_total = e->elapsed(0, 9);
    }
    return PS_HIP_OK;
}

int ps_hip_calc_forces_async(ps_hip_engine_t* e, float* x, float* y, float* z, float* mass,
                             float* fx, float* fy, float* fz, uint32_t* id, size_t n,
                             uint32_t* codes, float cx, float cy, float cz, float hw,
                             void* stream) {
    if (!e || e->pending) return PS_HIP_EINVAL;
    if (n == 0) return PS_HIP_OK;  // nothing to sort or build (H:1592)
    if (n > e->max_p) return PS_HIP_EINVAL;
    void* arrs[9] = {x, y, z, mass, fx, fy, fz, id, codes};
    for (void* p : arrs)
        if (!is_device_ptr(p)) return PS_HIP_EINVAL;  // host arrays would need a wait
    PS_RC(async_ensure(e, n));
    std::memset(&e->st, 0, sizeof(e->st));
    e->n = n;
    e->root = make_float4(cx, cy, cz, hw);
    e->x = x;
    e->y = y;
    e->z = z;
    e->m = mass;
    e->fx = fx;
    e->fy = fy;
    e->fz = fz;
    e->id = id;
    e->code = codes;
    e->bound = true;
    e->built = false;
    e->l2l_pending = false;
    hipStream_t cs = reinterpret_cast<hipStream_t>(stream);
    int rc = async_front(e, cs);
    if (rc == PS_HIP_OK) rc = async_back(e, cs);
    if (rc != PS_HIP_OK) {
        (void)hipStreamSynchronize(cs);
        return rc;
    }
    e->pending = true;
    e->async_s = cs;
    return PS_HIP_OK;
}

int ps_hip_calc_forces_wait(ps_hip_engine_t* e) {
    if (!e) return PS_HIP_EINVAL;
    if (!e->pending) return PS_HIP_OK;
    e->pending = false;
    PS_CK(wait_event(e->ev_async));
    const uint32_t flags = e->h_small[12];
    const unsigned long long total =
        *reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(e->h_small) + 64);
    if (flags & 1u) return PS_HIP_EINVAL;  // codes must be 30-bit Morton codes
    if (total > e->node_cap) return PS_HIP_EOOM;  // the caller's node capacity (H:1742-1743)
    for (int attempt = 0;; ++attempt) {
        const PsTravState hst = *e->h_tstate;
        if (!hst.overflow) break;
        // an arena was too small: grow past the need and run the tick again
        // from the traversal (the sorted arrays and the tree stand)
        if (attempt >= 4 || hst.arena_ctr > 0xFFFFFFFFull || hst.mask_ctr > 0xFFFFFFFFull)
            return PS_HIP_EOOM;
        PS_CK(grow(e->arena_b, std::max<size_t>(2 * e->arena_b.bytes, 5 * hst.arena_ctr), e->s));
        PS_CK(grow(e->masks_b, std::max<size_t>(2 * e->masks_b.bytes, 10 * hst.mask_ctr), e->s));
        ps_fill2_u32_count(reinterpret_cast<uint32_t*>(e->loc()), 4, e->p2p_len(), 1, 0u,
                           e->totals(), e->node_alloc, e->async_s);
        PS_RC(async_back(e, e->async_s));
        PS_CK(wait_event(e->ev_async));
    }
    std::memcpy(e->ls, e->h_small, sizeof(e->ls));
    e->n_nodes = (uint32_t)total;
    e->max_depth = 0;
    for (int d = 0; d <= 10; ++d)
        if (e->ls[d + 1] > e->ls[d]) e->max_depth = (uint32_t)d;
    e->built = true;
    e->st.n_particles = e->n;
    e->st.n_nodes = e->n_nodes;
    e->st.max_depth = e->max_depth;
    std::memcpy(e->st.level_start, e->ls, sizeof(e->ls));
    e->st.m2l_pairs = e->h_tstate->m2l_pairs;
    e->st.list_entries = e->h_tstate->arena_ctr - 1;
    e->st.iterations = e->max_depth + 1;
    return PS_HIP_OK;
}

int ps_hip_engine_export_tree(ps_hip_engine_t* e, size_t cap, float* geom4, float* mp4,
                              float* loc4, uint32_t* link2, uint32_t* nend, uint32_t* parent,
                              uint32_t* n_nodes) {
    if (!e || !e->built) return PS_HIP_EINVAL;
    if (n_nodes) *n_nodes = e->n_nodes;
    const size_t M = e->n_nodes;
    if (cap < M) return PS_HIP_EOOM;
    if (loc4) PS_RC(materialize_locals(e));
    PS_CK(wait_stream(e->s));
    if (geom4) PS_CK(hipMemcpy(geom4, e->geom(), 16 * M, hipMemcpyDeviceToHost));
    if (mp4) PS_CK(hipMemcpy(mp4, e->mp(), 16 * M, hipMemcpyDeviceToHost));
    if (loc4) PS_CK(hipMemcpy(loc4, e->loc(), 16 * M, hipMemcpyDeviceToHost));
    if (link2) PS_CK(hipMemcpy(link2, e->link(), 8 * M, hipMemcpyDeviceToHost));
    if (nend) PS_CK(hipMemcpy(nend, e->nend(), 4 * M, hipMemcpyDeviceToHost));
    if (parent) PS_CK(hipMemcpy(parent, e->parent(), 4 * M, hipMemcpyDeviceToHost));
    return PS_HIP_OK;
}

int ps_hip_debug_bind(ps_hip_engine_t* e, float* x, float* y, float* z, float* mass, float* fx,
                      float* fy, float* fz, uint32_t* id, size_t n, uint32_t* codes, float cx,
                      float cy, float cz, float hw) {
    if (!e || n == 0 || n > e->max_p) return PS_HIP_EINVAL;
    void* arrs[9] = {x, y, z, mass, fx, fy, fz, id, codes};
    for (void* p : arrs)
        if (!is_device_ptr(p)) return PS_HIP_EINVAL;
    e->n = n;
    e->root = make_float4(cx, cy, cz, hw);
    PS_RC(bind_all(e, x, y, z, mass, fx, fy, fz, id, n, codes, false));
    // pos4 for stages that skip the sort
    ps_launch_pack_pos4(e->x, e->y, e->z, e->m, n, e->pos4(), e->s);
    e->bound = true;
    e->built = false;
    return map_err(hipGetLastError());
}

int ps_hip_debug_stage(ps_hip_engine_t* e, int stage) {
    if (!e || !e->bound) return PS_HIP_EINVAL;
    if (stage > kBuild && !e->built) return PS_HIP_EINVAL;
    int rc;
    switch (stage) {
        case kSort: rc = stage_sort(e); break;
        case kBuild: rc = stage_build(e); break;
        case kUpward: rc = stage_upward(e); break;
        case kInteract:
            for (int attempt = 0;; ++attempt) {
                rc = stage_traverse(e);
                if (rc != PS_HIP_OK) break;
                PS_CK(join_side(e));
                PS_CK(wait_stream(e->s));
                bool redo = false;
                rc = trav_finish(e, attempt, &redo);
                if (rc != PS_HIP_OK || !redo) break;
            }
            break;
        case kDownward:
            rc = stage_l2l(e);
            if (rc == PS_HIP_OK) rc = stage_p2p(e, 1);
            break;
        case kP2P: rc = stage_p2p(e, 2); break;
        case kTickTail:
            for (int attempt = 0;; ++attempt) {
                rc = run_from_traverse(e);
                if (rc != PS_HIP_OK) break;
                PS_CK(wait_stream(e->s));
                bool redo = false;
                rc = trav_finish(e, attempt, &redo);
                if (rc != PS_HIP_OK || !redo) break;
            }
            break;
        default: return PS_HIP_EINVAL;
    }
    const int rj = finish_copy(e);
    if (rc != PS_HIP_OK) return rc;
    if (rj != PS_HIP_OK) return rj;
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_debug_set_local(ps_hip_engine_t* e, uint32_t node, const float* field3) {
    if (!e || !e->built || !field3 || node >= e->n_nodes) return PS_HIP_EINVAL;
    float v[4] = {0.f, field3[0], field3[1], field3[2]};
    PS_CK(hipMemcpyAsync(e->loc() + node, v, 16, hipMemcpyHostToDevice, e->s));
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_op_m2l(const float* tgeom4, const float* sgeom4, const float* smp4, float theta,
                  float* out4) {
    if (!tgeom4 || !sgeom4 || !smp4 || !out4) return PS_HIP_EINVAL;
    float* d = nullptr;
    PS_CK(dev_malloc(reinterpret_cast<void**>(&d), 16));
    hipError_t r = ps_op_m2l(tgeom4, sgeom4, smp4, theta, d, nullptr);
    if (r == hipSuccess) r = hipMemcpy(out4, d, 16, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return map_err(r);
}

int ps_hip_op_morton_exhaustive(uint64_t* out2) {
    if (!out2) return PS_HIP_EINVAL;
    unsigned long long* d = nullptr;
    PS_CK(dev_malloc(reinterpret_cast<void**>(&d), 16));
    hipError_t r = ps_op_morton_exhaustive(d, nullptr);
    if (r == hipSuccess) r = hipMemcpy(out2, d, 16, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return map_err(r);
}

int ps_hip_op_copy(const void* src, void* dst, size_t bytes) {
    if (!src || !dst || (bytes & 15u) || !is_device_ptr(src) || !is_device_ptr(dst))
        return PS_HIP_EINVAL;
    PS_CK(ps_op_copy(src, dst, bytes, nullptr));
    PS_CK(hipDeviceSynchronize());
    return PS_HIP_OK;
}

int ps_hip_op_bandwidth(int kind, const void* src, void* dst, const uint32_t* idx, size_t n) {
    if (kind < 0 || kind > 3 || !dst || !is_device_ptr(dst)) return PS_HIP_EINVAL;
    if (kind != 2 && (!src || !is_device_ptr(src))) return PS_HIP_EINVAL;
    if (kind == 3 && (!idx || !is_device_ptr(idx))) return PS_HIP_EINVAL;
    PS_CK(ps_op_bandwidth(kind, src, dst, idx, n, nullptr));
    PS_CK(hipDeviceSynchronize());
    return PS_HIP_OK;
}

int ps_hip_direct_f64(const float* x, const float* y, const float* z, const float* mass,
                      size_t n, const uint32_t* targets, size_t nt, double* out3) {
    if (!x || !y || !z || !mass || !targets || !out3 || n == 0) return PS_HIP_EINVAL;
    if (!is_device_ptr(x) || !is_device_ptr(y) || !is_device_ptr(z) || !is_device_ptr(mass))
        return PS_HIP_EINVAL;
    for (size_t k = 0; k < nt; ++k)
        if (targets[k] >= n) return PS_HIP_EINVAL;
    if (nt == 0) return PS_HIP_OK;
    float4* pos4 = nullptr;
    uint32_t* dt = nullptr;
    double* dout = nullptr;
    hipError_t r = dev_malloc(reinterpret_cast<void**>(&pos4), 16 * n);
    if (r == hipSuccess) r = dev_malloc(reinterpret_cast<void**>(&dt), 4 * nt);
    if (r == hipSuccess) r = dev_malloc(reinterpret_cast<void**>(&dout), 24 * nt);
    if (r == hipSuccess) r = hipMemcpy(dt, targets, 4 * nt, hipMemcpyHostToDevice);
    if (r == hipSuccess) {
        ps_launch_pack_pos4(x, y, z, mass, n, pos4, nullptr);
        ps_launch_direct_f64(pos4, n, dt, nt, dout, nullptr);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipMemcpy(out3, dout, 24 * nt, hipMemcpyDeviceToHost);
    if (pos4) (void)hipFree(pos4);
    if (dt) (void)hipFree(dt);
    if (dout) (void)hipFree(dout);
    return map_err(r);
}

int ps_hip_step_reset_ids(ps_hip_engine_t* e, uint32_t* id, size_t n) {
    if (!e || !id || !is_device_ptr(id)) return PS_HIP_EINVAL;
    ps_launch_iota(id, n, e->s);
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_step_gather(ps_hip_engine_t* e, const uint32_t* id, float* vx, float* vy, float* vz,
                       float* tmp3, size_t n) {
    if (!e || !id || !vx || !vy || !vz || !tmp3) return PS_HIP_EINVAL;
    ps_launch_gather_vel(id, vx, vy, vz, tmp3, tmp3 + n, tmp3 + 2 * n, n, e->s);
    PS_CK(hipMemcpyAsync(vx, tmp3, 4 * n, hipMemcpyDeviceToDevice, e->s));
    PS_CK(hipMemcpyAsync(vy, tmp3 + n, 4 * n, hipMemcpyDeviceToDevice, e->s));
    PS_CK(hipMemcpyAsync(vz, tmp3 + 2 * n, 4 * n, hipMemcpyDeviceToDevice, e->s));
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_step_euler(ps_hip_engine_t* e, float* x, float* y, float* z, float* vx, float* vy,
                      float* vz, const float* fx, const float* fy, const float* fz,
                      const float* mass, float dt, size_t n) {
    if (!e || !x || !y || !z || !vx || !vy || !vz || !fx || !fy || !fz || !mass)
        return PS_HIP_EINVAL;
    ps_launch_euler(x, y, z, vx, vy, vz, fx, fy, fz, mass, dt, n, e->s);
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

int ps_hip_step_frame(ps_hip_engine_t* e, uint32_t* id, float* vx, float* vy, float* vz,
                      float* tmp3, float* x, float* y, float* z, const float* fx,
                      const float* fy, const float* fz, const float* mass, float dt, size_t n) {
    if (!e || !id || !vx || !vy || !vz || !tmp3 || !x || !y || !z || !fx || !fy || !fz || !mass)
        return PS_HIP_EINVAL;
    ps_launch_frame(id, vx, vy, vz, tmp3, x, y, z, fx, fy, fz, mass, dt, n, e->s);
    PS_CK(hipGetLastError());
    PS_CK(wait_stream(e->s));
    return PS_HIP_OK;
}

void* ps_hip_malloc(size_t bytes) {
    // the caller's own memory: not counted among the library's allocations
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

int ps_hip_free(void* p) {
    if (!p) return PS_HIP_OK;
    return map_err(hipFree(p));
}

int ps_hip_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    if (!dst || !src) return PS_HIP_EINVAL;
    hipMemcpyKind k = kind == 1   ? hipMemcpyHostToDevice
                      : kind == 2 ? hipMemcpyDeviceToHost
                      : kind == 3 ? hipMemcpyDeviceToDevice
                                  : hipMemcpyDefault;
    return map_err(hipMemcpy(dst, src, bytes, k));
}

int ps_hip_synchronize(void) { return map_err(hipDeviceSynchronize()); }

int ps_hip_stream_create(void** out) {
    if (!out) return PS_HIP_EINVAL;
    hipStream_t s = nullptr;
    PS_CK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = s;
    return PS_HIP_OK;
}

int ps_hip_stream_destroy(void* stream) {
    if (!stream) return PS_HIP_EINVAL;
    return map_err(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)));
}

int ps_hip_stream_synchronize(void* stream) {
    return map_err(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
}

int ps_hip_device_name(char* buf, size_t len) {
    if (!buf || len == 0) return PS_HIP_EINVAL;
    int dev = 0;
    PS_CK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PS_CK(hipGetDeviceProperties(&p, dev));
    std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return PS_HIP_OK;
}

}  // extern "C"
