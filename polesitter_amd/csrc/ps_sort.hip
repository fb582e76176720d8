// ps_sort.hip -- stable LSD radix sort of the particles by Morton code
// (replaces ps_impl_sort_particles, H:1588-1693).
//
// Keys are 30-bit Morton codes, so three passes of 10-bit digits cover them
// (the reference uses four 8-bit passes over 32 bits; any stable sort yields
// the same permutation).  Each pass:
//   upsweep   per-tile digit histogram in LDS -> table[digit][tile]
//   row scan  one block per digit: exclusive scan of its row (where each
//             tile's run starts inside the digit's output) and the digit's
//             total; each downsweep block adds the exclusive scan of the
//             1024 totals (where the digit's output starts) -- bucket-major,
//             tile-minor, the same offset order as H:1661-1667
//   downsweep the tile is ranked stably inside LDS (each wavefront ranks its
//             own sixteenth in order: wave-wide match of equal digits with 10
//             ballots and the wave's running digit counts, then one scan of
//             the counts over the waves -- three barriers per tile), then
//             written out digit run by digit run, so consecutive lanes store
//             to consecutive addresses; the payload (id, x, y, z, mass) is
//             gathered from the tile's own input range, which is L2-resident.
// Tiles are 4096 keys (upsweep: 256 threads x 16; downsweep: 1024 threads
// x 4, sixteen wavefronts of 256 keys each ranked in four rounds).
// Ping-pong: io -> t0 -> t1 -> io, so the sorted data ends in the caller's
// arrays (H:1688-1690).  Many systems sorted at once (contiguous segments,
// one per system) are sorted segment by segment in the same three passes:
// the tiles are cut at the segments' ends, and a tile's digit runs go to
// its segment's start + the digit's start inside the segment + the tile's
// offset in it, so no key leaves its segment (ps_radix_sort_seg).  With the
// system id as a second key instead, a fourth, stable pass by it groups the
// systems again.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kRadixBits = 10;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kBlock = 256;   // upsweep
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
constexpr int kDBlock = 1024;  // downsweep: sixteen wavefronts rank one tile
constexpr int kDItems = kTile / kDBlock;
constexpr int kWaves = kDBlock / 64;
constexpr int kDigitsPer = kRadix / kDBlock;  // digits of the tile histogram per thread
static_assert(kDigitsPer >= 1 && kRadix % kDBlock == 0, "digits per thread");
static_assert(kRadix == kDBlock, "the downsweep scans the digit totals one per thread");
constexpr int kRowBlock = 256;  // row scan

// Tile t of the keys: [vstart[t], vstart[t + 1]) when the tiles are cut at
// segment ends (vstart set), else [t kTile, min(n, (t + 1) kTile)).
__device__ __forceinline__ void tile_range(const uint32_t* vstart, size_t n, uint32_t t,
                                           size_t& base, uint32_t& cnt) {
    if (vstart) {
        base = vstart[t];
        cnt = vstart[t + 1] - vstart[t];
    } else {
        base = (size_t)t * kTile;
        cnt = (uint32_t)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    }
}

__global__ __launch_bounds__(kBlock) void sort_upsweep_kernel(const uint32_t* __restrict__ key,
                                                              size_t n, int shift,
                                                              uint32_t* __restrict__ table,
                                                              uint32_t n_tiles,
                                                              const uint32_t* __restrict__ vstart) {
    __shared__ uint32_t h[kRadix];
    for (int d = threadIdx.x; d < kRadix; d += kBlock) h[d] = 0;
    __syncthreads();
    size_t base;
    uint32_t cnt;
    tile_range(vstart, n, blockIdx.x, base, cnt);
#pragma unroll 4
    for (int r = 0; r < kItems; ++r) {
        const uint32_t j = (uint32_t)(r * kBlock) + threadIdx.x;
        if (j < cnt) atomicAdd(&h[(key[base + j] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kBlock)
        table[(size_t)d * n_tiles + blockIdx.x] = h[d];
}

// Row d of the digit-major table (the digit's count in every tile) becomes
// its exclusive scan, in place; dtot[d] gets the row's total.  Thread k sums
// a contiguous run of the row, the block scans the sums, and the run is
// rewritten from there.
__global__ __launch_bounds__(kRowBlock) void sort_rowscan_kernel(uint32_t* __restrict__ table,
                                                                 uint32_t n_tiles,
                                                                 uint32_t* __restrict__ dtot) {
    __shared__ uint32_t wsum[kRowBlock / 64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t* row = table + (size_t)blockIdx.x * n_tiles;
    const uint32_t per = (n_tiles + kRowBlock - 1) / kRowBlock;
    const uint32_t lo = threadIdx.x * per < n_tiles ? threadIdx.x * per : n_tiles;
    const uint32_t hi = lo + per < n_tiles ? lo + per : n_tiles;
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t t = lo; t < hi; ++t) sum += row[t];
    uint32_t wtot;
    uint32_t ex = ps_wave_excl_scan(sum, &wtot);
    if (lane == 0) wsum[wave] = wtot;
    __syncthreads();
    uint32_t all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kRowBlock / 64; ++w) {
        if (w < wave) ex += wsum[w];
        all += wsum[w];
    }
#pragma unroll 8
    for (uint32_t t = lo; t < hi; ++t) {
        const uint32_t c = row[t];
        row[t] = ex;
        ex += c;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = all;
}

// Segment s's digit starts, one digit per thread: its tiles are
// [seg_tile[s], seg_tile[s + 1]), so digit d has P(b, d) - P(a, d) keys in it
// (P: the scanned table, P(n_tiles, d) = dtot[d]); the digit's run starts at
// off[s] + (the segment's keys of smaller digits).  sbase[s][d] is that start
// minus P(a, d), so a tile of the segment adds its own P(t, d).
__global__ __launch_bounds__(kRadix) void sort_segbase_kernel(
    const uint32_t* __restrict__ table, uint32_t n_tiles, const uint32_t* __restrict__ dtot,
    const uint32_t* __restrict__ seg_tile, const uint32_t* __restrict__ off,
    uint32_t* __restrict__ sbase) {
    __shared__ uint32_t wsum[kRadix / 64];
    const uint32_t sg = blockIdx.x, d = threadIdx.x;
    const uint32_t lane = d & 63u, wave = d >> 6;
    const uint32_t a = seg_tile[sg], b = seg_tile[sg + 1];
    const uint32_t pa = a < n_tiles ? table[(size_t)d * n_tiles + a] : dtot[d];
    const uint32_t pb = b < n_tiles ? table[(size_t)d * n_tiles + b] : dtot[d];
    uint32_t wt;
    uint32_t ex = ps_wave_excl_scan(pb - pa, &wt);
    if (lane == 0) wsum[wave] = wt;
    __syncthreads();
    for (uint32_t w = 0; w < wave; ++w) ex += wsum[w];
    sbase[(size_t)sg * kRadix + d] = off[sg] + ex - pa;
}

__global__ __launch_bounds__(kDBlock) void sort_downsweep_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ iin,
    const float* __restrict__ xin, const float* __restrict__ yin, const float* __restrict__ zin,
    const float* __restrict__ min_, uint32_t* __restrict__ kout, uint32_t* __restrict__ iout,
    float* __restrict__ xout, float* __restrict__ yout, float* __restrict__ zout,
    float* __restrict__ mout, float4* __restrict__ pos4, const uint32_t* __restrict__ ain,
    uint32_t* __restrict__ aout, size_t n, int shift, const uint32_t* __restrict__ offs,
    const uint32_t* __restrict__ dtot, uint32_t n_tiles, uint32_t* __restrict__ kcopy,
    const uint32_t* __restrict__ vstart, const uint32_t* __restrict__ vseg,
    const uint32_t* __restrict__ sbase) {
    __shared__ uint32_t keys[kTile];
    __shared__ uint16_t sidx[kTile];
    __shared__ uint32_t tstart[kRadix];
    __shared__ uint32_t goff[kRadix];
    __shared__ uint16_t wcnt[kWaves][kRadix];
    __shared__ uint32_t wsum[kWaves];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    size_t base;
    uint32_t valid_n;
    tile_range(vstart, n, blockIdx.x, base, valid_n);

    // where digit tid's run of this tile goes: the digit's start (exclusive
    // scan of the totals, one digit per thread; or the tile's segment's
    // start for the digit) + the tile's offset in it
    {
        const uint32_t d = (uint32_t)tid;
        const uint32_t own = offs[(size_t)d * n_tiles + blockIdx.x];
        tstart[d] = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcnt[w][d] = 0;
        if (sbase) {
            goff[d] = sbase[(size_t)vseg[blockIdx.x] * kRadix + d] + own;
        } else {
            const uint32_t c = dtot[d];
            uint32_t wt;
            uint32_t ex = ps_wave_excl_scan(c, &wt);
            if (lane == 0) wsum[wave] = wt;
            __syncthreads();
            for (int w = 0; w < wave; ++w) ex += wsum[w];
            goff[d] = ex + own;
        }
    }
    __syncthreads();

    // 1. load the tile, count digits
#pragma unroll
    for (int r = 0; r < kDItems; ++r) {
        uint32_t idx = (uint32_t)(r * kDBlock + tid);
        uint32_t k = 0xFFFFFFFFu;
        if (idx < valid_n) {
            k = kin[base + idx];
            atomicAdd(&tstart[(k >> shift) & (kRadix - 1)], 1u);
        }
        keys[idx] = k;
    }
    __syncthreads();

    // 2. exclusive scan of the tile histogram (kDigitsPer digits per thread)
    {
        uint32_t a[kDigitsPer];
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kDigitsPer; ++q) {
            a[q] = tstart[kDigitsPer * tid + q];
            s += a[q];
        }
        uint32_t inc = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t ex = inc - s;
        for (int w = 0; w < wave; ++w) ex += wsum[w];
#pragma unroll
        for (int q = 0; q < kDigitsPer; ++q) {
            tstart[kDigitsPer * tid + q] = ex;
            ex += a[q];
        }
    }
    __syncthreads();

    // 3. stable ranking.  Wave w ranks the tile's keys [w Q, (w+1) Q),
    //    Q = kTile / kWaves, in order (rounds of 64): a key's rank among the
    //    equal digits of its wave so far comes from the round's digit match
    //    and the wave's own running digit counts in wcnt[w] (no barriers);
    //    then one exclusive scan over the waves per digit, and the key's
    //    tile position is tstart[d] + (its digit's count in earlier waves) +
    //    that rank -- the stable order, since the waves hold consecutive
    //    ranges.
    constexpr int kQ = kTile / kWaves;
    constexpr int kRounds = kQ / 64;
    const uint64_t lt = ps_lanemask_lt();
    uint32_t dr[kRounds];  // per round: digit << 16 | rank inside the wave
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t idx = (uint32_t)(wave * kQ + r * 64 + lane);
        const bool valid = idx < valid_n;
        const uint32_t d = (keys[idx] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(valid && bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const uint32_t cnt = (uint32_t)__popcll(peers);
        // every lane of digit d reads the wave's count before its leader
        // (the lowest lane of the digit) adds this round's keys
        const uint32_t before = valid ? (uint32_t)wcnt[wave][d] : 0u;
        if (valid && rank == 0) wcnt[wave][d] = (uint16_t)(before + cnt);
        dr[r] = valid ? (d << 16) | (before + rank) : 0xFFFFFFFFu;
    }
    __syncthreads();
    // exclusive scan over the waves, per digit, in place
    for (int d = tid; d < kRadix; d += kDBlock) {
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = (uint16_t)acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        if (dr[r] == 0xFFFFFFFFu) continue;
        const uint32_t d = dr[r] >> 16;
        const uint32_t pos = tstart[d] + (uint32_t)wcnt[wave][d] + (dr[r] & 0xFFFFu);
        sidx[pos] = (uint16_t)(wave * kQ + r * 64 + lane);
    }
    __syncthreads();

    // 4. write out digit runs: position p of the tile goes to
    //    goff[d] + (p - tstart[d])
#pragma unroll
    for (int r = 0; r < kDItems; ++r) {
        uint32_t p = (uint32_t)(r * kDBlock + tid);
        if (p >= valid_n) break;
        uint32_t sidx_p = sidx[p];
        uint32_t k = keys[sidx_p];
        uint32_t d = (k >> shift) & (kRadix - 1);
        size_t g = (size_t)goff[d] + (p - tstart[d]);
        size_t src = base + sidx_p;
        kout[g] = k;
        iout[g] = iin[src];
        if (kcopy) kcopy[g] = k;
        if (xin) {  // null: keys and index only
            float vx = xin[src], vy = yin[src], vz = zin[src], vm = min_[src];
            xout[g] = vx;
            yout[g] = vy;
            zout[g] = vz;
            mout[g] = vm;
            if (pos4) pos4[g] = make_float4(vx, vy, vz, vm);
        }
        if (aout) aout[g] = ain[src];
    }
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// ---- incremental sort: the previous tick's sorted order, a few keys changed ----
//
// prev = the keys the last sort produced, position by position (sorted).  A
// particle whose key (low 30 bits, the bits the radix passes sort) still
// equals prev at its position is unchanged; the unchanged ones, in position
// order, are sorted by (key, position) already.  The changed ones are
// compacted and sorted stably by key, i.e. by (key, position).  Merging the
// two by (key, position) gives the stable order of all -- the permutation
// the full radix sort gives, for any input, as long as prev is sorted.
// Output position of an unchanged particle i: its rank among the unchanged
// + the number of changed ones before (key_i, i); of the changed one of
// rank r: r + the number of unchanged ones before (key, position), found by
// two binary searches in prev (the unchanged ones are prev's entries minus
// the changed positions).
constexpr uint32_t kKeyMask = 0x3FFFFFFFu;
constexpr int kMBlock = 1024;   // merge: positions per block
constexpr int kSlice = 4096;    // changed keys of a block's key range held in LDS

__device__ __forceinline__ unsigned long long packed_key(uint32_t key, uint32_t pos) {
    return ((unsigned long long)(key & kKeyMask) << 32) | pos;
}

// first index in a[0, n) whose value is >= v (a sorted)
__device__ __forceinline__ uint32_t lower_bound_u64(const unsigned long long* a, uint32_t n,
                                                    unsigned long long v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// both bounds of key k in prev[0, n): a = first index with key >= k, b = first
// with key > k; the two binary searches step together so their loads overlap
__device__ __forceinline__ void bound_keys(const uint32_t* prev, size_t n, uint32_t k, size_t& a,
                                           size_t& b) {
    size_t lo0 = 0, hi0 = n, lo1 = 0, hi1 = n;
    while (lo0 < hi0 || lo1 < hi1) {
        const size_t m0 = lo0 + ((hi0 - lo0) >> 1), m1 = lo1 + ((hi1 - lo1) >> 1);
        const bool s0 = lo0 < hi0, s1 = lo1 < hi1;
        const uint32_t v0 = s0 ? prev[m0] & kKeyMask : 0u;
        const uint32_t v1 = s1 ? prev[m1] & kKeyMask : 0u;
        if (s0) {
            if (v0 < k)
                lo0 = m0 + 1;
            else
                hi0 = m0;
        }
        if (s1) {
            if (v1 <= k)
                lo1 = m1 + 1;
            else
                hi1 = m1;
        }
    }
    a = lo0;
    b = lo1;
}

// Changed flags: one 64-bit word per 64 positions (mask[w]); the positions
// come in blocks of kMBlock (16 words), and bcnt[b] is block b's count.
constexpr int kWordsPerBlock = kMBlock / 64;
constexpr uint32_t kSmallSort = 4096;  // changed keys sorted in one block's LDS

__global__ __launch_bounds__(kMBlock) void delta_mask_kernel(const uint32_t* __restrict__ key,
                                                             const uint32_t* __restrict__ prev,
                                                             size_t n,
                                                             unsigned long long* __restrict__ mask,
                                                             uint32_t* __restrict__ bcnt) {
    __shared__ uint32_t wc[kWordsPerBlock];
    const size_t i = (size_t)blockIdx.x * kMBlock + threadIdx.x;
    const bool ch = i < n && ((key[i] ^ prev[i]) & kKeyMask) != 0u;
    const unsigned long long b = __ballot(ch);
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wc[wave] = (uint32_t)__popcll(b);
        mask[i >> 6] = b;  // words past n read as zero
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < kWordsPerBlock; ++w) t += wc[w];
        bcnt[blockIdx.x] = t;
    }
}

// boff[b] = changed positions before block b (b = 0..nb), *count = all of
// them; one block
__global__ __launch_bounds__(1024) void delta_scan_kernel(const uint32_t* __restrict__ bcnt,
                                                          uint32_t nb, uint32_t* __restrict__ boff,
                                                          unsigned long long* __restrict__ count) {
    __shared__ uint32_t wsum[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = tid * per < nb ? tid * per : nb;
    const uint32_t hi = lo + per < nb ? lo + per : nb;
    uint32_t sum = 0;
    for (uint32_t t = lo; t < hi; ++t) sum += bcnt[t];
    uint32_t wt;
    uint32_t ex = ps_wave_excl_scan(sum, &wt);
    if (lane == 0) wsum[wave] = wt;
    __syncthreads();
    uint32_t all = 0;
#pragma unroll
    for (uint32_t w = 0; w < 16; ++w) {
        if (w < wave) ex += wsum[w];
        all += wsum[w];
    }
    for (uint32_t t = lo; t < hi; ++t) {
        boff[t] = ex;
        ex += bcnt[t];
    }
    if (tid == 0) {
        boff[nb] = all;
        *count = all;
    }
}

// number of changed positions before p (p <= n)
__device__ __forceinline__ uint32_t changed_before(const unsigned long long* mask,
                                                   const uint32_t* boff, size_t p) {
    const size_t b = p / kMBlock;
    const size_t w0 = b * kWordsPerBlock, w = p >> 6;
    uint32_t c = boff[b];
    for (size_t q = w0; q < w; ++q) c += (uint32_t)__popcll(mask[q]);
    const uint32_t bits = (uint32_t)(p & 63);
    if (bits) c += (uint32_t)__popcll(mask[w] & ((1ull << bits) - 1ull));
    return c;
}

// the block's word prefixes in LDS (thread w < 16 reads word w): wp[w] =
// changed positions of the block before word w
__device__ __forceinline__ void block_word_prefix(const unsigned long long* mask, size_t nwords,
                                                  uint32_t* wp, unsigned long long* wm) {
    if (threadIdx.x < kWordsPerBlock) {
        const size_t w = (size_t)blockIdx.x * kWordsPerBlock + threadIdx.x;
        wm[threadIdx.x] = w < nwords ? mask[w] : 0ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int q = 0; q < kWordsPerBlock; ++q) {
            wp[q] = c;
            c += (uint32_t)__popcll(wm[q]);
        }
    }
    __syncthreads();
}

// the changed keys and their positions, compacted in position order
__global__ __launch_bounds__(kMBlock) void delta_compact_kernel(
    const uint32_t* __restrict__ key, size_t n, const unsigned long long* __restrict__ mask,
    const uint32_t* __restrict__ boff, uint32_t* __restrict__ ckey, uint32_t* __restrict__ cpos) {
    __shared__ uint32_t wp[kWordsPerBlock];
    __shared__ unsigned long long wm[kWordsPerBlock];
    block_word_prefix(mask, (n + 63) / 64, wp, wm);
    const size_t i = (size_t)blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long m = wm[wl];
    if (!((m >> lane) & 1ull)) return;
    const uint32_t c = boff[blockIdx.x] + wp[wl] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    ckey[c] = key[i];
    cpos[c] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void delta_pack_kernel(const uint32_t* __restrict__ ckey,
                                                         const uint32_t* __restrict__ cpos,
                                                         uint32_t nc,
                                                         unsigned long long* __restrict__ cpk) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r < nc) cpk[r] = packed_key(ckey[r], cpos[r]);
}

// at most kSmallSort changed keys: packed (key, position) values are
// distinct, so one bitonic sort in LDS gives their stable order
__global__ __launch_bounds__(1024) void delta_small_sort_kernel(const uint32_t* __restrict__ ckey,
                                                                const uint32_t* __restrict__ cpos,
                                                                uint32_t nc,
                                                                unsigned long long* __restrict__ cpk) {
    __shared__ unsigned long long a[kSmallSort];
    uint32_t P = 2;
    while (P < nc) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += 1024)
        a[i] = i < nc ? packed_key(ckey[i], cpos[i]) : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += 1024) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long u = a[i], v = a[ixj];
                    const bool asc = (i & k) == 0;
                    if ((u > v) == asc) {
                        a[i] = v;
                        a[ixj] = u;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < nc; i += 1024) cpk[i] = a[i];
}

// A whole sort of n <= kSmallSort particles in one block: (key bits 0..29,
// position) packed into distinct 64-bit values, one bitonic sort in LDS,
// then every particle moved once, in place -- all reads before a barrier,
// all writes after it.  The order by (key, position) is the stable order,
// the permutation of the LSD passes.
__global__ __launch_bounds__(1024) void small_sort_kernel(PsSortArrays io, float4* __restrict__ pos4,
                                                          uint32_t* __restrict__ kcopy,
                                                          uint32_t n) {
    __shared__ unsigned long long a[kSmallSort];
    uint32_t P = 2;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += 1024)
        a[i] = i < n ? packed_key(io.key[i], i) : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += 1024) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long u = a[i], v = a[ixj];
                    const bool asc = (i & k) == 0;
                    if ((u > v) == asc) {
                        a[i] = v;
                        a[ixj] = u;
                    }
                }
            }
            __syncthreads();
        }
    }
    constexpr uint32_t R = kSmallSort / 1024;
    uint32_t kk[R], id[R];
    float x[R], y[R], z[R], m[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t p = r * 1024 + threadIdx.x;
        if (p >= n) continue;
        const uint32_t src = (uint32_t)a[p];
        kk[r] = io.key[src];
        id[r] = io.id[src];
        if (io.x) {
            x[r] = io.x[src];
            y[r] = io.y[src];
            z[r] = io.z[src];
            m[r] = io.m[src];
        }
    }
    __syncthreads();  // every particle is read before any is written
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t p = r * 1024 + threadIdx.x;
        if (p >= n) continue;
        io.key[p] = kk[r];
        io.id[p] = id[r];
        if (kcopy) kcopy[p] = kk[r];
        if (io.x) {
            io.x[p] = x[r];
            io.y[p] = y[r];
            io.z[p] = z[r];
            io.m[p] = m[r];
            if (pos4) pos4[p] = make_float4(x[r], y[r], z[r], m[r]);
        }
    }
}

__device__ __forceinline__ void move_particle(const PsSortArrays& in, const PsSortArrays& out,
                                              float4* pos4, uint32_t* kcopy, size_t src,
                                              size_t dst) {
    const uint32_t k = in.key[src];
    const float vx = in.x[src], vy = in.y[src], vz = in.z[src], vm = in.m[src];
    out.key[dst] = k;
    out.id[dst] = in.id[src];
    out.x[dst] = vx;
    out.y[dst] = vy;
    out.z[dst] = vz;
    out.m[dst] = vm;
    if (pos4) pos4[dst] = make_float4(vx, vy, vz, vm);
    if (kcopy) kcopy[dst] = k;
}

// the unchanged particles: block b takes positions [b kMBlock, (b+1) kMBlock);
// the changed keys that can fall among them are one slice of the sorted
// changed keys, [lo, hi), held in LDS when it fits
// Where each merge block's changed keys start among the sorted ones: block b
// gets the first (key, position) at or after its first position's, block nb
// gets nc -- every block's binary search at once, ahead of the merge.
__global__ __launch_bounds__(256) void delta_range_kernel(const uint32_t* __restrict__ prev,
                                                          size_t n, uint32_t nb,
                                                          const unsigned long long* __restrict__ cpk,
                                                          uint32_t nc,
                                                          uint32_t* __restrict__ range) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b > nb) return;
    const size_t q = (size_t)b * kMBlock;
    range[b] = q < n ? lower_bound_u64(cpk, nc, packed_key(prev[q], (uint32_t)q)) : nc;
}

__global__ __launch_bounds__(kMBlock) void delta_merge_kernel(
    PsSortArrays in, PsSortArrays out, float4* __restrict__ pos4, uint32_t* __restrict__ kcopy,
    const uint32_t* __restrict__ prev, size_t n, const unsigned long long* __restrict__ mask,
    const uint32_t* __restrict__ boff, const unsigned long long* __restrict__ cpk, uint32_t nc,
    const uint32_t* __restrict__ range) {
    __shared__ unsigned long long slice[kSlice];
    __shared__ uint32_t wp[kWordsPerBlock];
    __shared__ unsigned long long wm[kWordsPerBlock];
    block_word_prefix(mask, (n + 63) / 64, wp, wm);
    const size_t b0 = (size_t)blockIdx.x * kMBlock;
    const size_t b1 = b0 + kMBlock < n ? b0 + kMBlock : n;
    const uint32_t tid = threadIdx.x;
    // the block's changed keys lie in [lo, hi): the probes of its positions
    // are at or after its first position's and before the next block's
    const uint32_t lo = range[blockIdx.x], hi = range[blockIdx.x + 1];
    const bool in_lds = hi - lo <= (uint32_t)kSlice;
    if (in_lds)
        for (uint32_t j = tid; j < hi - lo; j += kMBlock) slice[j] = cpk[lo + j];
    __syncthreads();
    const size_t i = b0 + tid;
    if (i >= b1) return;
    const uint32_t wl = tid >> 6, lane = tid & 63;
    const unsigned long long m = wm[wl];
    if ((m >> lane) & 1ull) return;  // changed: delta_scatter_kernel
    const uint32_t rank_u = (uint32_t)i - (boff[blockIdx.x] + wp[wl] +
                                           (uint32_t)__popcll(m & ((1ull << lane) - 1ull)));
    const unsigned long long probe = packed_key(prev[i], (uint32_t)i);
    const uint32_t before = in_lds ? lo + lower_bound_u64(slice, hi - lo, probe)
                                   : lo + lower_bound_u64(cpk + lo, hi - lo, probe);
    move_particle(in, out, pos4, kcopy, i, (size_t)rank_u + before);
}

// the changed particles, by their rank r in the sorted changed keys
__global__ __launch_bounds__(256) void delta_scatter_kernel(
    PsSortArrays in, PsSortArrays out, float4* __restrict__ pos4, uint32_t* __restrict__ kcopy,
    const uint32_t* __restrict__ prev, size_t n, const unsigned long long* __restrict__ mask,
    const uint32_t* __restrict__ boff, const unsigned long long* __restrict__ cpk, uint32_t nc) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nc) return;
    const unsigned long long v = cpk[r];
    const uint32_t k = (uint32_t)(v >> 32), src = (uint32_t)v;
    // the unchanged entries before (k, src): prev's positions p with
    // (prev[p], p) < (k, src), minus the changed positions among them
    size_t a, b;
    bound_keys(prev, n, k, a, b);
    const size_t p = src < a ? a : (src > b ? b : src);
    const size_t dst = (size_t)r + (p - changed_before(mask, boff, p));
    move_particle(in, out, pos4, kcopy, src, dst);
}

// out = in for the six arrays of n particles (16 bytes per load when every
// array is 16-byte aligned, vec)
__global__ __launch_bounds__(256) void copy6_kernel(PsSortArrays in, PsSortArrays out, size_t n,
                                                    int vec) {
    const size_t nv = vec ? n >> 2 : 0;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t v = (size_t)blockIdx.x * 256 + threadIdx.x; v < nv; v += stride) {
        reinterpret_cast<uint4*>(out.key)[v] = reinterpret_cast<const uint4*>(in.key)[v];
        reinterpret_cast<uint4*>(out.id)[v] = reinterpret_cast<const uint4*>(in.id)[v];
        reinterpret_cast<float4*>(out.x)[v] = reinterpret_cast<const float4*>(in.x)[v];
        reinterpret_cast<float4*>(out.y)[v] = reinterpret_cast<const float4*>(in.y)[v];
        reinterpret_cast<float4*>(out.z)[v] = reinterpret_cast<const float4*>(in.z)[v];
        reinterpret_cast<float4*>(out.m)[v] = reinterpret_cast<const float4*>(in.m)[v];
    }
    for (size_t i = (nv << 2) + (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        out.key[i] = in.key[i];
        out.id[i] = in.id[i];
        out.x[i] = in.x[i];
        out.y[i] = in.y[i];
        out.z[i] = in.z[i];
        out.m[i] = in.m[i];
    }
}

}  // namespace

size_t ps_sort_tmp_bytes(size_t n) {
    size_t tiles = (n + kTile - 1) / kTile;
    size_t table = (size_t)kRadix * (tiles ? tiles : 1);
    const size_t delta = ps_sort_delta_tmp_bytes(n);
    const size_t radix = align256(table * sizeof(uint32_t)) + align256(kRadix * sizeof(uint32_t));
    // the incremental sort keeps its flag words while a radix sort of the
    // changed keys uses the tables after them
    return delta + radix;
}

hipError_t ps_radix_sort(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                         size_t n, void* tmp, hipStream_t s, uint32_t* kcopy) {
    if (n == 0) return hipSuccess;
    if (n <= kSmallSort && !(io.aux && t0.aux && t1.aux)) {  // one block, one launch
        small_sort_kernel<<<1, 1024, 0, s>>>(io, pos4, kcopy, (uint32_t)n);
        return hipGetLastError();
    }
    const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
    const size_t table_n = (size_t)kRadix * tiles;
    uint32_t* table = reinterpret_cast<uint32_t*>(tmp);
    uint32_t* dtot =
        reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(tmp) + align256(table_n * sizeof(uint32_t)));
    // three passes over the code (io -> t0 -> t1 -> io); with a second key,
    // io -> t0 -> t1 -> t0 by code, then t0 -> io by aux (code carried)
    const bool two = io.aux && t0.aux && t1.aux;
    const PsSortArrays* src[4] = {&io, &t0, &t1, &t0};
    const PsSortArrays* dst[4] = {&t0, &t1, two ? &t0 : &io, &io};
    const int passes = two ? 4 : 3;
    for (int pass = 0; pass < passes; ++pass) {
        const bool by_aux = pass == 3;
        const int shift = by_aux ? 0 : pass * kRadixBits;
        const PsSortArrays& a = *src[pass];
        const PsSortArrays& b = *dst[pass];
        const uint32_t* kin = by_aux ? a.aux : a.key;
        uint32_t* kout = by_aux ? b.aux : b.key;
        const uint32_t* ain = two ? (by_aux ? a.key : a.aux) : nullptr;
        uint32_t* aout = two ? (by_aux ? b.key : b.aux) : nullptr;
        sort_upsweep_kernel<<<tiles, kBlock, 0, s>>>(kin, n, shift, table, tiles, nullptr);
        sort_rowscan_kernel<<<kRadix, kRowBlock, 0, s>>>(table, tiles, dtot);
        sort_downsweep_kernel<<<tiles, kDBlock, 0, s>>>(
            kin, a.id, a.x, a.y, a.z, a.m, kout, b.id, b.x, b.y, b.z, b.m,
            pass == passes - 1 ? pos4 : nullptr, ain, aout, n, shift, table, dtot, tiles,
            pass == passes - 1 ? kcopy : nullptr, nullptr, nullptr, nullptr);
    }
    return hipGetLastError();
}

size_t ps_sort_seg_tiles(const uint32_t* off, uint32_t n_sys, uint32_t* vstart, uint32_t* vseg,
                         uint32_t* seg_tile) {
    size_t t = 0;
    for (uint32_t sg = 0; sg < n_sys; ++sg) {
        if (seg_tile) seg_tile[sg] = (uint32_t)t;
        for (uint32_t b = off[sg]; b < off[sg + 1]; b += (uint32_t)kTile, ++t) {
            if (vstart) vstart[t] = b;
            if (vseg) vseg[t] = sg;
        }
    }
    if (seg_tile) seg_tile[n_sys] = (uint32_t)t;
    if (vstart) vstart[t] = n_sys ? off[n_sys] : 0u;
    return t;
}

size_t ps_sort_seg_tmp_bytes(size_t n_tiles, uint32_t n_sys) {
    const size_t table = (size_t)kRadix * (n_tiles ? n_tiles : 1);
    return align256(table * sizeof(uint32_t)) + align256(kRadix * sizeof(uint32_t)) +
           align256((size_t)kRadix * (n_sys ? n_sys : 1) * sizeof(uint32_t));
}

hipError_t ps_radix_sort_seg(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                             size_t n, PsSortSegs segs, void* tmp, hipStream_t s) {
    if (n == 0 || segs.n_tiles == 0) return hipSuccess;
    const uint32_t tiles = segs.n_tiles;
    const size_t table_n = (size_t)kRadix * tiles;
    char* p = reinterpret_cast<char*>(tmp);
    uint32_t* table = reinterpret_cast<uint32_t*>(p);
    p += align256(table_n * sizeof(uint32_t));
    uint32_t* dtot = reinterpret_cast<uint32_t*>(p);
    p += align256(kRadix * sizeof(uint32_t));
    uint32_t* sbase = reinterpret_cast<uint32_t*>(p);
    // three passes over the code, io -> t0 -> t1 -> io; every key stays in
    // its segment (the segments' other arrays, aux included, need no move)
    const PsSortArrays* src[3] = {&io, &t0, &t1};
    const PsSortArrays* dst[3] = {&t0, &t1, &io};
    for (int pass = 0; pass < 3; ++pass) {
        const int shift = pass * kRadixBits;
        const PsSortArrays& a = *src[pass];
        const PsSortArrays& b = *dst[pass];
        sort_upsweep_kernel<<<tiles, kBlock, 0, s>>>(a.key, n, shift, table, tiles, segs.vstart);
        sort_rowscan_kernel<<<kRadix, kRowBlock, 0, s>>>(table, tiles, dtot);
        sort_segbase_kernel<<<segs.n_sys, kRadix, 0, s>>>(table, tiles, dtot, segs.seg_tile,
                                                          segs.off, sbase);
        sort_downsweep_kernel<<<tiles, kDBlock, 0, s>>>(
            a.key, a.id, a.x, a.y, a.z, a.m, b.key, b.id, b.x, b.y, b.z, b.m,
            pass == 2 ? pos4 : nullptr, nullptr, nullptr, n, shift, table, dtot, tiles, nullptr,
            segs.vstart, segs.vseg, sbase);
    }
    return hipGetLastError();
}

namespace {
struct DeltaTmp {
    unsigned long long* mask;  // one word per 64 positions, padded to whole blocks
    uint32_t* bcnt;            // per block of kMBlock positions
    uint32_t* boff;            // exclusive prefix of bcnt, nb + 1 entries
    uint32_t* range;           // per block: its first changed key in the sorted ones, nb + 1
    void* rest;                // the changed keys' radix tables
};
size_t delta_blocks(size_t n) { return (n + kMBlock - 1) / kMBlock; }
DeltaTmp delta_tmp(void* tmp, size_t n) {
    const size_t nb = delta_blocks(n);
    char* p = reinterpret_cast<char*>(tmp);
    DeltaTmp d;
    d.mask = reinterpret_cast<unsigned long long*>(p);
    p += align256(nb * kWordsPerBlock * 8 + 8);
    d.bcnt = reinterpret_cast<uint32_t*>(p);
    p += align256(nb * 4 + 4);
    d.boff = reinterpret_cast<uint32_t*>(p);
    p += align256(nb * 4 + 8);
    d.range = reinterpret_cast<uint32_t*>(p);
    p += align256(nb * 4 + 8);
    d.rest = p;
    return d;
}
}  // namespace

size_t ps_sort_delta_tmp_bytes(size_t n) {
    const size_t nb = delta_blocks(n);
    return align256(nb * kWordsPerBlock * 8 + 8) + align256(nb * 4 + 4) +
           2 * align256(nb * 4 + 8);
}

hipError_t ps_sort_delta_count(const uint32_t* key, const uint32_t* prev, size_t n, void* tmp,
                               uint32_t* ckey, uint32_t* cpos, unsigned long long* count,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    DeltaTmp d = delta_tmp(tmp, n);
    const uint32_t nb = (uint32_t)delta_blocks(n);
    delta_mask_kernel<<<nb, kMBlock, 0, s>>>(key, prev, n, d.mask, d.bcnt);
    delta_scan_kernel<<<1, 1024, 0, s>>>(d.bcnt, nb, d.boff, count);
    delta_compact_kernel<<<nb, kMBlock, 0, s>>>(key, n, d.mask, d.boff, ckey, cpos);
    return hipGetLastError();
}

hipError_t ps_sort_delta_finish(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                                size_t n, void* tmp, const uint32_t* prev, uint32_t* ckey,
                                uint32_t* cpos, uint32_t nc, uint32_t* kcopy, hipStream_t s,
                                hipStream_t s_copy, hipEvent_t ev_sorted, hipEvent_t ev_copied) {
    if (n == 0) return hipSuccess;
    DeltaTmp d = delta_tmp(tmp, n);
    // the changed (key, position) pairs sorted stably by key: t0 and t1 key
    // and id arrays are the scratch, the result ends in (ckey, cpos)
    unsigned long long* cpk = reinterpret_cast<unsigned long long*>(t1.z);  // t1.z and t1.m
    if (nc > kSmallSort) {
        PsSortArrays c = {ckey, cpos, nullptr, nullptr, nullptr, nullptr, nullptr};
        PsSortArrays c0 = {t0.key, t0.id, nullptr, nullptr, nullptr, nullptr, nullptr};
        PsSortArrays c1 = {t1.key, t1.id, nullptr, nullptr, nullptr, nullptr, nullptr};
        // the radix tables after the flag words, which the merge still reads
        hipError_t e = ps_radix_sort(c, c0, c1, nullptr, nc, d.rest, s, nullptr);
        if (e != hipSuccess) return e;
        delta_pack_kernel<<<(nc + 255) / 256, 256, 0, s>>>(ckey, cpos, nc, cpk);
    } else if (nc) {
        delta_small_sort_kernel<<<1, 1024, 0, s>>>(ckey, cpos, nc, cpk);
    }
    // the merged order into t0, then back into the caller's arrays
    const uint32_t mb = (uint32_t)delta_blocks(n);
    delta_range_kernel<<<(mb + 1 + 255) / 256, 256, 0, s>>>(prev, n, mb, cpk, nc, d.range);
    delta_merge_kernel<<<mb, kMBlock, 0, s>>>(io, t0, pos4, kcopy, prev, n, d.mask, d.boff, cpk,
                                              nc, d.range);
    if (nc)
        delta_scatter_kernel<<<(nc + 255) / 256, 256, 0, s>>>(io, t0, pos4, kcopy, prev, n,
                                                              d.mask, d.boff, cpk, nc);
    size_t cb = (n / 4 + 255) / 256 + 1;
    if (cb > 8192) cb = 8192;
    if (cb == 0) cb = 1;
    const void* ptrs[12] = {t0.key, t0.id, t0.x, t0.y, t0.z, t0.m,
                            io.key, io.id, io.x, io.y, io.z, io.m};
    int vec = 1;
    for (const void* q : ptrs)
        if (reinterpret_cast<uintptr_t>(q) & 15u) vec = 0;
    if (s_copy && ev_sorted && ev_copied) {  // beside the work that follows on s
        hipError_t e = hipEventRecord(ev_sorted, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(s_copy, ev_sorted, 0);
        if (e != hipSuccess) return e;
        copy6_kernel<<<(unsigned)cb, 256, 0, s_copy>>>(t0, io, n, vec);
        return hipEventRecord(ev_copied, s_copy);
    }
    copy6_kernel<<<(unsigned)cb, 256, 0, s>>>(t0, io, n, vec);
    return hipGetLastError();
}
