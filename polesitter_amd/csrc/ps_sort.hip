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
// arrays (H:1688-1690).  Many systems
// sorted at once carry their system id as a second key: a fourth, stable pass
// by it groups the systems again, each in code order.
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

__global__ __launch_bounds__(kBlock) void sort_upsweep_kernel(const uint32_t* __restrict__ key,
                                                              size_t n, int shift,
                                                              uint32_t* __restrict__ table,
                                                              uint32_t n_tiles) {
    __shared__ uint32_t h[kRadix];
    for (int d = threadIdx.x; d < kRadix; d += kBlock) h[d] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll 4
    for (int r = 0; r < kItems; ++r) {
        size_t i = base + (size_t)r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & (kRadix - 1)], 1u);
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

__global__ __launch_bounds__(kDBlock) void sort_downsweep_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ iin,
    const float* __restrict__ xin, const float* __restrict__ yin, const float* __restrict__ zin,
    const float* __restrict__ min_, uint32_t* __restrict__ kout, uint32_t* __restrict__ iout,
    float* __restrict__ xout, float* __restrict__ yout, float* __restrict__ zout,
    float* __restrict__ mout, float4* __restrict__ pos4, const uint32_t* __restrict__ ain,
    uint32_t* __restrict__ aout, size_t n, int shift, const uint32_t* __restrict__ offs,
    const uint32_t* __restrict__ dtot, uint32_t n_tiles) {
    __shared__ uint32_t keys[kTile];
    __shared__ uint16_t sidx[kTile];
    __shared__ uint32_t tstart[kRadix];
    __shared__ uint32_t goff[kRadix];
    __shared__ uint16_t wcnt[kWaves][kRadix];
    __shared__ uint32_t wsum[kWaves];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * kTile;
    const uint32_t valid_n = (uint32_t)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);

    // where digit tid's run of this tile goes: the digit's start (exclusive
    // scan of the totals, one digit per thread) + the tile's offset in it
    {
        const uint32_t d = (uint32_t)tid;
        const uint32_t own = offs[(size_t)d * n_tiles + blockIdx.x];
        const uint32_t c = dtot[d];
        uint32_t wt;
        uint32_t ex = ps_wave_excl_scan(c, &wt);
        if (lane == 0) wsum[wave] = wt;
        tstart[d] = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcnt[w][d] = 0;
        __syncthreads();
        for (int w = 0; w < wave; ++w) ex += wsum[w];
        goff[d] = ex + own;
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
        float vx = xin[src], vy = yin[src], vz = zin[src], vm = min_[src];
        kout[g] = k;
        iout[g] = iin[src];
        xout[g] = vx;
        yout[g] = vy;
        zout[g] = vz;
        mout[g] = vm;
        if (pos4) pos4[g] = make_float4(vx, vy, vz, vm);
        if (aout) aout[g] = ain[src];
    }
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t ps_sort_tmp_bytes(size_t n) {
    size_t tiles = (n + kTile - 1) / kTile;
    size_t table = (size_t)kRadix * (tiles ? tiles : 1);
    return align256(table * sizeof(uint32_t)) + align256(kRadix * sizeof(uint32_t));
}

hipError_t ps_radix_sort(PsSortArrays io, PsSortArrays t0, PsSortArrays t1, float4* pos4,
                         size_t n, void* tmp, hipStream_t s) {
    if (n == 0) return hipSuccess;
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
        sort_upsweep_kernel<<<tiles, kBlock, 0, s>>>(kin, n, shift, table, tiles);
        sort_rowscan_kernel<<<kRadix, kRowBlock, 0, s>>>(table, tiles, dtot);
        sort_downsweep_kernel<<<tiles, kDBlock, 0, s>>>(
            kin, a.id, a.x, a.y, a.z, a.m, kout, b.id, b.x, b.y, b.z, b.m,
            pass == passes - 1 ? pos4 : nullptr, ain, aout, n, shift, table, dtot, tiles);
    }
    return hipGetLastError();
}
