// ps_scan.hip -- device-wide exclusive prefix sums (reduce-then-scan).
//
// Tiles of 2048 elements (256 threads x 8).  Pass 1 reduces each tile and adds
// its sum into exact 64-bit totals; the tile sums are scanned recursively; pass
// 2 rescans each tile from its offset.  Used for the radix-sort digit table,
// the per-depth node counts of the tree build and the traversal's per-entry
// output counts.
#include "ps_kernels.h"
#include "ps_device.h"

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kWaves = kBlock / 64;

struct OpsU32 {
    using T = uint32_t;
    static constexpr int C = 1;
    __device__ static T zero() { return 0u; }
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static unsigned long long comp(T v, int) { return v; }
    __device__ static T shfl_up(T v, int d) { return __shfl_up(v, d, 64); }
    __device__ static T shfl_down(T v, int d) { return __shfl_down(v, d, 64); }
};

struct OpsU4 {
    using T = uint4;
    static constexpr int C = 4;
    __device__ static T zero() { return make_uint4(0, 0, 0, 0); }
    __device__ static T add(T a, T b) {
        return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    }
    __device__ static unsigned long long comp(T v, int c) {
        return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
    }
    __device__ static T shfl_up(T v, int d) {
        return make_uint4(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64), __shfl_up(v.z, d, 64),
                          __shfl_up(v.w, d, 64));
    }
    __device__ static T shfl_down(T v, int d) {
        return make_uint4(__shfl_down(v.x, d, 64), __shfl_down(v.y, d, 64),
                          __shfl_down(v.z, d, 64), __shfl_down(v.w, d, 64));
    }
};

// tile sums + exact 64-bit totals
template <class Ops>
__global__ __launch_bounds__(kBlock) void scan_reduce_kernel(const typename Ops::T* __restrict__ in,
                                                             size_t n, typename Ops::T* bsum,
                                                             unsigned long long* total) {
    using T = typename Ops::T;
    __shared__ T wsum[kWaves];
    const size_t base = (size_t)blockIdx.x * kTile;
    T acc = Ops::zero();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        size_t i = base + (size_t)k * kBlock + threadIdx.x;
        if (i < n) acc = Ops::add(acc, in[i]);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) acc = Ops::add(acc, Ops::shfl_down(acc, d));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wsum[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = Ops::zero();
        for (int w = 0; w < kWaves; ++w) s = Ops::add(s, wsum[w]);
        bsum[blockIdx.x] = s;
        if (total) {
            for (int c = 0; c < Ops::C; ++c) atomicAdd(&total[c], Ops::comp(s, c));
        }
    }
}

// device-side element count: tiles at or beyond *n_dev contribute zeros
template <class Ops>
__global__ __launch_bounds__(kBlock) void scan_reduce_dev_kernel(
    const typename Ops::T* __restrict__ in, const uint32_t* __restrict__ n_dev,
    typename Ops::T* bsum, unsigned long long* total) {
    using T = typename Ops::T;
    __shared__ T wsum[kWaves];
    const size_t n = *n_dev;
    const size_t base = (size_t)blockIdx.x * kTile;
    if (base >= n) {
        if (threadIdx.x == 0) bsum[blockIdx.x] = Ops::zero();
        return;
    }
    T acc = Ops::zero();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        size_t i = base + (size_t)k * kBlock + threadIdx.x;
        if (i < n) acc = Ops::add(acc, in[i]);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) acc = Ops::add(acc, Ops::shfl_down(acc, d));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wsum[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = Ops::zero();
        for (int w = 0; w < kWaves; ++w) s = Ops::add(s, wsum[w]);
        bsum[blockIdx.x] = s;
        if (total) {
            for (int c = 0; c < Ops::C; ++c) atomicAdd(&total[c], Ops::comp(s, c));
        }
    }
}

// rescan each tile from its offset; total (a single tile only): its exact
// 64-bit sums, stored
template <class Ops>
__global__ __launch_bounds__(kBlock) void scan_down_kernel(const typename Ops::T* in,
                                                           typename Ops::T* out, size_t n,
                                                           const typename Ops::T* boff,
                                                           unsigned long long* total = nullptr) {
    using T = typename Ops::T;
    __shared__ T tile[kTile];
    __shared__ T wsum[kWaves];
    __shared__ unsigned long long tsum[Ops::C][kWaves];
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        size_t i = base + (size_t)k * kBlock + threadIdx.x;
        tile[k * kBlock + threadIdx.x] = (i < n) ? in[i] : Ops::zero();
    }
    __syncthreads();
    T local[kItems];
    T s = Ops::zero();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        local[k] = tile[threadIdx.x * kItems + k];
        s = Ops::add(s, local[k]);
    }
    // inclusive wave scan of the per-thread sums
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (total) {  // block-uniform
#pragma unroll
        for (int c = 0; c < Ops::C; ++c) {
            unsigned long long v = 0;
#pragma unroll
            for (int k = 0; k < kItems; ++k) v += Ops::comp(local[k], c);
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
            if (lane == 0) tsum[c][wave] = v;
        }
    }
    T inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = Ops::shfl_up(inc, d);
        if (lane >= d) inc = Ops::add(inc, o);
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    T run = boff ? boff[blockIdx.x] : Ops::zero();
    for (int w = 0; w < wave; ++w) run = Ops::add(run, wsum[w]);
    T excl = Ops::shfl_up(inc, 1);
    if (lane > 0) run = Ops::add(run, excl);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        tile[threadIdx.x * kItems + k] = run;
        run = Ops::add(run, local[k]);
    }
    __syncthreads();
    if (total && threadIdx.x < (unsigned)Ops::C) {
        unsigned long long v = 0;
        for (int w = 0; w < kWaves; ++w) v += tsum[threadIdx.x][w];
        total[threadIdx.x] = v;
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        size_t i = base + (size_t)k * kBlock + threadIdx.x;
        if (i < n) out[i] = tile[k * kBlock + threadIdx.x];
    }
}

template <class Ops>
__global__ __launch_bounds__(kBlock) void scan_down_dev_kernel(const typename Ops::T* in,
                                                               typename Ops::T* out,
                                                               const uint32_t* __restrict__ n_dev,
                                                               const typename Ops::T* boff) {
    using T = typename Ops::T;
    __shared__ T tile[kTile];
    __shared__ T wsum[kWaves];
    const size_t n = *n_dev;
    const size_t base = (size_t)blockIdx.x * kTile;
    if (base >= n) return;  // block-uniform
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        size_t i = base + (size_t)k * kBlock + threadIdx.x;
        tile[k * kBlock + threadIdx.x] = (i < n) ? in[i] : Ops::zero();
    }
    __syncthreads();
    T local[kItems];
    T s = Ops::zero();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        local[k] = tile[threadIdx.x * kItems + k];
        s = Ops::add(s, local[k]);
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    T inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = Ops::shfl_up(inc, d);
        if (lane >= d) inc = Ops::add(inc, o);
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    T run = boff ? boff[blockIdx.x] : Ops::zero();
    for (int w = 0; w < wave; ++w) run = Ops::add(run, wsum[w]);
    T excl = Ops::shfl_up(inc, 1);
    if (lane > 0) run = Ops::add(run, excl);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        tile[threadIdx.x * kItems + k] = run;
        run = Ops::add(run, local[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        size_t i = base + (size_t)k * kBlock + threadIdx.x;
        if (i < n) out[i] = tile[k * kBlock + threadIdx.x];
    }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// p[0 .. per * min(*count, cap)) = v, the count read on the device
// p[0 .. per * m) = v and p2[0 .. per2 * m) = v (p2 may be null), with
// m = min(*count, cap)
__global__ void fill_u32_count_kernel(uint32_t* p, uint32_t per, uint32_t* p2, uint32_t per2,
                                      uint32_t v, const unsigned long long* count, size_t cap) {
    const unsigned long long c = *count;
    const size_t m = (size_t)(c < cap ? c : cap);
    const size_t n = (size_t)per * m, n2 = p2 ? (size_t)per2 * m : 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
        p2[i] = v;
}

// Sum of n u64 values into *out: each block sums its slice (kRedV
// independent loads per thread per round); a single block stores the sum,
// several add theirs to *out, which the launcher clears first.
constexpr int kRedThreads = 1024;
constexpr int kRedV = 8;
constexpr size_t kRedSlice = (size_t)kRedThreads * kRedV * 4;  // values per block

__global__ void clear_u64_kernel(unsigned long long* p) { *p = 0ull; }

__global__ __launch_bounds__(kRedThreads) void reduce_u64_kernel(const unsigned long long* in,
                                                                 size_t n,
                                                                 unsigned long long* out) {
    __shared__ unsigned long long ws[kRedThreads / 64];
    const size_t lo = (size_t)blockIdx.x * kRedSlice;
    const size_t hi = lo + kRedSlice < n ? lo + kRedSlice : n;
    unsigned long long v = 0;
    for (size_t base = lo; base < hi; base += (size_t)kRedThreads * kRedV) {
        unsigned long long t[kRedV];
#pragma unroll
        for (int j = 0; j < kRedV; ++j) {
            const size_t i = base + (size_t)j * kRedThreads + threadIdx.x;
            t[j] = i < hi ? in[i] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < kRedV; ++j) v += t[j];
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < kRedThreads / 64; ++w) s += ws[w];
        if (gridDim.x == 1)
            *out = s;
        else
            atomicAdd(out, s);
    }
}

__global__ void fill_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <class Ops>
size_t scan_tmp_bytes(size_t n) {
    size_t bytes = 0;
    size_t m = n;
    while (m > (size_t)kTile) {
        m = (m + kTile - 1) / kTile;
        bytes += 2 * align256(m * sizeof(typename Ops::T));
    }
    return bytes + 256;
}

template <class Ops>
hipError_t scan_impl(const typename Ops::T* in, typename Ops::T* out, size_t n,
                     unsigned long long* total, void* tmp, hipStream_t s) {
    using T = typename Ops::T;
    size_t nb = (n + kTile - 1) / kTile;
    if (nb <= 1) {  // one launch: the tile's scan and its totals
        if (n == 0) {
            if (total) fill_u64_kernel<<<1, 64, 0, s>>>(total, Ops::C);
            return hipGetLastError();
        }
        scan_down_kernel<Ops><<<1, kBlock, 0, s>>>(in, out, n, nullptr, total);
        return hipGetLastError();
    }
    if (total) fill_u64_kernel<<<1, 64, 0, s>>>(total, Ops::C);
    T* bsum = reinterpret_cast<T*>(tmp);
    T* boff = reinterpret_cast<T*>(reinterpret_cast<char*>(tmp) + align256(nb * sizeof(T)));
    void* next = reinterpret_cast<char*>(tmp) + 2 * align256(nb * sizeof(T));
    scan_reduce_kernel<Ops><<<(unsigned)nb, kBlock, 0, s>>>(in, n, bsum, total);
    hipError_t e = scan_impl<Ops>(bsum, boff, nb, nullptr, next, s);
    if (e != hipSuccess) return e;
    scan_down_kernel<Ops><<<(unsigned)nb, kBlock, 0, s>>>(in, out, n, boff);
    return hipGetLastError();
}

template <class Ops>
hipError_t scan_dev_impl(const typename Ops::T* in, typename Ops::T* out, const uint32_t* n_dev,
                         size_t n_max, unsigned long long* total, void* tmp, hipStream_t s) {
    using T = typename Ops::T;
    if (total) fill_u64_kernel<<<1, 64, 0, s>>>(total, Ops::C);
    if (n_max == 0) return hipGetLastError();
    size_t nb = (n_max + kTile - 1) / kTile;
    T* bsum = reinterpret_cast<T*>(tmp);
    T* boff = reinterpret_cast<T*>(reinterpret_cast<char*>(tmp) + align256(nb * sizeof(T)));
    void* next = reinterpret_cast<char*>(tmp) + 2 * align256(nb * sizeof(T));
    scan_reduce_dev_kernel<Ops><<<(unsigned)nb, kBlock, 0, s>>>(in, n_dev, bsum, total);
    if (nb > 1) {
        hipError_t e = scan_impl<Ops>(bsum, boff, nb, nullptr, next, s);
        if (e != hipSuccess) return e;
    }
    scan_down_dev_kernel<Ops><<<(unsigned)nb, kBlock, 0, s>>>(in, out, n_dev,
                                                               nb > 1 ? boff : nullptr);
    return hipGetLastError();
}

}  // namespace

size_t ps_scan_dev_tmp_bytes_u4(size_t n_max) {
    size_t nb = (n_max + kTile - 1) / kTile;
    return 2 * align256(nb * sizeof(uint4)) + scan_tmp_bytes<OpsU4>(nb) + 256;
}
size_t ps_scan_dev_tmp_bytes_u32(size_t n_max) {
    size_t nb = (n_max + kTile - 1) / kTile;
    return 2 * align256(nb * sizeof(uint32_t)) + scan_tmp_bytes<OpsU32>(nb) + 256;
}

hipError_t ps_scan_dev_u4(const uint4* in, uint4* out, const uint32_t* n_dev, size_t n_max,
                          unsigned long long* total64, void* tmp, hipStream_t s) {
    return scan_dev_impl<OpsU4>(in, out, n_dev, n_max, total64, tmp, s);
}
hipError_t ps_scan_dev_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev,
                           size_t n_max, unsigned long long* total64, void* tmp, hipStream_t s) {
    return scan_dev_impl<OpsU32>(in, out, n_dev, n_max, total64, tmp, s);
}

size_t ps_scan_tmp_bytes_u32(size_t n) { return scan_tmp_bytes<OpsU32>(n); }
size_t ps_scan_tmp_bytes_u4(size_t n) { return scan_tmp_bytes<OpsU4>(n); }

hipError_t ps_scan_u32(const uint32_t* in, uint32_t* out, size_t n, unsigned long long* total64,
                       void* tmp, hipStream_t s) {
    return scan_impl<OpsU32>(in, out, n, total64, tmp, s);
}

hipError_t ps_scan_u4(const uint4* in, uint4* out, size_t n, unsigned long long* total64,
                      void* tmp, hipStream_t s) {
    return scan_impl<OpsU4>(in, out, n, total64, tmp, s);
}

void ps_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_u32_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, v, n);
}

void ps_fill_u32_count(uint32_t* p, uint32_t v, const unsigned long long* count, uint32_t per,
                       size_t cap, hipStream_t s) {
    if (cap == 0 || per == 0) return;
    size_t blocks = (per * cap + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_u32_count_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, per, nullptr, 0u, v, count, cap);
}

void ps_fill2_u32_count(uint32_t* p, uint32_t per, uint32_t* p2, uint32_t per2, uint32_t v,
                        const unsigned long long* count, size_t cap, hipStream_t s) {
    if (cap == 0 || per + per2 == 0) return;
    size_t blocks = ((size_t)(per > per2 ? per : per2) * cap + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_u32_count_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, per, p2, per2, v, count, cap);
}

void ps_reduce_u64(const unsigned long long* in, size_t n, unsigned long long* out,
                   hipStream_t s) {
    const unsigned blocks = (unsigned)((n + kRedSlice - 1) / kRedSlice);
    if (blocks > 1) clear_u64_kernel<<<1, 1, 0, s>>>(out);
    reduce_u64_kernel<<<blocks > 1 ? blocks : 1u, kRedThreads, 0, s>>>(in, n, out);
}
