// ps_prepare.hip -- ps_prepare_particles on the device (H:1911-2170).
//
// Two HBM-bound passes: a min/max reduction over x, y and z (one cube, exact,
// order independent), then Morton quantisation + bit interleave, one thread
// per particle.  Algorithmic traffic: 12 B read per particle per pass + 4 B
// of codes written.
#include <float.h>

#include <mutex>

#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxParts = 1024;

__device__ __forceinline__ void wave_minmax(float& lo, float& hi) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = fminf(lo, __shfl_down(lo, d, 64));
        hi = fmaxf(hi, __shfl_down(hi, d, 64));
    }
}

__global__ __launch_bounds__(kBlock) void bbox_partial_kernel(const float* __restrict__ x,
                                                              const float* __restrict__ y,
                                                              const float* __restrict__ z,
                                                              size_t n, float2* part) {
    __shared__ float slo[kBlock / 64], shi[kBlock / 64];
    float lo = FLT_MAX, hi = -FLT_MAX;
    size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        float a = x[i], b = y[i], c = z[i];
        lo = fminf(lo, fminf(a, fminf(b, c)));
        hi = fmaxf(hi, fmaxf(a, fmaxf(b, c)));
    }
    wave_minmax(lo, hi);
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            lo = fminf(lo, slo[w]);
            hi = fmaxf(hi, shi[w]);
        }
        part[blockIdx.x] = make_float2(lo, hi);
    }
}

// the cube's parameters from its min and max: (min, max, range, 1023/range)
__device__ __forceinline__ void cube_params(float lo, float hi, float* prm) {
#pragma clang fp contract(off)
    float range = hi - lo;  // H:2008-2011
    if (range < 0.001f) range = 0.001f;
    prm[0] = lo;
    prm[1] = hi;
    prm[2] = range;
    prm[3] = 1023.0f / range;  // H:2017
}

__global__ __launch_bounds__(kBlock) void bbox_final_kernel(const float2* part, int np,
                                                            float* prm) {
    __shared__ float slo[kBlock / 64], shi[kBlock / 64];
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (int i = threadIdx.x; i < np; i += kBlock) {
        lo = fminf(lo, part[i].x);
        hi = fmaxf(hi, part[i].y);
    }
    wave_minmax(lo, hi);
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            lo = fminf(lo, slo[w]);
            hi = fmaxf(hi, shi[w]);
        }
        cube_params(lo, hi, prm);
    }
}

__device__ __forceinline__ uint32_t q_trunc(float v, float lo, float range) {
#pragma clang fp contract(off)
    // H:2133-2153: (int)(((v - min) / range) * 1023), clamped to [0, 1023]
    int q = (int)(((v - lo) / range) * 1023.0f);
    q = q < 0 ? 0 : q;
    q = q > 1023 ? 1023 : q;
    return (uint32_t)q;
}

__device__ __forceinline__ uint32_t q_rne(float v, float lo, float scale) {
#pragma clang fp contract(off)
    // H:2046-2058: (v - min) * fl(1023/range), min(., 1023), max(0, .), then
    // round-to-nearest-even conversion
    float t = (v - lo) * scale;
    t = t < 1023.0f ? t : 1023.0f;
    t = 0.0f > t ? 0.0f : t;
    return (uint32_t)(int)rintf(t);
}

// particle i's code in the cube (lo, range, scale = 1023/range)
__device__ __forceinline__ uint32_t particle_code(float vx, float vy, float vz, float lo,
                                                 float range, float scale, bool rne) {
    uint32_t qx, qy, qz;
    if (rne) {
        qx = q_rne(vx, lo, scale);
        qy = q_rne(vy, lo, scale);
        qz = q_rne(vz, lo, scale);
    } else {
        qx = q_trunc(vx, lo, range);
        qy = q_trunc(vy, lo, range);
        qz = q_trunc(vz, lo, range);
    }
    return ps_morton3(qx, qy, qz);
}

__global__ __launch_bounds__(kBlock) void morton_kernel(const float* __restrict__ x,
                                                        const float* __restrict__ y,
                                                        const float* __restrict__ z, size_t n,
                                                        const float* __restrict__ prm,
                                                        uint32_t* __restrict__ codes, int mode,
                                                        size_t n_vec) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    codes[i] = particle_code(x[i], y[i], z[i], prm[0], prm[2], prm[3], mode == 1 && i < n_vec);
}

// The whole prepare pass of n <= kSmallPrep particles in one block: the
// min/max reduction, the cube's parameters (to prm), and the codes.
constexpr int kSmallBlock = 1024;
constexpr size_t kSmallPrep = 4096;
__global__ __launch_bounds__(kSmallBlock) void prepare_small_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    uint32_t n, float* __restrict__ prm, uint32_t* __restrict__ codes, int mode) {
    __shared__ float slo[kSmallBlock / 64], shi[kSmallBlock / 64];
    __shared__ float sp[4];
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (uint32_t i = threadIdx.x; i < n; i += kSmallBlock) {
        const float a = x[i], b = y[i], c = z[i];
        lo = fminf(lo, fminf(a, fminf(b, c)));
        hi = fmaxf(hi, fmaxf(a, fmaxf(b, c)));
    }
    wave_minmax(lo, hi);
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kSmallBlock / 64; ++w) {
            lo = fminf(lo, slo[w]);
            hi = fmaxf(hi, shi[w]);
        }
        cube_params(lo, hi, sp);
        for (int k = 0; k < 4; ++k) prm[k] = sp[k];
    }
    __syncthreads();
    const uint32_t n_vec = n - (n % 8u);
    for (uint32_t i = threadIdx.x; i < n; i += kSmallBlock)
        codes[i] = particle_code(x[i], y[i], z[i], sp[0], sp[2], sp[3], mode == 1 && i < n_vec);
}

// ---- many systems in one set of arrays: system s is [off[s], off[s+1]) ----

constexpr int kSegBlock = 1024;

// one block per system: its bounding cube, root cell and quantisation
// parameters (as bbox_final_kernel), prm[s] = (min, max, range, 1023/range),
// root[s] = (min + range/2 (x3), range/2)
__global__ __launch_bounds__(kSegBlock) void bbox_seg_kernel(const float* __restrict__ x,
                                                             const float* __restrict__ y,
                                                             const float* __restrict__ z,
                                                             const uint32_t* __restrict__ off,
                                                             float4* __restrict__ prm,
                                                             float4* __restrict__ root) {
    __shared__ float slo[kSegBlock / 64], shi[kSegBlock / 64];
    const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (uint32_t i = b + threadIdx.x; i < e; i += kSegBlock) {
        float a = x[i], c = y[i], d = z[i];
        lo = fminf(lo, fminf(a, fminf(c, d)));
        hi = fmaxf(hi, fmaxf(a, fmaxf(c, d)));
    }
    wave_minmax(lo, hi);
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma clang fp contract(off)
        for (int w = 1; w < kSegBlock / 64; ++w) {
            lo = fminf(lo, slo[w]);
            hi = fmaxf(hi, shi[w]);
        }
        float range = hi - lo;  // H:2008-2011
        if (range < 0.001f) range = 0.001f;
        prm[blockIdx.x] = make_float4(lo, hi, range, 1023.0f / range);
        const float half = range * 0.5f;  // README.md:104-126: root at min + range/2
        const float c = lo + half;
        root[blockIdx.x] = make_float4(c, c, c, half);
    }
}

// sys[i] = the system of particle i
__global__ __launch_bounds__(kSegBlock) void sys_fill_kernel(const uint32_t* __restrict__ off,
                                                             uint32_t* __restrict__ sys) {
    const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
    for (uint32_t i = b + threadIdx.x; i < e; i += kSegBlock) sys[i] = blockIdx.x;
}

// Morton codes with each system's own cube; the AVX2 mode covers the first
// 8 floor(N_s / 8) particles of each system, as a separate prepare call would
__global__ __launch_bounds__(kBlock) void morton_seg_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    size_t n, const uint32_t* __restrict__ sys, const uint32_t* __restrict__ off,
    const float4* __restrict__ prm, uint32_t* __restrict__ codes, int mode) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = sys[i];
    const float4 p = prm[s];
    const uint32_t b = off[s], e = off[s + 1];
    const uint32_t nvec = b + ((e - b) & ~7u);
    float vx = x[i], vy = y[i], vz = z[i];
    uint32_t qx, qy, qz;
    if (mode == 1 && i < nvec) {
        qx = q_rne(vx, p.x, p.w);
        qy = q_rne(vy, p.x, p.w);
        qz = q_rne(vz, p.x, p.w);
    } else {
        qx = q_trunc(vx, p.x, p.z);
        qy = q_trunc(vy, p.x, p.z);
        qz = q_trunc(vz, p.x, p.z);
    }
    codes[i] = ps_morton3(qx, qy, qz);
}

}  // namespace

void ps_launch_prepare_seg(const float* x, const float* y, const float* z, size_t n,
                           const uint32_t* off, uint32_t n_sys, float4* prm, float4* root,
                           uint32_t* sys, uint32_t* codes, int mode, hipStream_t s) {
    if (n == 0 || n_sys == 0) return;
    bbox_seg_kernel<<<n_sys, kSegBlock, 0, s>>>(x, y, z, off, prm, root);
    sys_fill_kernel<<<n_sys, kSegBlock, 0, s>>>(off, sys);
    morton_seg_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(x, y, z, n, sys,
                                                                              off, prm, codes,
                                                                              mode);
}

bool ps_launch_prepare_small(const float* x, const float* y, const float* z, size_t n,
                             float* prm, uint32_t* codes, int mode, hipStream_t s) {
    if (n == 0 || n > kSmallPrep) return false;
    prepare_small_kernel<<<1, kSmallBlock, 0, s>>>(x, y, z, (uint32_t)n, prm, codes, mode);
    return true;
}

void ps_launch_bbox(const float* x, const float* y, const float* z, size_t n, float2* part,
                    float* prm, hipStream_t s) {
    size_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > (size_t)kMaxParts) blocks = kMaxParts;
    if (blocks == 0) blocks = 1;
    bbox_partial_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(x, y, z, n, part);
    bbox_final_kernel<<<1, kBlock, 0, s>>>(part, (int)blocks, prm);
}

void ps_launch_morton(const float* x, const float* y, const float* z, size_t n,
                      const float* prm, uint32_t* codes, int mode, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + kBlock - 1) / kBlock;
    morton_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(x, y, z, n, prm, codes, mode,
                                                      n - (n % 8u));
}

// ---- the prepare pass without allocation ----------------------------------
// Its scratch (the bbox partials of at most kMaxParts blocks, then the four
// parameters min, max, range, 1023/range) lives in this file-scope device
// array: every device the library runs on holds one copy until the process
// ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol): each device's address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
static_assert(PS_PREP_SCRATCH_BYTES >= kMaxParts * sizeof(float2) + 4 * sizeof(float),
              "the no-alloc scratch holds the partials and the parameters");
static __device__ __align__(256) unsigned char scratch[PS_PREP_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

hipError_t ps_prepare_scratch(void** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t ps_prepare_noalloc(const float* x, const float* y, const float* z, size_t n,
                              uint32_t* codes, int mode, float* out3, hipStream_t s) {
    void* ptr = nullptr;
    hipError_t e = ps_prepare_scratch(&ptr);
    if (e != hipSuccess) return e;
    float2* part = reinterpret_cast<float2*>(ptr);  // bbox partials first
    float* prm = reinterpret_cast<float*>(reinterpret_cast<char*>(ptr) +
                                          kMaxParts * sizeof(float2));
    // every partial the final pass reads is written by this call's first
    // pass, and the parameters by its final pass: nothing needs clearing
    if (!ps_launch_prepare_small(x, y, z, n, prm, codes, mode, s)) {
        ps_launch_bbox(x, y, z, n, part, prm, s);
        ps_launch_morton(x, y, z, n, prm, codes, mode, s);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(out3, prm, 3 * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}
