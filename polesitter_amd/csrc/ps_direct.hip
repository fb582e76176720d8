// ps_direct.hip -- FP64 direct sum on the device, the error oracle of every
// reported tick (tests/accuracy.c:42-72 on a sample of targets), plus the
// (x, y, z, m) packing used by it and by the tick.
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// one block per target: lanes stride over all sources, FP64 throughout
__global__ __launch_bounds__(kBlock) void direct_f64_kernel(const float4* __restrict__ pos4,
                                                            size_t n,
                                                            const uint32_t* __restrict__ targets,
                                                            double* __restrict__ out3) {
#pragma clang fp contract(off)
    __shared__ double red[3][kWaves];
    const uint32_t i = targets[blockIdx.x];
    const float4 p = pos4[i];
    const double eps2 = (double)0.1f;  // SOFTENING_SQ is the float 0.1F
    const double px = p.x, py = p.y, pz = p.z, pm = p.w;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (size_t j = threadIdx.x; j < n; j += kBlock) {
        if (j == i) continue;
        const float4 q = pos4[j];
        double dx = (double)q.x - px;
        double dy = (double)q.y - py;
        double dz = (double)q.z - pz;
        double r2 = dx * dx + dy * dy + dz * dz + eps2;
        double inv = 1.0 / sqrt(r2);
        double inv3 = inv * inv * inv;
        double sc = pm * (double)q.w * inv3;
        ax += dx * sc;
        ay += dy * sc;
        az += dz * sc;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        ax += __shfl_down(ax, d, 64);
        ay += __shfl_down(ay, d, 64);
        az += __shfl_down(az, d, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = ax;
        red[1][w] = ay;
        red[2][w] = az;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sx = 0, sy = 0, sz = 0;
        for (int k = 0; k < kWaves; ++k) {
            sx += red[0][k];
            sy += red[1][k];
            sz += red[2][k];
        }
        out3[3 * (size_t)blockIdx.x + 0] = sx;
        out3[3 * (size_t)blockIdx.x + 1] = sy;
        out3[3 * (size_t)blockIdx.x + 2] = sz;
    }
}

__global__ void pack_pos4_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                 const float* __restrict__ z, const float* __restrict__ m,
                                 size_t n, float4* __restrict__ pos4) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) pos4[i] = make_float4(x[i], y[i], z[i], m[i]);
}

}  // namespace

void ps_launch_direct_f64(const float4* pos4, size_t n, const uint32_t* targets, size_t nt,
                          double* out3, hipStream_t s) {
    if (nt == 0) return;
    direct_f64_kernel<<<(unsigned)nt, kBlock, 0, s>>>(pos4, n, targets, out3);
}

void ps_launch_pack_pos4(const float* x, const float* y, const float* z, const float* m,
                         size_t n, float4* pos4, hipStream_t s) {
    if (n == 0) return;
    pack_pos4_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(x, y, z, m, n,
                                                                             pos4);
}
