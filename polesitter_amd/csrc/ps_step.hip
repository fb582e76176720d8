// ps_step.hip -- the caller's frame step around the tick, on the device
// (README.md:94-161): reset ids, gather the caller's external velocity
// arrays by the new sort order, semi-implicit Euler -- as separate steps or
// fused (gather + Euler, then copy back + reset ids: two kernels).
#include "ps_device.h"
#include "ps_kernels.h"

namespace {

constexpr int kBlock = 256;

__global__ void iota_kernel(uint32_t* id, size_t n) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) id[i] = (uint32_t)i;
}

// out[i] = v[id[i]]: the velocity that belonged to the particle now at i
__global__ void gather_kernel(const uint32_t* __restrict__ id, const float* __restrict__ vx,
                              const float* __restrict__ vy, const float* __restrict__ vz,
                              float* __restrict__ ox, float* __restrict__ oy,
                              float* __restrict__ oz, size_t n) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t k = id[i];
    ox[i] = vx[k];
    oy[i] = vy[k];
    oz[i] = vz[k];
}

// v += (f / m) * dt; x += v * dt, skipping massless particles (README.md:149-161)
__global__ void euler_kernel(float* __restrict__ x, float* __restrict__ y,
                             float* __restrict__ z, float* __restrict__ vx,
                             float* __restrict__ vy, float* __restrict__ vz,
                             const float* __restrict__ fx, const float* __restrict__ fy,
                             const float* __restrict__ fz, const float* __restrict__ m, float dt,
                             size_t n) {
#pragma clang fp contract(off)
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float mi = m[i];
    if (mi <= 0.0f) return;
    float a = vx[i] + (fx[i] / mi) * dt;
    float b = vy[i] + (fy[i] / mi) * dt;
    float c = vz[i] + (fz[i] / mi) * dt;
    vx[i] = a;
    vy[i] = b;
    vz[i] = c;
    x[i] = x[i] + a * dt;
    y[i] = y[i] + b * dt;
    z[i] = z[i] + c * dt;
}

// the frame step fused, first kernel: t[i] = v[id[i]] (the velocity of the
// particle now at i), then for m > 0: t += (f / m) dt, x += t dt
__global__ void frame_kernel(const uint32_t* __restrict__ id, const float* __restrict__ vx,
                             const float* __restrict__ vy, const float* __restrict__ vz,
                             float* __restrict__ tx, float* __restrict__ ty,
                             float* __restrict__ tz, float* __restrict__ x, float* __restrict__ y,
                             float* __restrict__ z, const float* __restrict__ fx,
                             const float* __restrict__ fy, const float* __restrict__ fz,
                             const float* __restrict__ m, float dt, size_t n) {
#pragma clang fp contract(off)
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = id[i];
    float a = vx[k], b = vy[k], c = vz[k];
    const float mi = m[i];
    if (mi > 0.0f) {
        a = a + (fx[i] / mi) * dt;
        b = b + (fy[i] / mi) * dt;
        c = c + (fz[i] / mi) * dt;
        x[i] = x[i] + a * dt;
        y[i] = y[i] + b * dt;
        z[i] = z[i] + c * dt;
    }
    tx[i] = a;
    ty[i] = b;
    tz[i] = c;
}

// second kernel: v = t, id[i] = i for the next frame
__global__ void frame_finish_kernel(const float* __restrict__ tx, const float* __restrict__ ty,
                                    const float* __restrict__ tz, float* __restrict__ vx,
                                    float* __restrict__ vy, float* __restrict__ vz,
                                    uint32_t* __restrict__ id, size_t n) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    vx[i] = tx[i];
    vy[i] = ty[i];
    vz[i] = tz[i];
    id[i] = (uint32_t)i;
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void ps_launch_iota(uint32_t* id, size_t n, hipStream_t s) {
    if (n) iota_kernel<<<blocks_for(n), kBlock, 0, s>>>(id, n);
}

void ps_launch_gather_vel(const uint32_t* id, const float* vx, const float* vy,
                          const float* vz, float* ox, float* oy, float* oz, size_t n,
                          hipStream_t s) {
    if (n) gather_kernel<<<blocks_for(n), kBlock, 0, s>>>(id, vx, vy, vz, ox, oy, oz, n);
}

void ps_launch_euler(float* x, float* y, float* z, float* vx, float* vy, float* vz,
                     const float* fx, const float* fy, const float* fz, const float* m,
                     float dt, size_t n, hipStream_t s) {
    if (n)
        euler_kernel<<<blocks_for(n), kBlock, 0, s>>>(x, y, z, vx, vy, vz, fx, fy, fz, m, dt,
                                                      n);
}

void ps_launch_frame(uint32_t* id, float* vx, float* vy, float* vz, float* tmp3, float* x,
                     float* y, float* z, const float* fx, const float* fy, const float* fz,
                     const float* m, float dt, size_t n, hipStream_t s) {
    if (!n) return;
    frame_kernel<<<blocks_for(n), kBlock, 0, s>>>(id, vx, vy, vz, tmp3, tmp3 + n, tmp3 + 2 * n, x,
                                                  y, z, fx, fy, fz, m, dt, n);
    frame_finish_kernel<<<blocks_for(n), kBlock, 0, s>>>(tmp3, tmp3 + n, tmp3 + 2 * n, vx, vy,
                                                         vz, id, n);
}
