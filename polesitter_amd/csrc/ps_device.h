// ps_device.h -- device-side definitions shared by the gfx950 kernels.
//
// Node layout in HBM (structure of arrays, canonical level order: depth-major,
// ascending Morton prefix inside a depth, so the children of one node are
// contiguous and in octant order):
//   geom[n]   float4  (x, y, z, half_width)          ps_node_t x..half_width, H:364-367
//   mp[n]     float4  (m0, px, py, pz)               ps_node_t multipole,     H:370
//   loc[n]    float4  (unused, Ex, Ey, Ez)           ctx->local_exp,          H:496
//   link[n]   uint2   internal: (first child, slot mask)
//                     leaf:     (first particle, PS_LEAF)                     H:371-381
//   nend[n]   uint32  one past the node's last particle
//   parent[n] uint32
// Particles after the sort: the caller's SoA arrays plus pos4[i] =
// (x, y, z, mass) and leaf_of[i] (the leaf that holds particle i).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PS_MAX_DEPTH 10            // H:882
#define PS_LEAF      0x80000000u   // link.y flag of a leaf
#define PS_NONE      0xFFFFFFFFu
#define PS_SOFT2     0.1f          // softening eps^2, H:1065,1302,1433

// ---------------------------------------------------------------------------
// Morton code, H:777-799: bit k of each 10-bit axis goes to bit 3k (+0 x,
// +1 y, +2 z).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ps_spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t ps_morton3(uint32_t x, uint32_t y, uint32_t z) {
    return ps_spread10(x) | (ps_spread10(y) << 1) | (ps_spread10(z) << 2);
}

// ---------------------------------------------------------------------------
// MAC, H:1056-1064 and H:1265-1273.  The comparison decides discrete pair
// sets, so the float operation sequence must match the reference exactly:
// no fused multiply-add, left-to-right sums.  dx is source minus target.
// theta2 is fl(theta * theta), as the reference evaluates it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool ps_far(float4 t, float4 s, float theta2) {
#pragma clang fp contract(off)
    float dx = s.x - t.x;
    float dy = s.y - t.y;
    float dz = s.z - t.z;
    float d2 = (dx * dx + dy * dy) + dz * dz;
    float hs = t.w + s.w;
    return d2 > theta2 * (hs * hs);
}

// ---------------------------------------------------------------------------
// M2L, H:1065-1092: first-order (monopole + dipole) softened field of the
// source cell at the target centre,
//   E += m0 d / r^3 + p / r^3 - 3 (p.d) d / r^5,   r^2 = |d|^2 + eps^2,
// with d = source - target and d2 = |d|^2 exactly as the MAC computed it.
// The reference's sqrtf and two divisions become one v_rsq_f32 (1 ulp) and
// products: the field is compared within a tolerance, never bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ps_m2l_d(float dx, float dy, float dz, float d2, float4 smp,
                                         float& ex, float& ey, float& ez) {
    const float ir = __builtin_amdgcn_rsqf(d2 + PS_SOFT2);
    const float ir2 = ir * ir;
    const float ir3 = ir2 * ir;
    const float ir5 = ir3 * ir2;
    const float mdr = smp.y * dx + smp.z * dy + smp.w * dz;
    const float a = smp.x * ir3 - 3.0f * mdr * ir5;  // coefficient of d
    ex += a * dx + smp.y * ir3;
    ey += a * dy + smp.z * ir3;
    ez += a * dz + smp.w * ir3;
}

// ps_m2l_d for every lane, adding only where `on`: elsewhere the inverse
// distance is 0, so each term added is +-0.  An accumulator that starts at +0
// never holds -0 (+0 + -0 = +0 in round-to-nearest), so adding +-0 leaves it
// unchanged bit for bit: the same sums as calling ps_m2l_d where `on` only.
__device__ __forceinline__ void ps_m2l_dm(float dx, float dy, float dz, float d2, float4 smp,
                                          bool on, float& ex, float& ey, float& ez) {
    const float r = __builtin_amdgcn_rsqf(d2 + PS_SOFT2);
    const float ir = on ? r : 0.0f;
    const float ir2 = ir * ir;
    const float ir3 = ir2 * ir;
    const float ir5 = ir3 * ir2;
    const float mdr = smp.y * dx + smp.z * dy + smp.w * dz;
    const float a = smp.x * ir3 - 3.0f * mdr * ir5;
    ex += a * dx + smp.y * ir3;
    ey += a * dy + smp.z * ir3;
    ez += a * dz + smp.w * ir3;
}

__device__ __forceinline__ void ps_m2l(float4 t, float4 s, float4 smp, float& ex, float& ey,
                                       float& ez) {
    float dx, dy, dz, d2;
    {
#pragma clang fp contract(off)
        dx = s.x - t.x;
        dy = s.y - t.y;
        dz = s.z - t.z;
        d2 = (dx * dx + dy * dy) + dz * dz;
    }
    ps_m2l_d(dx, dy, dz, d2, smp, ex, ey, ez);
}

// MAC that also returns the separation it evaluated (for M2L on far pairs)
__device__ __forceinline__ bool ps_far_d(float4 t, float4 s, float theta2, float& dx, float& dy,
                                         float& dz, float& d2) {
#pragma clang fp contract(off)
    dx = s.x - t.x;
    dy = s.y - t.y;
    dz = s.z - t.z;
    d2 = (dx * dx + dy * dy) + dz * dz;
    const float hs = t.w + s.w;
    return d2 > theta2 * (hs * hs);
}

// inclusive prefix sum over the wavefront with DPP lane moves (no LDS
// round trip): row shifts by 1, 2, 4, 8 give each row of 16 lanes its own
// prefix (lanes shifted in from outside the row read 0), then the last lane
// of row 0 (row 2) is broadcast into row 1 (row 3), and lane 31 into rows 2
// and 3 (rows outside a broadcast's row mask take 0)
__device__ __forceinline__ uint32_t ps_wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// exclusive prefix over the wave; *total gets the wave-wide sum
__device__ __forceinline__ uint32_t ps_wave_excl_scan(uint32_t v, uint32_t* total) {
    const uint32_t inc = ps_wave_incl_scan(v);
    *total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    return inc - v;
}

// merge two compensated sums (s, c): value = s - c
__device__ __forceinline__ void ps_kahan_merge(float& s, float& c, float s2, float c2) {
#pragma clang fp contract(off)
    const float t = s + s2;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (s2 - bp);  // exact: t + err == s + s2
    s = t;
    c = c + c2 - err;
}

// ---------------------------------------------------------------------------
// One softened particle pair of the P2P inner loop, H:1427-1443:
// a += m_j * d / (|d|^2 + eps^2)^(3/2), d = source - target.  v_rsq_f32 is
// the FP32 reciprocal square root (within 1 ulp); contraction is allowed
// here: the pair sum is compared within a tolerance, never bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ps_pair(float4 t, float4 s, float& ax, float& ay, float& az) {
    float dx = s.x - t.x;
    float dy = s.y - t.y;
    float dz = s.z - t.z;
    float d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, PS_SOFT2)));
    float inv = __builtin_amdgcn_rsqf(d2);
    float f = s.w * (inv * inv * inv);
    ax = fmaf(dx, f, ax);
    ay = fmaf(dy, f, ay);
    az = fmaf(dz, f, az);
}

// Kahan-compensated accumulation of one term.
__device__ __forceinline__ void ps_kahan_add(float& sum, float& comp, float term) {
#pragma clang fp contract(off)
    float y = term - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// ---------------------------------------------------------------------------
// The traversal's node record (written by the upward pass): the centre, and
// in the fourth word
//   bits 0..4  -depth, 5-bit two's complement (the half width is the root's
//              times 2^-depth, exactly: each level halves it)
//   bits 5..8  the node's list-record bits: internal flag, id count - 1 (a
//              leaf stands for itself, an internal node for its children)
//   bits 9..   what the node adds to a target's next list when near: its id
//              count | internal << 16
// ---------------------------------------------------------------------------
constexpr uint32_t PS_TNODE_OUT_SHIFT = 9;
__device__ __forceinline__ uint32_t ps_tnode_word(int depth, bool leaf, uint32_t nk) {
    const uint32_t intl = leaf ? 0u : 1u;
    const uint32_t cnt = leaf ? 1u : nk;
    return ((uint32_t)(-depth) & 0x1Fu) | (intl << 5) | ((cnt - 1u) << 6) |
           ((cnt | (intl << 16)) << PS_TNODE_OUT_SHIFT);
}
// -depth
__device__ __forceinline__ int ps_tnode_nexp(uint32_t w) { return ((int)(w << 27)) >> 27; }
// bits 28..31 of a list record: internal flag, id count - 1
__device__ __forceinline__ uint32_t ps_tnode_rec_bits(uint32_t w) { return (w & 0x1E0u) << 23; }

__device__ __forceinline__ uint32_t ps_child_count(uint2 link) {
    return (link.y & PS_LEAF) ? 0u : (uint32_t)__popc(link.y & 0xFFu);
}

__device__ __forceinline__ uint64_t ps_lanemask_lt() {
    return (__lane_id() == 0) ? 0ull : ((~0ull) >> (64 - __lane_id()));
}

// ---------------------------------------------------------------------------
// Lane groups inside a wavefront.
// ---------------------------------------------------------------------------
// DPP lane exchanges inside a row of 16 lanes (one VALU op, no LDS round
// trip): quad_perm [1,0,3,2] and [2,3,0,1] are xor 1 and xor 2; the half-row
// and row mirrors pair lane i with 7 - i and 15 - i.  In a butterfly sum taken
// in the order 1, 2, 4 (8) every lane of a 4 (8) lane block holds the same
// value after the earlier steps, so the mirror partner gives exactly what
// xor 4 (xor 8) would.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __builtin_bit_cast(float, dpp_u32<CTRL>(__builtin_bit_cast(uint32_t, v)));
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
constexpr int kDppShr1 = 0x111, kDppShr2 = 0x112, kDppShr4 = 0x114, kDppShr8 = 0x118;

// Inclusive prefix sums over groups of 32 or 64 lanes with DPP (see
// ps_wave_incl_scan): rows of 16 by shifts, row 0's (row 2's) last lane into
// row 1 (row 3), and for 64 lanes lane 31 into rows 2 and 3.  The group's
// total is its last lane's, read into a scalar.
template <int G>
__device__ __forceinline__ uint32_t wide_incl_scan_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppShr1, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppShr2, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppShr4, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppShr8, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    if (G == 64)
        v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}
// one DPP step of a float scan: x moved by CTRL within the rows ROWS (0 in
// lanes that take nothing)
template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ float dpp_step_f32(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, ROWS, 0xF, BOUND));
}
template <int G>
__device__ __forceinline__ float wide_incl_scan_f32(float v) {
    v += dpp_step_f32<kDppShr1, 0xF, true>(v);
    v += dpp_step_f32<kDppShr2, 0xF, true>(v);
    v += dpp_step_f32<kDppShr4, 0xF, true>(v);
    v += dpp_step_f32<kDppShr8, 0xF, true>(v);
    v += dpp_step_f32<0x142, 0xA, false>(v);  // row_bcast:15
    if (G == 64) v += dpp_step_f32<0x143, 0xC, false>(v);  // row_bcast:31
    return v;
}
// the last lane's value of each group of 32 or 64 lanes, in all its lanes
template <int G>
__device__ __forceinline__ uint32_t wide_last_u32(uint32_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
    if (G == 64) return hi;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    return __lane_id() < 32 ? lo : hi;
}

// sum over the G lanes of one group (G a power of two <= 64), all with DPP
// lane moves; the same value in every lane
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    if (G <= 16) {
        v += dpp_f32<kDppXor1>(v);
        v += dpp_f32<kDppXor2>(v);
        v += dpp_f32<kDppHalfMirror>(v);
        if (G == 16) v += dpp_f32<kDppMirror>(v);
        return v;
    }
    return __builtin_bit_cast(float,
                              wide_last_u32<G>(__builtin_bit_cast(uint32_t, wide_incl_scan_f32<G>(v))));
}

// inclusive prefix sum over the lanes of one group (gl: lane in the group)
// and the group's total, valid in every lane
template <int G>
__device__ __forceinline__ uint32_t group_scan_u32(uint32_t v, uint32_t gl, uint32_t& total) {
    uint32_t inc = v;
    if (G <= 16) {
        // row shifts cross into the neighbouring group for the first lanes:
        // those take nothing
        uint32_t t = dpp_u32<kDppShr1>(inc);
        inc += gl >= 1u ? t : 0u;
        t = dpp_u32<kDppShr2>(inc);
        inc += gl >= 2u ? t : 0u;
        t = dpp_u32<kDppShr4>(inc);
        inc += gl >= 4u ? t : 0u;
        if (G == 16) {
            t = dpp_u32<kDppShr8>(inc);
            inc += gl >= 8u ? t : 0u;
        }
        uint32_t s = v;
        s += dpp_u32<kDppXor1>(s);
        s += dpp_u32<kDppXor2>(s);
        s += dpp_u32<kDppHalfMirror>(s);
        if (G == 16) s += dpp_u32<kDppMirror>(s);
        total = s;
        return inc;
    }
    inc = wide_incl_scan_u32<G>(inc);
    total = wide_last_u32<G>(inc);
    return inc;
}

template <int G>
__device__ __forceinline__ unsigned long long group_bits(bool pred, uint32_t gshift) {
    const unsigned long long b = __ballot(pred) >> gshift;
    return G == 64 ? b : (b & ((1ull << G) - 1ull));
}
// the same from a wavefront mask already at hand
template <int G>
__device__ __forceinline__ unsigned long long group_bits_of(unsigned long long m, uint32_t gshift) {
    const unsigned long long b = m >> gshift;
    return G == 64 ? b : (b & ((1ull << G) - 1ull));
}

template <int G>
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t v) {
    if (G <= 16) {
        v += dpp_u32<kDppXor1>(v);
        v += dpp_u32<kDppXor2>(v);
        v += dpp_u32<kDppHalfMirror>(v);
        if (G == 16) v += dpp_u32<kDppMirror>(v);
        return v;
    }
    return wide_last_u32<G>(wide_incl_scan_u32<G>(v));
}

// orders the LDS accesses of one wavefront (its lanes run in step: no
// hardware barrier is needed, only that the compiler keeps the order)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
