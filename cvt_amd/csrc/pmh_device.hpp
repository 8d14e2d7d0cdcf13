// pmh_device.hpp -- device-side plane algebra, MWC64X and small helpers for the
// gfx950 PatchMatch stereo kernels.
//
// Every function states the reference code it follows (paths relative to the
// mherb/cvt tree).  The library is compiled with -ffp-contract=off: a fused
// multiply-add appears only where written as __builtin_fmaf, so the rounding of
// each expression is fixed by the source text.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmh {

constexpr int kViewSamples = 4;        // VIEWSAMPLES, pmhstereo.cl:40
constexpr int kPropDraws = 21;         // 6*2 + 3 + 2*3 draws of pmhstereo.cl:355-389
constexpr uint32_t kPropStride = 45;   // RNG_init stride, pmhstereo.cl:327
constexpr uint32_t kInitStride = 3;    // RNG_init stride, pmhstereo.cl:241
constexpr uint32_t kMVStride = 33;     // RNG_init stride, pmmvstereo.cl:237

constexpr uint32_t kMwcA = 4294883355u;               // RNG.cl:105
constexpr uint64_t kMwcM = 18446383549859758079ull;   // RNG.cl:106
constexpr uint64_t kMwcBase = 4077358422479273989ull; // RNG.cl:93
// 2^64 mod M = 2^64 - M: folds the high word of a 128-bit product.
constexpr uint64_t kMwcFold = 0x0000000000000000ull - kMwcM;

// Per-stride skip table: pow[j] = A^(stride * 2^j) mod M.
struct SkipTable {
    uint64_t pow[40];
};

struct ViewRec {            // one view-propagation slot record per target pixel
    int n;                  // scatters received (reference VIEWPROP_t.n)
    int src[kViewSamples];  // source x of the first min(n, 4), ascending
};

// ---------------------------------------------------------------- MWC64X ----
// a*b mod M for a, b < M, using 2^64 == kMwcFold (mod M).
__device__ __forceinline__ uint64_t mwc_mulmod(uint64_t a, uint64_t b) {
    uint64_t lo = a * b;
    uint64_t hi = __umul64hi(a, b);
    while (hi) {
        uint64_t plo = hi * kMwcFold;
        uint64_t phi = __umul64hi(hi, kMwcFold);
        uint64_t s = plo + lo;
        hi = phi + (s < lo ? 1ull : 0ull);
        lo = s;
    }
    return lo >= kMwcM ? lo - kMwcM : lo;
}

// RNG_init(linpos, stride) of RNG.cl:151-154: seed value BASEID, skipped by
// linpos*stride steps, i.e. BASEID * A^(stride*linpos) mod M.
__device__ __forceinline__ void mwc_seed(const SkipTable& t, uint64_t linpos, uint32_t& x,
                                         uint32_t& c) {
    uint64_t v = kMwcBase;
    for (int j = 0; j < 40; ++j) {
        if ((linpos >> j) == 0) break;
        if ((linpos >> j) & 1ull) v = mwc_mulmod(v, t.pow[j]);
    }
    x = (uint32_t)(v / kMwcA);
    c = (uint32_t)(v % kMwcA);
}

// MWC64X_NextFloat (RNG.cl:142-147) with MWC64X_Step (RNG.cl:108-118).
__device__ __forceinline__ float mwc_next_float(uint32_t& x, uint32_t& c) {
    uint32_t res = x ^ c;
    uint32_t xn = kMwcA * x + c;
    uint32_t carry = xn < c ? 1u : 0u;
    uint32_t cn = __umulhi(kMwcA, x) + carry;
    x = xn;
    c = cn;
    return 2.3283064365386962890625e-10f * (float)res;
}

// ---------------------------------------------------------- plane algebra ----
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    return fminf(fmaxf(v, lo), hi);
}

__device__ __forceinline__ bool finite3(const float4& s) {
    return isfinite(s.x) && isfinite(s.y) && isfinite(s.z);
}

// nd_state_viewprop, pmhstereo.cl:119-124
__device__ __forceinline__ float4 viewprop(const float4& s) {
    return make_float4(1.0f / s.x, -s.y / s.x, -s.z / s.x, 0.0f);
}

// nd_state_transform, pmhstereo.cl:126-129 (x component)
__device__ __forceinline__ float transform_x(const float4& s, float cx, float cy) {
    return __builtin_fmaf(s.x, cx, __builtin_fmaf(s.y, cy, s.z));
}

// nd_state_init, pmhstereo.cl:50-68 (u0: depth, u1: nx, u2: ny)
__device__ __forceinline__ float4 state_init(float u0, float u1, float u2, float cx, float cy,
                                             int lr, float normmul, float depthmax) {
    float z = u0 * depthmax;
    float nx = ((u1 - 0.5f) * normmul) * 0.95f;
    float ny = ((u2 - 0.5f) * normmul) * 0.95f;
    nx = clampf(nx, -0.95f, 0.95f);
    ny = clampf(ny, -0.95f, 0.95f);
    float nf = fmaxf(__builtin_sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nf;
    ny = ny / nf;
    float nz = __builtin_sqrtf((1.0f - nx * nx) - ny * ny);
    float4 r = make_float4(1.0f - (-nx / nz), 0.0f - (-ny / nz),
                           0.0f - ((nx * cx + ny * cy) / nz + z), 0.0f);
    return lr ? r : viewprop(r);
}

// nd_state_to_ref_normal_depth, pmhstereo.cl:70-86 -> (nx, ny, nz, depth)
__device__ __forceinline__ float4 to_ref_nd(const float4& s, float cx, float cy, int lr) {
    float4 t = lr ? s : viewprop(s);
    float t0 = 1.0f - t.x, t1 = 0.0f - t.y, t2 = 0.0f - t.z;
    float w = (t0 * cx + t1 * cy) + t2;
    float z = 1.0f / __builtin_sqrtf((t0 * t0 + t1 * t1) + 1.0f);
    float x = -t0 * z;
    float y = -t1 * z;
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) return make_float4(0.f, 0.f, 1.f, 0.f);
    return make_float4(x, y, z, w);
}

// nd_state_refine, pmhstereo.cl:88-117, right view returns
// viewprop((1,0,0,0) - viewprop(s)) as in :114-115.
__device__ __forceinline__ float4 refine(float u0, float u1, float u2, const float4& s, float cx,
                                         float cy, float depthmax, int lr) {
    float4 st = lr ? s : viewprop(s);
    st = make_float4(1.0f - st.x, 0.0f - st.y, 0.0f - st.z, 0.0f - st.w);
    if (!lr) return viewprop(st);
    float z = (st.x * cx + st.y * cy) + st.z;
    float nz = 1.0f / __builtin_sqrtf((st.x * st.x + st.y * st.y) + 1.0f);
    float nx = -st.x * nz;
    float ny = -st.y * nz;
    z += (u0 - 0.5f) * 2.0f;
    z = clampf(z, 0.0f, depthmax);
    nx += (u1 - 0.5f) * 0.1f;
    ny += (u2 - 0.5f) * 0.1f;
    nx = clampf(nx, -0.95f, 0.95f);
    ny = clampf(ny, -0.95f, 0.95f);
    float nf = fmaxf(__builtin_sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nf;
    ny = ny / nf;
    nz = __builtin_sqrtf((1.0f - nx * nx) - ny * ny);
    float r0 = -nx / nz, r1 = -ny / nz, r2 = (nx * cx + ny * cy) / nz + z;
    return make_float4(1.0f - r0, 0.0f - r1, 0.0f - r2, 0.0f);
}

// nd_state_to_normal, pmhstereo.cl:144-153
__device__ __forceinline__ float3 to_normal(const float4& s) {
    float t0 = 1.0f - s.x, t1 = 0.0f - s.y;
    float z = 1.0f / __builtin_sqrtf((t0 * t0 + t1 * t1) + 1.0f);
    return make_float3(-t0 * z, -t1 * z, z);
}

// smoothDistance, pmhstereo.cl:266-283: squared distance of one state to the
// smooth prior (the reference evaluates it for candidate and self together).
__device__ __forceinline__ float smooth_dist(const float4& s, const float4& sm, float cx, float cy,
                                             float dw, int lr) {
    float4 A = to_ref_nd(s, cx, cy, lr);
    float d0 = A.x - sm.x, d1 = A.y - sm.y, d2 = A.z - sm.z, d3 = A.w - sm.w;
    return ((d0 * (1.0f * d0) + d1 * (1.0f * d1)) + d2 * (0.0f * d2)) + d3 * (dw * d3);
}

// Acceptance test of pmhstereo.cl:340 (and :352,360,368,378,388):
// cand.w + theta*sd.x <= self.w + theta*sd.y, where sd is (0,0) unless both
// states are finite.
__device__ __forceinline__ bool accept(const float4& cand, const float4& self, const float4& sm,
                                       float cx, float cy, float dw, int lr, float theta) {
    float sx = 0.0f, sy = 0.0f;
    if (finite3(cand) && finite3(self)) {
        sx = smooth_dist(cand, sm, cx, cy, dw, lr);
        sy = smooth_dist(self, sm, cx, cy, dw, lr);
    }
    return cand.w + theta * sx <= self.w + theta * sy;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float len4(const float4& v) {
    return __builtin_sqrtf(((v.x * v.x + v.y * v.y) + v.z * v.z) + v.w * v.w);
}
__device__ __forceinline__ float len3(const float4& v) {
    return __builtin_sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
}

}  // namespace pmh
