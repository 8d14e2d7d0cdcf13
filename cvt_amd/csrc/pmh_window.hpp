// pmh_window.hpp -- device code of the window cost shared by the forward
// kernels (pmh_kernels.hip) and the adjoint kernels (pmh_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "pmh_device.hpp"
#include "pmh_internal.hpp"

namespace pmh {

// ------------------------------------------------------------------ helpers --
// Block index -> tile index so that the blocks dispatched to one XCD (round
// robin over 8) cover one contiguous band of tiles and share its L2 lines.
__device__ __forceinline__ int xcd_tile(int nb) {
    const int per = (nb + 7) >> 3;
    const int b = blockIdx.x;
    return (b & 7) * per + (b >> 3);
}
static inline int xcd_grid(int nb) { return ((nb + 7) >> 3) << 3; }

// Flattened work list over a batch of items (pairs or pair views): tile t
// belongs to item i with prefix[i] <= t < prefix[i + 1].  A block takes
// `chunk` consecutive tiles (chunk > 1: a persistent grid smaller than the
// work); blocks are renumbered so that the blocks of one XCD take one
// contiguous range of tiles.  The grid is a multiple of 8.
__device__ __forceinline__ int item_of(const Tiles& T, int t) {
    int lo = 0, hi = T.items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T.prefix[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

//
// Persistent mode (chunk 0): the grid holds as many blocks as the device runs
// at once; the work list is cut into 8 bands, one per XCD, and a block takes
// the next tile of its XCD's band (one atomic per tile on T.queue[band]),
// then of the other bands once its own is done.  The last block to finish
// clears the counters, so the next launch (or graph replay) finds them at
// zero without a memset.
//
// The tile body f is called from one place for both modes: a body called
// from two is too big to inline twice and becomes a function call (its
// arguments on the stack, registers saved around it).
template <class F>
__device__ __forceinline__ void for_tiles(const Tiles& T, F&& f) {
    __shared__ int next_tile;
    const bool persistent = T.chunk <= 0;  // uniform over the grid
    const int home = (int)blockIdx.x & 7;
    const int per_band = (T.total + 7) >> 3;
    const int lb = home * ((int)gridDim.x >> 3) + ((int)blockIdx.x >> 3);
    int t = persistent ? 0 : lb * T.chunk;
    const int t_end = persistent ? 0 : min(t + T.chunk, T.total);
    for (;;) {
        int cur;
        if (persistent) {
            if (threadIdx.x == 0) {
                int got_t = -1;
                for (int k = 0; k < 8 && got_t < 0; ++k) {
                    const int b = (home + k) & 7;
                    const int n = min(per_band, T.total - b * per_band);
                    if (n <= 0) continue;
                    const int got = atomicAdd(&T.queue[b], 1);
                    if (got < n) got_t = b * per_band + got;
                }
                next_tile = got_t;
            }
            __syncthreads();
            cur = __builtin_amdgcn_readfirstlane(next_tile);
            __syncthreads();  // every thread has read next_tile before it is written again
        } else {
            cur = t < t_end ? t : -1;
            ++t;
        }
        if (cur < 0) break;
        const int tt = cur * T.stride;
        const int i = __builtin_amdgcn_readfirstlane(item_of(T, tt));
        f(i, __builtin_amdgcn_readfirstlane(tt - T.prefix[i]));
    }
    if (persistent && threadIdx.x == 0) {
        // every other block has stopped taking tiles once this count is reached
        if (atomicAdd(&T.queue[8], 1) == (int)gridDim.x - 1)
            for (int b = 0; b < kTileQueueWords; ++b) atomicExch(&T.queue[b], 0);
    }
}

// ----------------------------------------------------------- window cost ----
// Window cost of one plane for the pixel (x, y): patch_eval_color_grad_weighted
// (pmhstereo.cl:170-228) for gray inputs, summed in the reference's dy-outer,
// dx-inner order; taps outside the image or whose x' leaves [0, W) add exact
// zeros.
//
// Execution shape: one wave owns one row segment of pixels (64 lanes, every
// pixel or every other pixel), so the window row qy is wave-uniform.
// plane_cost sweeps the window for one plane; plane_costs sweeps it once for
// several planes of the same pixels (the four-wave propagate kernels cost a
// wave's independent candidates that way).  The
// bilinear gathers are the bandwidth problem: per window row the wave first
// copies the span of the other view's row that its 64 lanes x (2p+1) taps can
// reach into LDS with coalesced loads (each texel once instead of ~2p+1
// times), then gathers from LDS.  Reads of the pixel's own view stay on the
// L1 path, so L1 and LDS share the traffic.  A span wider than the stage
// falls back to gathers through the L1 (random planes with steep slopes).
// Global reads use buffer resources with 32-bit byte offsets into three
// planes (G, dG, (I, dI)); (I, dI) holds one channel for gray inputs and
// RGBA for colour inputs.
constexpr int kWave = 64;
constexpr int kSplitWavesStage = 4;  // waves of a propagate block that share a stage

// Lane of the calling thread within its wave.
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Orders one wave's LDS writes before its later reads (and reads before later
// writes).  A wave's LDS operations are performed in program order, so this
// only has to keep the compiler from moving them across; a block of several
// waves that stage independently uses it instead of a block barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}

// 16-, 8- and 4-byte loads through a buffer resource (offsets in bytes).  The
// whole vector is bit-cast at once: subscripting the builtin's result element
// by element lowers to a single-dword load with the other lanes undefined.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 load16(__amdgpu_buffer_rsrc_t rs, unsigned off) {
    const f32x4 v = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 load8(__amdgpu_buffer_rsrc_t rs, unsigned off) {
    const f32x2 v = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float load4(__amdgpu_buffer_rsrc_t rs, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}
// The same with a wave-uniform byte offset `soff` added (kept out of the VALU).
__device__ __forceinline__ float4 load16s(__amdgpu_buffer_rsrc_t rs, unsigned off, unsigned soff) {
    const f32x4 v =
        __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, (int)soff, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float load4s(__amdgpu_buffer_rsrc_t rs, unsigned off, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, (int)soff, 0));
}

// Colour channels of an input image.  NC = 1: gray (the reference's
// CL_INTENSITY read returns (I, I, I, I), so its RGB sums are 3 x one
// channel); NC = 4: RGBA, sums over R, G, B left to right.  V is one texel's
// colour, the gather planes hold (I, I(x+1) - I(x)) per texel.
template <int NC>
struct Col;
template <>
struct Col<1> {
    using V = float;
    static constexpr unsigned kIdBytes = 8;  // (I, dI)
    static __device__ __forceinline__ V load(__amdgpu_buffer_rsrc_t rs, unsigned texel) {
        return load4(rs, texel * 4u);
    }
    static __device__ __forceinline__ V load_s(__amdgpu_buffer_rsrc_t rs, unsigned off,
                                               unsigned soff) {
        return load4s(rs, off, soff);
    }
    // |c0 - v1|; the reference's RGB sum is 3 |c0 - v1|, folded into the
    // gray weight exponent
    static __device__ __forceinline__ float dsum(V c0, V v1) { return fabsf(c0 - v1); }
    // 0.05 sum_rgb min(|v1 - (t0 + fa dt)|, 0.04), the RGB sum of equal
    // channels folded into the factor (0.15 = 3 x 0.05)
    static __device__ __forceinline__ float cterm(V v1, V t0, V dt, float fa) {
        return 0.15f * fminf(fabsf(v1 - __builtin_fmaf(fa, dt, t0)), 0.04f);
    }
};
template <>
struct Col<4> {
    using V = float4;
    static constexpr unsigned kIdBytes = 32;  // (I rgba, dI rgba)
    static __device__ __forceinline__ V load(__amdgpu_buffer_rsrc_t rs, unsigned texel) {
        return load16(rs, texel * 16u);
    }
    static __device__ __forceinline__ V load_s(__amdgpu_buffer_rsrc_t rs, unsigned off,
                                               unsigned soff) {
        return load16s(rs, off, soff);
    }
    // sum_rgb |c0 - v1|
    static __device__ __forceinline__ float dsum(V c0, V v1) {
        return (fabsf(c0.x - v1.x) + fabsf(c0.y - v1.y)) + fabsf(c0.z - v1.z);
    }
    static __device__ __forceinline__ float cterm(V v1, V t0, V dt, float fa) {
        const float r = fminf(fabsf(v1.x - __builtin_fmaf(fa, dt.x, t0.x)), 0.04f);
        const float g = fminf(fabsf(v1.y - __builtin_fmaf(fa, dt.y, t0.y)), 0.04f);
        const float b = fminf(fabsf(v1.z - __builtin_fmaf(fa, dt.z, t0.z)), 0.04f);
        return 0.05f * ((r + g) + b);
    }
};

// The support-factor table is written by the host and never by a kernel:
// read through the constant address space, its wave-uniform entries become
// scalar loads instead of vector loads that the tap loop would wait on.
using CFloat = const __attribute__((address_space(4))) float;
__device__ __forceinline__ CFloat* const_table(const float* p) { return (CFloat*)p; }

// Resources of one cost evaluation: this view (I1, G1), the other view's
// gather planes and the support-factor table.
struct CostCtx {
    __amdgpu_buffer_rsrc_t i1, g1, rec;
    unsigned dg_base, id_base;
    CFloat* F;    // support factors F(dx, dy)
    CFloat* Fw;   // weight exponents of this colour format (w = exp2(dsum * Fw))
    int W, H, P;
};

template <int NC>
__device__ __forceinline__ CostCtx cost_ctx(const float* I1, const float4* G1, const Gather& g2,
                                            const float* F, int W, int H, int P) {
    CostCtx c;
    const unsigned n = (unsigned)W * (unsigned)H;
    c.i1 = make_rsrc(I1, n * 4u * NC);
    c.g1 = make_rsrc(G1, n * 16u);
    c.rec = make_rsrc(g2.rec, g2.n * (32u + Col<NC>::kIdBytes));
    c.dg_base = g2.n * 16u;
    c.id_base = g2.n * 32u;
    c.F = const_table(F);
    c.Fw = c.F + (NC == kGray ? 1 : 2) * (2 * P + 1) * (2 * P + 1);
    c.W = W;
    c.H = H;
    c.P = P;
    return c;
}

// Gather-plane reads of texel tx (global path).
template <int NC>
__device__ __forceinline__ void gather_texel(const CostCtx& cc, unsigned tx, float4& tg,
                                             float4& tdg, typename Col<NC>::V& t0,
                                             typename Col<NC>::V& dt) {
    tg = load16(cc.rec, tx * 16u);
    tdg = load16(cc.rec, tx * 16u + cc.dg_base);
    if constexpr (NC == 1) {
        const float2 v = load8(cc.rec, tx * 8u + cc.id_base);
        t0 = v.x;
        dt = v.y;
    } else {
        t0 = load16(cc.rec, tx * 32u + cc.id_base);
        dt = load16(cc.rec, tx * 32u + cc.id_base + 16u);
    }
}

// LDS copy of the other view's texels [s0, s0 + span] of one row (span <= kS;
// one texel past the span, clamped to the row, so that every staged cell has
// its right neighbour).  Only G and the colour are staged, 20 B per texel gray
// and 32 B RGBA; a cell's differences dG = G(x+1) - G(x) and dI = I(x+1) - I(x)
// are taken from the neighbour in LDS, the same subtraction that writes the
// gather planes, so the same bits.  The stage width trades LDS per wave
// (occupancy) against how often a steep plane's span overflows to the L1 path.
template <int kS, int NC>
struct GatherStageT {
    using V = typename Col<NC>::V;
    V c[kS + 1];
    float4 g[kS + 1];
    using LdsF4 = const __attribute__((address_space(3))) f32x4;
    using LdsF1 = const __attribute__((address_space(3))) float;
    static __device__ __forceinline__ unsigned lds_addr(const void* p) {
        return (unsigned)(size_t)(const __attribute__((address_space(3))) void*)p;
    }
    // Cell of texel s0 + j (0 <= j < kS).
    __device__ __forceinline__ void get(int j, float4& tg, float4& tdg, V& t0, V& dt) const {
        cell(lds_addr(g) + ((unsigned)j << 4), lds_addr(c) + (unsigned)j * (unsigned)sizeof(V),
             tg, tdg, t0, dt);
    }
    // LDS addresses of texel 0 of a row staged from s0 (wave-uniform): a tap
    // at texel i0 then costs one shift-add per array.  Computed once per row
    // (readfirstlane keeps them in scalar registers; folded into the
    // instruction offset instead, they cost a subtraction per tap).
    struct RowBase {
        unsigned g, c;
    };
    __device__ __forceinline__ RowBase row_base(int s0) const {
        const int gb = __builtin_amdgcn_readfirstlane((int)(lds_addr(g) - (unsigned)s0 * 16u));
        const int cb = __builtin_amdgcn_readfirstlane(
            (int)(lds_addr(c) - (unsigned)s0 * (unsigned)sizeof(V)));
        return {(unsigned)gb, (unsigned)cb};
    }
    static __device__ __forceinline__ void get_texel(const RowBase& rb, int i0, float4& tg,
                                                     float4& tdg, V& t0, V& dt) {
        cell(rb.g + ((unsigned)i0 << 4), rb.c + (unsigned)i0 * (unsigned)sizeof(V), tg, tdg, t0,
             dt);
    }
    static __device__ __forceinline__ void cell(unsigned ga, unsigned ca, float4& tg, float4& tdg,
                                                V& t0, V& dt) {
        LdsF4* gp = (LdsF4*)(size_t)ga;
        const f32x4 ga4 = gp[0], gb4 = gp[1];
        const float4 a = make_float4(ga4.x, ga4.y, ga4.z, ga4.w);
        const float4 b = make_float4(gb4.x, gb4.y, gb4.z, gb4.w);
        tg = a;
        tdg = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, b.w - a.w);
        V u, v;
        if constexpr (NC == 1) {
            LdsF1* cp = (LdsF1*)(size_t)ca;
            u = cp[0];
            v = cp[1];
        } else {
            LdsF4* cp = (LdsF4*)(size_t)ca;
            const f32x4 u4 = cp[0], v4 = cp[1];
            u = make_float4(u4.x, u4.y, u4.z, u4.w);
            v = make_float4(v4.x, v4.y, v4.z, v4.w);
        }
        t0 = u;
        if constexpr (NC == 1) {
            dt = v - u;
        } else {
            dt = make_float4(v.x - u.x, v.y - u.y, v.z - u.z, v.w - u.w);
        }
    }
};

// Minimum / maximum over the 64 lanes of a wave (all lanes active), returned
// wave-uniform.  DPP scan of GFX9: row_shr 1, 2, 4, 8 within each 16-lane
// row, then row_bcast:15 and row_bcast:31 carry the rows' results upward, so
// lane 63 holds the whole wave's; lanes whose DPP source lies outside the row
// keep their value (the identity's role).  Six VALU operations and a
// readlane, where shuffles would be six dependent LDS permutes.
__device__ __forceinline__ int wave_min_i(int v) {
    const int I = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max_i(int v) {
    const int I = (int)0x80000000;
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// Matching cost of one tap: this view's colour v1 and gradients g1 against the
// other view's bilinear cell (G, dG, I, dI) at fraction fa.
template <int NC>
__device__ __forceinline__ float tap_cost(typename Col<NC>::V v1, const float4 g1, float fa,
                                          const float4 tg, const float4 tdg,
                                          typename Col<NC>::V t0, typename Col<NC>::V dt) {
    const float cterm = Col<NC>::cterm(v1, t0, dt, fa);
    const float m0 = fminf(fabsf(g1.x - __builtin_fmaf(fa, tdg.x, tg.x)), 0.01f);
    const float m1 = fminf(fabsf(g1.y - __builtin_fmaf(fa, tdg.y, tg.y)), 0.01f);
    const float m2 = fminf(fabsf(g1.z - __builtin_fmaf(fa, tdg.z, tg.z)), 0.01f);
    const float m3 = fminf(fabsf(g1.w - __builtin_fmaf(fa, tdg.w, tg.w)), 0.01f);
    const float gs = ((m0 + m1) + m2) + m3;
    return __builtin_fmaf(0.95f, gs, cterm);
}

// The taps of one window row.  FETCH(tx_in_row, valid, G, dG, I, dI) returns
// the other view's bilinear cell at texel tx of the row.  Taps of a lane
// without a pixel (active = false) never count, so their texels need not lie
// in the staged span.
template <int NC, class FETCH>
__device__ __forceinline__ void row_taps(const CostCtx& cc, int x, bool active, int P,
                                         unsigned rowo, float rb, float a,
                                         typename Col<NC>::V c0,
                                         CFloat* Fwr, float park, float& ws, float& rt,
                                         FETCH fetch) {
    using V = typename Col<NC>::V;
    const int W = cc.W;
    const float Wf = (float)W;
#pragma unroll 5
    for (int dx = -P; dx <= P; ++dx) {
        const int qx = x + dx;
        const bool inq = (unsigned)qx < (unsigned)W;
        const unsigned q = rowo + (unsigned)(inq ? qx : x);
        const V v1 = Col<NC>::load(cc.i1, q);
        const float4 g1 = load16(cc.g1, q * 16u);
        const float xp = __builtin_fmaf(a, (float)qx, rb);
        const bool valid = active && (xp >= 0.0f) && (xp < Wf);
        const float xs = valid ? xp : park;
        const float fa = __builtin_amdgcn_fractf(xs);
        float4 tg, tdg;
        V t0, dt;
        fetch((int)xs, valid, tg, tdg, t0, dt);
        const float wt = __builtin_amdgcn_exp2f(Col<NC>::dsum(c0, v1) * Fwr[dx]);
        const float wv = (valid && inq) ? wt : 0.0f;
        const float C = tap_cost<NC>(v1, g1, fa, tg, tdg, t0, dt);
        ws += wv;
        rt = __builtin_fmaf(wv, C, rt);
    }
}

// row_taps for an interior segment: all 64 lanes have a pixel whose window
// lies inside the image row, so inq is true for every tap and this view's
// addresses are a per-lane base plus a wave-uniform offset.  kWhole: the
// plane also maps every lane's whole window row into [0, W), so every tap is
// valid and the compares and selects drop out; otherwise a tap outside [0, W)
// reads texel `park` and does not count.  Same arithmetic in the same order as
// row_taps for counted taps, so the same bits.
template <int NC, bool kWhole, class FETCH>
__device__ __forceinline__ void row_taps_inner(const CostCtx& cc, int x, int P, unsigned rowo,
                                               float rb, float a, typename Col<NC>::V c0,
                                               CFloat* Fwr, float park, float& ws, float& rt,
                                               FETCH fetch) {
    using V = typename Col<NC>::V;
    const unsigned q0 = rowo + (unsigned)(x - P);  // the lane's first tap
    const unsigned vi = q0 * (4u * NC), vg = q0 * 16u;
    const float Wf = (float)cc.W;
    float qf = (float)(x - P);  // (float)qx, stepped exactly (|qx| < 2^24)
#pragma unroll 5
    for (int k = 0; k <= 2 * P; ++k, qf += 1.0f) {
        const V v1 = Col<NC>::load_s(cc.i1, vi, (unsigned)k * (4u * NC));
        const float4 g1 = load16s(cc.g1, vg, (unsigned)k * 16u);
        const float xp = __builtin_fmaf(a, qf, rb);
        const bool valid = kWhole || ((xp >= 0.0f) && (xp < Wf));
        const float xs = valid ? xp : park;
        const float fa = __builtin_amdgcn_fractf(xs);
        float4 tg, tdg;
        V t0, dt;
        fetch((int)xs, tg, tdg, t0, dt);
        const float wt = __builtin_amdgcn_exp2f(Col<NC>::dsum(c0, v1) * Fwr[k - P]);
        const float wv = valid ? wt : 0.0f;
        const float C = tap_cost<NC>(v1, g1, fa, tg, tdg, t0, dt);
        ws += wv;
        rt = __builtin_fmaf(wv, C, rt);
    }
}

// Stages the other view's texels of row `rowo` that the wave's taps can reach
// (x' = a qx + rb for |qx - x| <= P is monotone in qx, so the window ends
// bound it).  Returns false if no lane has a valid tap in the row; otherwise
// s0 is the first staged texel and `staged` says whether the span fit the
// stage (if not, taps gather through the L1 instead); `whole` says whether
// every lane has a pixel whose window row maps into [0, W) (then every tap of
// the row is valid: x' is monotone between the window ends).
template <int kS, int NC>
__device__ __forceinline__ bool stage_row(const CostCtx& cc, int x, bool active, float a,
                                          float rb, unsigned rowo, GatherStageT<kS, NC>& st,
                                          int& s0, bool& staged, bool& whole) {
    const int W = cc.W, P = cc.P;
    const float Wf = (float)W;
    const float xa = __builtin_fmaf(a, (float)(x - P), rb);
    const float xb = __builtin_fmaf(a, (float)(x + P), rb);
    const float lo = fminf(xa, xb), hi = fmaxf(xa, xb);
    const bool any = active && hi >= 0.0f && lo < Wf;  // false for NaN
    whole = __ballot(active && lo >= 0.0f && hi < Wf) == ~0ull;
    s0 = wave_min_i(any ? (int)fmaxf(lo, 0.0f) : 0x7fffffff);
    const int s1 = wave_max_i(any ? min((int)fminf(hi, Wf), W - 1) : -1);
    if (s1 < s0) return false;  // no valid tap in the wave: the row adds zeros
    const int span = s1 - s0 + 1;
    staged = span <= kS;
    if (staged) {
        wave_lds_sync();  // the previous row's stage is no longer read
        for (int i = lane_id(); i <= span; i += kWave) {
            const unsigned tx = rowo + (unsigned)min(s0 + i, W - 1);
            st.g[i] = load16(cc.rec, tx * 16u);
            if constexpr (NC == 1) {
                st.c[i] = load4(cc.rec, tx * 8u + cc.id_base);
            } else {
                st.c[i] = load16(cc.rec, tx * 32u + cc.id_base);
            }
        }
        wave_lds_sync();
    }
    return true;
}

// One window row (qy = y + dy, inside the image) of one plane: stages the
// row's span and adds its taps to ws and rt in tap order.
template <int kS, int NC>
__device__ __forceinline__ void plane_row(const CostCtx& cc, int x, bool active, bool interior,
                                          int qy, CFloat* Fwr, typename Col<NC>::V c0,
                                          const float4 pl, GatherStageT<kS, NC>& st, float& ws,
                                          float& rt) {
    using V = typename Col<NC>::V;
    const int P = cc.P;
    const unsigned rowo = (unsigned)(qy * cc.W);
    const float rb = __builtin_fmaf(pl.y, (float)qy, pl.z);
    int s0;
    bool staged, whole;
    if (!stage_row(cc, x, active, pl.x, rb, rowo, st, s0, staged, whole)) return;
    // taps that do not count read texel s0 (staged, in the row)
    const float park = (float)s0;
    if (staged && interior) {
        const auto base = st.row_base(s0);
        auto fetch = [&](int i0, float4& tg, float4& tdg, V& t0, V& dt) {
            st.get_texel(base, i0, tg, tdg, t0, dt);
        };
        if (whole)
            row_taps_inner<NC, true>(cc, x, P, rowo, rb, pl.x, c0, Fwr, park, ws, rt, fetch);
        else
            row_taps_inner<NC, false>(cc, x, P, rowo, rb, pl.x, c0, Fwr, park, ws, rt, fetch);
    } else if (staged) {
        const auto base = st.row_base(s0);
        row_taps<NC>(cc, x, active, P, rowo, rb, pl.x, c0, Fwr, park, ws, rt,
                     [&](int i0, bool, float4& tg, float4& tdg, V& t0, V& dt) {
                         st.get_texel(base, i0, tg, tdg, t0, dt);
                     });
    } else {
        row_taps<NC>(cc, x, active, P, rowo, rb, pl.x, c0, Fwr, park, ws, rt,
                     [&](int i0, bool, float4& tg, float4& tdg, V& t0, V& dt) {
                         gather_texel<NC>(cc, (unsigned)i0 + rowo, tg, tdg, t0, dt);
                     });
    }
}

// Every lane of the wave takes part (the block is one wave; lanes without a
// pixel pass active = false and cost a clamped duplicate).
template <int kS, int NC>
__device__ __forceinline__ float plane_cost(const CostCtx& cc, int x, bool active, int y,
                                            typename Col<NC>::V c0, const float4 pl,
                                            GatherStageT<kS, NC>& st) {
    const int W = cc.W, P = cc.P;
    const int S = 2 * P + 1;
    float ws = 0.0f, rt = 0.0f;
    // every lane has a pixel whose window stays inside the image row
    const bool interior = __ballot(active && x - P >= 0 && x + P < W) == ~0ull;
    for (int dy = -P; dy <= P; ++dy) {
        const int qy = y + dy;
        if (qy < 0 || qy >= cc.H) continue;
        plane_row(cc, x, active, interior, qy, cc.Fw + (dy + P) * S + P, c0, pl, st, ws, rt);
    }
    const bool fin = isfinite(pl.x) && isfinite(pl.y) && isfinite(pl.z);
    return (!fin || ws <= 1.1f) ? 1e5f : rt / ws;
}

// The taps of one window row for N planes at once, the union of the planes'
// spans staged from s0.  This view's texels and the support weight of a tap
// do not depend on the plane, so they are loaded and computed once per tap;
// each plane then reads its own cell and adds to its own sums in tap order,
// the arithmetic of row_taps_inner (interior segment: every lane has a pixel
// whose window lies inside the image row) or of row_taps (kEdge: a segment at
// the image's left or right edge).  kWhole: every plane maps every lane's
// window row into [0, W).
template <int kS, int NC, int N, bool kWhole, bool kEdge>
__device__ __forceinline__ void row_taps_multi(const CostCtx& cc, int x, bool active,
                                               unsigned rowo, const float (&a)[N],
                                               const float (&rb)[N], typename Col<NC>::V c0,
                                               CFloat* Fwr, int s0,
                                               const GatherStageT<kS, NC>& st, float (&ws)[N],
                                               float (&rt)[N]) {
    static_assert(!(kEdge && kWhole), "an edge segment has taps outside the image");
    using V = typename Col<NC>::V;
    const int P = cc.P, W = cc.W;
    const unsigned q0 = rowo + (unsigned)(x - P);
    const unsigned vi = q0 * (4u * NC), vg = q0 * 16u;
    const float Wf = (float)W, park = (float)s0;
    const auto base = st.row_base(s0);
    float qf = (float)(x - P);
    constexpr int kUnroll = NC == 1 && !kEdge ? 2 : 1;
#pragma unroll kUnroll
    for (int k = 0; k <= 2 * P; ++k, qf += 1.0f) {
        V v1;
        float4 g1;
        bool inq = true;
        if constexpr (kEdge) {
            const int qx = x - P + k;
            inq = (unsigned)qx < (unsigned)W;
            const unsigned q = rowo + (unsigned)(inq ? qx : x);
            v1 = Col<NC>::load(cc.i1, q);
            g1 = load16(cc.g1, q * 16u);
        } else {
            v1 = Col<NC>::load_s(cc.i1, vi, (unsigned)k * (4u * NC));
            g1 = load16s(cc.g1, vg, (unsigned)k * 16u);
        }
        const float wt = __builtin_amdgcn_exp2f(Col<NC>::dsum(c0, v1) * Fwr[k - P]);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float xp = __builtin_fmaf(a[j], qf, rb[j]);
            const bool valid = kWhole || ((!kEdge || active) && (xp >= 0.0f) && (xp < Wf));
            const float xs = valid ? xp : park;
            float4 tg, tdg;
            V t0, dt;
            st.get_texel(base, (int)xs, tg, tdg, t0, dt);
            const float wv = (valid && inq) ? wt : 0.0f;
            const float C = tap_cost<NC>(v1, g1, __builtin_amdgcn_fractf(xs), tg, tdg, t0, dt);
            ws[j] += wv;
            rt[j] = __builtin_fmaf(wv, C, rt[j]);
        }
    }
}

// Stages texels [s0, s0 + span] of row `rowo` (span <= kS) for the wave.
template <int kS, int NC>
__device__ __forceinline__ void stage_span(const CostCtx& cc, unsigned rowo, int s0, int span,
                                           GatherStageT<kS, NC>& st) {
    wave_lds_sync();  // the previous row's stage is no longer read
    for (int i = lane_id(); i <= span; i += kWave) {
        const unsigned tx = rowo + (unsigned)min(s0 + i, cc.W - 1);
        st.g[i] = load16(cc.rec, tx * 16u);
        if constexpr (NC == 1) {
            st.c[i] = load4(cc.rec, tx * 8u + cc.id_base);
        } else {
            st.c[i] = load16(cc.rec, tx * 32u + cc.id_base);
        }
    }
    wave_lds_sync();
}

// plane_cost of N planes of the same pixels: out[j] has the bits of
// plane_cost(..., pl[j], ...).  A window row whose union span (over the
// planes and the lanes with a pixel) fits the stage is staged once and swept
// by row_taps_multi; otherwise each plane takes the row alone (plane_row).
// A plane's sums are built in the same tap order whichever way a row is
// taken, and a tap that does not count adds exact zeros (its parked cell is
// finite), so the bits do not depend on it.
template <int kS, int NC, int N>
__device__ __forceinline__ void plane_costs(const CostCtx& cc, int x, bool active, int y,
                                            typename Col<NC>::V c0, const float4 (&pl)[N],
                                            GatherStageT<kS, NC>& st, float (&out)[N]) {
    const int W = cc.W, P = cc.P;
    const int S = 2 * P + 1;
    const float Wf = (float)W;
    float ws[N], rt[N], a[N], b[N], c[N], rb[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        ws[j] = 0.0f;
        rt[j] = 0.0f;
        a[j] = pl[j].x;
        b[j] = pl[j].y;
        c[j] = pl[j].z;
    }
    const bool interior = __ballot(active && x - P >= 0 && x + P < W) == ~0ull;
    for (int dy = -P; dy <= P; ++dy) {
        const int qy = y + dy;
        if (qy < 0 || qy >= cc.H) continue;
        const unsigned rowo = (unsigned)(qy * W);
        CFloat* Fwr = cc.Fw + (dy + P) * S + P;
        int lo_l = 0x7fffffff, hi_l = -1;
        bool whole = true;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            rb[j] = __builtin_fmaf(b[j], (float)qy, c[j]);
            const float xa = __builtin_fmaf(a[j], (float)(x - P), rb[j]);
            const float xb = __builtin_fmaf(a[j], (float)(x + P), rb[j]);
            const float lo = fminf(xa, xb), hi = fmaxf(xa, xb);
            if (active && hi >= 0.0f && lo < Wf) {  // false for NaN
                lo_l = min(lo_l, (int)fmaxf(lo, 0.0f));
                hi_l = max(hi_l, min((int)fminf(hi, Wf), W - 1));
            }
            whole = whole && active && lo >= 0.0f && hi < Wf;
        }
        const int s0 = wave_min_i(lo_l);
        const int s1 = wave_max_i(hi_l);
        if (s1 < s0) continue;  // no plane has a valid tap in the row
        if (s1 - s0 + 1 <= kS) {
            stage_span(cc, rowo, s0, s1 - s0 + 1, st);
            if (!interior)
                row_taps_multi<kS, NC, N, false, true>(cc, x, active, rowo, a, rb, c0, Fwr, s0,
                                                       st, ws, rt);
            else if (__ballot(whole) == ~0ull)
                row_taps_multi<kS, NC, N, true, false>(cc, x, active, rowo, a, rb, c0, Fwr, s0,
                                                       st, ws, rt);
            else
                row_taps_multi<kS, NC, N, false, false>(cc, x, active, rowo, a, rb, c0, Fwr, s0,
                                                        st, ws, rt);
            continue;
        }
#pragma unroll
        for (int j = 0; j < N; ++j)
            plane_row(cc, x, active, interior, qy, Fwr, c0, pl[j], st, ws[j], rt[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const bool fin = isfinite(pl[j].x) && isfinite(pl[j].y) && isfinite(pl[j].z);
        out[j] = (!fin || ws[j] <= 1.1f) ? 1e5f : rt[j] / ws[j];
    }
}

// ------------------------------------------------- block-shared stage ----
// The four waves of a propagate block cost candidates of the same row
// segment, so their window rows are the same rows: one LDS stage per block
// holds the union of every wave's span of the other view's row, and the waves
// sweep the window rows in lockstep.  A stage shared four ways can afford to
// carry the bilinear differences with each texel (G, dG, I, dI: the gather
// planes' contents, 40 bytes), so a tap reads its cell and does not rebuild
// dG = G(x+1) - G(x), dI = I(x+1) - I(x); and it can be about twice as wide as
// the four per-wave stages together, so fewer rows overflow.  Gray inputs only.
//
// The stage keeps G, dG and (I, dI) in three arrays indexed by texel (16-, 16-
// and 8-byte reads per cell); one 40-byte record per texel, read as 8-byte
// pieces, kept the LDS 70% busier.  A tap's texel depends on its lane's own
// candidate plane, so the lanes' addresses are as good as random within the
// span and bank conflicts (about half of the LDS-active cycles) do not follow
// a lane stride: an XOR swizzle of the texel index that separates lanes two
// texels apart measured no fewer conflicts and cost two VALU per plane-tap.
constexpr int kBlockMaxN = 5;  // planes one wave sweeps together

template <int kS>
struct BlockStage {
    float4 g[kS];   // G
    float4 d[kS];   // dG
    float2 id[kS];  // (I, dI)
    static constexpr int kQuarter = kS / kSplitWavesStage;  // per-wave part (overflow rows)
    static __device__ __forceinline__ unsigned addr(const void* p) {
        return (unsigned)(size_t)(const __attribute__((address_space(3))) void*)p;
    }
};

// LDS byte addresses of a stage region that starts at slot `first` and holds
// texels from s0: the G array's and the (I, dI) array's, offset so that texel
// i sits at address + 16 i (resp. 8 i).
struct RecBase {
    unsigned g, id;
};
template <int kS>
__device__ __forceinline__ RecBase rec_base(const BlockStage<kS>& st, int first, int s0) {
    const unsigned g = BlockStage<kS>::addr(st.g) + (unsigned)(first - s0) * 16u;
    const unsigned id = BlockStage<kS>::addr(st.id) + (unsigned)(first - s0) * 8u;
    int gs = __builtin_amdgcn_readfirstlane((int)g), ids = __builtin_amdgcn_readfirstlane((int)id);
    // opaque to the optimizer: otherwise it factors both addresses as
    // (i0 - s0) << k, one more VALU operation per plane-tap than two
    // shift-and-adds on scalar bases
    asm volatile("" : "+s"(gs), "+s"(ids));
    return {(unsigned)gs, (unsigned)ids};
}

// The cell of texel i0 in a stage region (kS: the stage's slot count, the
// distance of the dG array from the G array).
template <int kS>
__device__ __forceinline__ void rec_cell(const RecBase& rb, int i0, float4& tg, float4& tdg,
                                         float& t0, float& dt) {
    using L4 = const __attribute__((address_space(3))) f32x4;
    using L2 = const __attribute__((address_space(3))) f32x2;
    L4* gp = (L4*)(size_t)(rb.g + ((unsigned)i0 << 4));
    const f32x4 g = gp[0], d = gp[kS];
    const f32x2 id = *(L2*)(size_t)(rb.id + ((unsigned)i0 << 3));
    tg = make_float4(g.x, g.y, g.z, g.w);
    tdg = make_float4(d.x, d.y, d.z, d.w);
    t0 = id.x;
    dt = id.y;
}

// Copies the gather-plane cells of texels [s0, s1] of row `rowo` into a stage
// region starting at slot `first`; `lane`/`lanes`: this thread's share.
template <int kS>
__device__ __forceinline__ void rec_stage(const CostCtx& cc, unsigned rowo, int s0, int s1,
                                          int first, int lane, int lanes, BlockStage<kS>& st) {
    for (int i = s0 + lane; i <= s1; i += lanes) {
        const unsigned tx = rowo + (unsigned)i;
        const int sl = first + i - s0;
        st.g[sl] = load16(cc.rec, tx * 16u);
        st.d[sl] = load16(cc.rec, tx * 16u + cc.dg_base);
        st.id[sl] = load8(cc.rec, tx * 8u + cc.id_base);
    }
}

// The taps of one window row for the first N planes, cells read from a stage
// region (rb) that holds the texels the row's valid taps reach.  The
// arithmetic of row_taps_multi (and so of plane_cost): kWhole, kEdge as there.
template <int kS, int N, bool kWhole, bool kEdge>
__device__ __forceinline__ void rec_taps(const CostCtx& cc, int x, bool active, unsigned rowo,
                                         const float (&a)[kBlockMaxN],
                                         const float (&rb)[kBlockMaxN], float c0, CFloat* Fwr,
                                         int s0, const RecBase& base, float (&ws)[kBlockMaxN],
                                         float (&rt)[kBlockMaxN]) {
    static_assert(!(kEdge && kWhole), "an edge segment has taps outside the image");
    const int P = cc.P, W = cc.W;
    const unsigned q0 = rowo + (unsigned)(x - P);
    const unsigned vi = q0 * 4u, vg = q0 * 16u;
    const float Wf = (float)W, park = (float)s0;
    float qf = (float)(x - P);
    // one plane (the refinement sweeps): deeper, so that this view's loads of
    // several taps are in flight at once (35 taps at P = 17 are 5 trips)
    constexpr int kUnroll = kEdge ? 1 : (N == 1 ? 7 : 2);
#pragma unroll kUnroll
    for (int k = 0; k <= 2 * P; ++k, qf += 1.0f) {
        float v1;
        float4 g1;
        bool inq = true;
        if constexpr (kEdge) {
            const int qx = x - P + k;
            inq = (unsigned)qx < (unsigned)W;
            const unsigned q = rowo + (unsigned)(inq ? qx : x);
            v1 = load4(cc.i1, q * 4u);
            g1 = load16(cc.g1, q * 16u);
        } else {
            v1 = load4s(cc.i1, vi, (unsigned)k * 4u);
            g1 = load16s(cc.g1, vg, (unsigned)k * 16u);
        }
        const float wt = __builtin_amdgcn_exp2f(Col<1>::dsum(c0, v1) * Fwr[k - P]);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float xp = __builtin_fmaf(a[j], qf, rb[j]);
            const bool valid = kWhole || ((!kEdge || active) && (xp >= 0.0f) && (xp < Wf));
            const float xs = valid ? xp : park;
            float4 tg, tdg;
            float t0, dt;
            rec_cell<kS>(base, (int)xs, tg, tdg, t0, dt);
            const float wv = (valid && inq) ? wt : 0.0f;
            const float C = tap_cost<1>(v1, g1, __builtin_amdgcn_fractf(xs), tg, tdg, t0, dt);
            ws[j] += wv;
            rt[j] = __builtin_fmaf(wv, C, rt[j]);
        }
    }
}

// rec_taps for the first N planes with the loop variant the segment and the
// planes allow.
template <int kS, int N>
__device__ __forceinline__ void rec_taps_any(const CostCtx& cc, int x, bool active, bool interior,
                                             bool whole, unsigned rowo,
                                             const float (&a)[kBlockMaxN],
                                             const float (&rb)[kBlockMaxN], float c0, CFloat* Fwr,
                                             int s0, const RecBase& base,
                                             float (&ws)[kBlockMaxN], float (&rt)[kBlockMaxN]) {
    if (!interior)
        rec_taps<kS, N, false, true>(cc, x, active, rowo, a, rb, c0, Fwr, s0, base, ws, rt);
    else if (whole)
        rec_taps<kS, N, true, false>(cc, x, active, rowo, a, rb, c0, Fwr, s0, base, ws, rt);
    else
        rec_taps<kS, N, false, false>(cc, x, active, rowo, a, rb, c0, Fwr, s0, base, ws, rt);
}

// Shared left/right neighbour planes.  In a red or black pass the pixel two
// columns right of pixel x (x + 1 in between) has x's right, up-right and
// down-right neighbours (x + 1, y + dy) as its left, up-left and down-left
// neighbours, and its window columns are x's shifted by two.  One sweep over
// the 2P + 3 columns x - P .. x + P + 2 therefore gives both pixels' costs of
// each of these three planes: a tap's matching cost depends on the plane and
// the column only, its support weight on the pixel (own colour, factor of the
// column offset).  Each pixel's sums take its own columns in its own order
// with its own weights (a column outside a pixel's window adds weight 0 and
// 0 x the finite cost, which leaves the sums' bits unchanged), so both costs
// have the bits of plane_cost.  The same holds for a Jacobi sweep (the pixel
// two columns right shares the neighbour between them).
constexpr int kExtPlanes = 3;

// Lane data of the wave that costs the shared planes.
struct ExtLane {
    int xt;        // the lane's column (may lie outside [0, W) for a lane without a pixel)
    bool own, nb;  // taps count for the pixel at xt / for the pixel at xt + 2
    float c0n;     // the colour of the pixel at xt + 2
};

// The taps of one window row for the kExtPlanes shared planes (j < 3: own
// and xt + 2 sums) and N - kExtPlanes planes of the lane's own pixel, cells
// from a stage region that holds every counted lane's columns xt - P ..
// xt + P + 2.  kWhole, kEdge as for rec_taps, over those columns.
template <int kS, int N, bool kWhole, bool kEdge>
__device__ __forceinline__ void rec_taps_ext(const CostCtx& cc, const ExtLane& e, unsigned rowo,
                                             const float (&a)[kBlockMaxN],
                                             const float (&rb)[kBlockMaxN], float c0, CFloat* Fwr,
                                             int s0, const RecBase& base,
                                             float (&ws)[kBlockMaxN], float (&rt)[kBlockMaxN],
                                             float (&wsn)[kBlockMaxN], float (&rtn)[kBlockMaxN]) {
    static_assert(!(kEdge && kWhole), "an edge segment has taps outside the image");
    static_assert(N >= kExtPlanes && N <= kBlockMaxN, "the shared planes come first");
    const int P = cc.P, W = cc.W;
    const float Wf = (float)W, park = (float)s0;
    const bool cnt = e.own || e.nb;
    const unsigned q0 = rowo + (unsigned)(e.xt - P);  // interior: every column inside the row
    const unsigned vi = q0 * 4u, vg = q0 * 16u;
    // column k of the sweep; kO: inside the own window (k <= 2P), kN: inside
    // the window of xt + 2 (k >= 2)
    auto col = [&](int k, float qf, auto o_, auto n_) {
        constexpr bool kO = decltype(o_)::value, kN = decltype(n_)::value;
        float v1;
        float4 g1;
        bool inq = true;
        if constexpr (kEdge) {
            const int qx = e.xt - P + k;
            inq = (unsigned)qx < (unsigned)W;
            const unsigned q = rowo + (unsigned)clampi(qx, 0, W - 1);
            v1 = load4(cc.i1, q * 4u);
            g1 = load16(cc.g1, q * 16u);
        } else {
            v1 = load4s(cc.i1, vi, (unsigned)k * 4u);
            g1 = load16s(cc.g1, vg, (unsigned)k * 16u);
        }
        float wo = 0.0f, wn = 0.0f;
        if constexpr (kO) {
            const float t = __builtin_amdgcn_exp2f(Col<1>::dsum(c0, v1) * Fwr[k - P]);
            wo = (e.own && inq) ? t : 0.0f;
        }
        if constexpr (kN) {
            const float t = __builtin_amdgcn_exp2f(Col<1>::dsum(e.c0n, v1) * Fwr[k - 2 - P]);
            wn = (e.nb && inq) ? t : 0.0f;
        }
        constexpr int J = kO ? N : kExtPlanes;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const float xp = __builtin_fmaf(a[j], qf, rb[j]);
            const bool valid = kWhole || (cnt && (xp >= 0.0f) && (xp < Wf));
            const float xs = valid ? xp : park;
            float4 tg, tdg;
            float t0, dt;
            rec_cell<kS>(base, (int)xs, tg, tdg, t0, dt);
            const float C = tap_cost<1>(v1, g1, __builtin_amdgcn_fractf(xs), tg, tdg, t0, dt);
            if constexpr (kO) {
                const float wv = valid ? wo : 0.0f;
                ws[j] += wv;
                rt[j] = __builtin_fmaf(wv, C, rt[j]);
            }
            if constexpr (kN) {
                if (j < kExtPlanes) {
                    const float wv = valid ? wn : 0.0f;
                    wsn[j] += wv;
                    rtn[j] = __builtin_fmaf(wv, C, rtn[j]);
                }
            }
        }
    };
    using T = std::true_type;
    using F = std::false_type;
    float qf = (float)(e.xt - P);  // (float)qx, stepped exactly (|qx| < 2^24)
    col(0, qf, T{}, F{});
    qf += 1.0f;
    col(1, qf, T{}, F{});
    qf += 1.0f;
#pragma unroll 2
    for (int k = 2; k <= 2 * P; ++k, qf += 1.0f) col(k, qf, T{}, T{});
    col(2 * P + 1, qf, F{}, T{});
    qf += 1.0f;
    col(2 * P + 2, qf, F{}, T{});
}

template <int kS, int N>
__device__ __forceinline__ void rec_taps_ext_any(const CostCtx& cc, const ExtLane& e,
                                                 bool interior, bool whole, unsigned rowo,
                                                 const float (&a)[kBlockMaxN],
                                                 const float (&rb)[kBlockMaxN], float c0,
                                                 CFloat* Fwr, int s0, const RecBase& base,
                                                 float (&ws)[kBlockMaxN], float (&rt)[kBlockMaxN],
                                                 float (&wsn)[kBlockMaxN],
                                                 float (&rtn)[kBlockMaxN]) {
    if (!interior)
        rec_taps_ext<kS, N, false, true>(cc, e, rowo, a, rb, c0, Fwr, s0, base, ws, rt, wsn, rtn);
    else if (whole)
        rec_taps_ext<kS, N, true, false>(cc, e, rowo, a, rb, c0, Fwr, s0, base, ws, rt, wsn, rtn);
    else
        rec_taps_ext<kS, N, false, false>(cc, e, rowo, a, rb, c0, Fwr, s0, base, ws, rt, wsn, rtn);
}

// Window costs of this wave's first n planes (n wave-uniform, 0 .. kBlockMaxN)
// with all waves of the block sweeping the window rows together: out[j] has
// the bits of plane_cost(pl[j]).  Every wave of the block must call it with
// the same (y, P) -- the barriers are block-wide.  A row whose union over the
// block overflows the stage is taken plane by plane by each wave on its own
// quarter of the stage (or by gathers through the L1 if the plane's span does
// not fit the quarter).  spans: LDS scratch of 2 x 4 x 2 ints.
// kExt: this wave's first kExtPlanes planes are the shared planes of lane
// data e (n = kExtPlanes or one more; x, active: e.xt clamped, e.own); outn[j] has the
// bits of plane_cost(pl[j]) at the pixel e.xt + 2.  Its overflow rows are
// gathered through the L1.
template <int kS, bool kExt>
__device__ __forceinline__ void block_costs(const CostCtx& cc, int x, bool active, int y, float c0,
                                            const float4 (&pl)[kBlockMaxN], int n,
                                            const ExtLane& e, BlockStage<kS>& st, int* spans,
                                            float (&out)[kBlockMaxN],
                                            float (&outn)[kBlockMaxN]) {
    const int W = cc.W, P = cc.P;
    const int S = 2 * P + 1;
    const float Wf = (float)W;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = lane_id();
    constexpr int kQ = BlockStage<kS>::kQuarter;
    float ws[kBlockMaxN], rt[kBlockMaxN], wsn[kBlockMaxN], rtn[kBlockMaxN], a[kBlockMaxN],
        b[kBlockMaxN], c[kBlockMaxN], rb[kBlockMaxN];
#pragma unroll
    for (int j = 0; j < kBlockMaxN; ++j) {
        ws[j] = 0.0f;
        rt[j] = 0.0f;
        wsn[j] = 0.0f;
        rtn[j] = 0.0f;
        a[j] = pl[j].x;
        b[j] = pl[j].y;
        c[j] = pl[j].z;
    }
    // the columns a lane's taps read, xl .. xr, and whether any of them counts
    const int xl = kExt ? e.xt - P : x - P;
    const int xr = kExt ? e.xt + P + 2 : x + P;
    const bool cnt = kExt ? (e.own || e.nb) : active;
    const bool interior = __ballot(cnt && xl >= 0 && xr < W) == ~0ull;
    int par = 0;
    for (int dy = -P; dy <= P; ++dy) {
        const int qy = y + dy;
        if (qy < 0 || qy >= cc.H) continue;  // the same rows for every wave
        const unsigned rowo = (unsigned)(qy * W);
        CFloat* Fwr = cc.Fw + (dy + P) * S + P;
        int lo_l = 0x7fffffff, hi_l = -1;
        bool whole = true;
#pragma unroll
        for (int j = 0; j < kBlockMaxN; ++j) {
            rb[j] = __builtin_fmaf(b[j], (float)qy, c[j]);
            if (j < n) {
                const float xa = __builtin_fmaf(a[j], (float)xl, rb[j]);
                const float xb = __builtin_fmaf(a[j], (float)xr, rb[j]);
                const float lo = fminf(xa, xb), hi = fmaxf(xa, xb);
                if (cnt && hi >= 0.0f && lo < Wf) {  // false for NaN
                    lo_l = min(lo_l, (int)fmaxf(lo, 0.0f));
                    hi_l = max(hi_l, min((int)fminf(hi, Wf), W - 1));
                }
                whole = whole && cnt && lo >= 0.0f && hi < Wf;
            }
        }
        const int s0w = wave_min_i(lo_l);
        const int s1w = wave_max_i(hi_l);
        int* sp = spans + par * (2 * kSplitWavesStage);
        if (lane == 0) {
            sp[2 * w] = s0w;
            sp[2 * w + 1] = s1w;
        }
        par ^= 1;
        __syncthreads();  // spans visible; every wave is done with the previous row's stage
        int s0 = sp[0], s1 = sp[1];
#pragma unroll
        for (int v = 1; v < kSplitWavesStage; ++v) {
            s0 = min(s0, sp[2 * v]);
            s1 = max(s1, sp[2 * v + 1]);
        }
        s0 = __builtin_amdgcn_readfirstlane(s0);
        s1 = __builtin_amdgcn_readfirstlane(s1);
        if (s1 < s0) continue;  // no wave has a valid tap in the row
        whole = __ballot(whole) == ~0ull;
        if (s1 - s0 + 1 <= kS) {
            rec_stage(cc, rowo, s0, s1, 0, (int)threadIdx.x, kSplitWavesStage * kWave, st);
            __syncthreads();
            const RecBase base = rec_base(st, 0, s0);
            if constexpr (kExt) {
                switch (n) {
                case 3: rec_taps_ext_any<kS, 3>(cc, e, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt, wsn, rtn); break;
                case 4: rec_taps_ext_any<kS, 4>(cc, e, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt, wsn, rtn); break;
                default: break;  // the shared planes' wave holds at most one more
                }
            } else {
                switch (n) {
                case 1: rec_taps_any<kS, 1>(cc, x, active, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt); break;
                case 2: rec_taps_any<kS, 2>(cc, x, active, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt); break;
                case 3: rec_taps_any<kS, 3>(cc, x, active, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt); break;
                case 4: rec_taps_any<kS, 4>(cc, x, active, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt); break;
                case 5: rec_taps_any<kS, 5>(cc, x, active, interior, whole, rowo, a, rb, c0, Fwr, s0, base, ws, rt); break;
                default: break;
                }
            }
            continue;
        }
        auto gather = [&](int i0, bool, float4& tg, float4& tdg, float& t0, float& dt) {
            gather_texel<1>(cc, (unsigned)i0 + rowo, tg, tdg, t0, dt);
        };
        if constexpr (kExt) {
            // overflow: every plane by gathers, each window on its own
#pragma unroll
            for (int j = 0; j < kBlockMaxN; ++j) {
                if (j >= n) break;
                row_taps<1>(cc, x, e.own, P, rowo, rb[j], a[j], c0, Fwr, 0.0f, ws[j], rt[j],
                            gather);
                if (j < kExtPlanes)
                    row_taps<1>(cc, clampi(e.xt + 2, 0, W - 1), e.nb, P, rowo, rb[j], a[j], e.c0n,
                                Fwr, 0.0f, wsn[j], rtn[j], gather);
            }
            continue;
        }
        // overflow: each wave takes its planes one at a time in its own quarter
#pragma unroll
        for (int j = 0; j < kBlockMaxN; ++j) {
            if (j >= n) break;
            const float xa = __builtin_fmaf(a[j], (float)(x - P), rb[j]);
            const float xb = __builtin_fmaf(a[j], (float)(x + P), rb[j]);
            const float lo = fminf(xa, xb), hi = fmaxf(xa, xb);
            const bool any = active && hi >= 0.0f && lo < Wf;
            const bool whole1 = __ballot(active && lo >= 0.0f && hi < Wf) == ~0ull;
            const int p0 = wave_min_i(any ? (int)fmaxf(lo, 0.0f) : 0x7fffffff);
            const int p1 = wave_max_i(any ? min((int)fminf(hi, Wf), W - 1) : -1);
            if (p1 < p0) continue;  // the plane adds zeros in this row
            float a1[kBlockMaxN] = {a[j]}, rb1[kBlockMaxN] = {rb[j]};
            float ws1[kBlockMaxN] = {ws[j]}, rt1[kBlockMaxN] = {rt[j]};
            if (p1 - p0 + 1 <= kQ) {
                wave_lds_sync();  // this wave's previous plane no longer reads the quarter
                rec_stage(cc, rowo, p0, p1, w * kQ, lane, kWave, st);
                wave_lds_sync();
                rec_taps_any<kS, 1>(cc, x, active, interior, whole1, rowo, a1, rb1, c0, Fwr, p0,
                                    rec_base(st, w * kQ, p0), ws1, rt1);
            } else {
                row_taps<1>(cc, x, active, P, rowo, rb1[0], a1[0], c0, Fwr, (float)p0, ws1[0],
                            rt1[0], gather);
            }
            ws[j] = ws1[0];
            rt[j] = rt1[0];
        }
    }
#pragma unroll
    for (int j = 0; j < kBlockMaxN; ++j) {
        const bool fin = isfinite(pl[j].x) && isfinite(pl[j].y) && isfinite(pl[j].z);
        out[j] = (!fin || ws[j] <= 1.1f) ? 1e5f : rt[j] / ws[j];
        outn[j] = (!fin || wsn[j] <= 1.1f) ? 1e5f : rtn[j] / wsn[j];
    }
}

// The pixel's own colour, the c0 of the window cost.
template <int NC>
__device__ __forceinline__ typename Col<NC>::V own_colour(const float* I1, int idx) {
    if constexpr (NC == 1) {
        return I1[idx];
    } else {
        return reinterpret_cast<const float4*>(I1)[idx];
    }
}

// Row segment of a block: 64 lanes on row y, every pixel (sweep 0) or every
// other pixel of one colour (sweep 1 red: x+y even, 2 black: x+y odd).
struct Seg {
    int x;        // this lane's pixel (clamped to W - 1)
    bool active;  // pixel exists
};

__device__ __forceinline__ Seg row_segment(int bx, int y, int W, int sweep) {
    const int lx = bx * kWave + lane_id();
    const int x = sweep == 0 ? lx : 2 * lx + ((y + (sweep - 1)) & 1);
    Seg s;
    s.active = x < W;
    s.x = x < W ? x : W - 1;
    return s;
}

}  // namespace pmh
