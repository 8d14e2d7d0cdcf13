// pmh_adjoint.hip -- adjoint of the window cost and the gradient-step plane
// refinement built on it.  Neither has a counterpart in the reference; both
// differentiate the forward cost of patch_eval_color_grad_weighted
// (pmhstereo.cl:170-228) as restated in pmh_window.hpp.
//
// cost = R / Sw with R = sum_q w_q C_q and Sw = sum_q w_q over valid taps q.
// The support weights w_q do not depend on the plane, so
//   d cost / d theta = (1/Sw) sum_q w_q dC_q/dx' dx'/dtheta,
// with x' = a qx + b qy + c and dC/dx' from the bilinear cell (piecewise
// linear: the derivative of min(|d|, t) is sign(d) below t, 0 above; sign(0)
// is 0).  Windows are recomputed; nothing of the forward pass is stored.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pmh_device.hpp"
#include "pmh_internal.hpp"
#include "pmh_window.hpp"

namespace pmh {

__device__ __forceinline__ float sgnf(float v) { return (v > 0.0f) - (v < 0.0f); }

// k sign(d) where |d| < lim, else 0 (sign(0) = 0): the derivative of
// k min(|d|, lim) in d.  copysign(k, d) equals k sign(d) for d != 0, so
// this is the same value as the sgnf form in fewer operations.
__device__ __forceinline__ float kink_grad(float d, float lim, float k) {
    return (fabsf(d) < lim && d != 0.0f) ? __builtin_copysignf(k, d) : 0.0f;
}
// sign(d) v where |d| < lim, else 0 (exact: a sign flip of v)
__device__ __forceinline__ float kink_sel(float d, float lim, float v) {
    return (fabsf(d) < lim && d != 0.0f) ? (d > 0.0f ? v : -v) : 0.0f;
}

// Per-tap pieces shared by the forward sums and the adjoint.
struct Tap {
    float wv;      // support weight, 0 for a tap that does not count
    float C;       // matching cost of the tap
    float dC_dv1;  // d C / d I1(q)   (= - d C / d I2(x')); gray inputs
    float4 dC_dg1; // d C / d G1(q)   (= - d C / d G2(x'))
    float dC_dx;   // d C / d x'
    float fa;      // bilinear weight of the right texel
    int i0;        // left texel of the bilinear cell
    float w;       // support weight before masking
    float v1;      // this view's value at the tap; gray inputs
    float4 dCv;    // d C / d I1_c(q) of the colour channels (RGB); RGBA inputs
    float4 v1c;    // this view's colour at the tap; RGBA inputs
};

// The cost terms of one tap and their derivatives from its inputs: this
// view's v1, g1, the other view's bilinear cell (tg, tdg, t0, dt) at fraction
// fa, the weight exponent Fw and whether the tap counts.  The cost terms are
// those of row_taps, in the same order (same bits).
template <int NC>
__device__ __forceinline__ Tap tap_math(typename Col<NC>::V c0, typename Col<NC>::V v1,
                                        const float4 g1, float fa, int i0, const float4 tg,
                                        const float4 tdg, typename Col<NC>::V t0,
                                        typename Col<NC>::V dt, float Fw, bool counts) {
    Tap t;
    t.i0 = i0;
    t.fa = fa;
    t.w = __builtin_amdgcn_exp2f(Col<NC>::dsum(c0, v1) * Fw);
    t.wv = counts ? t.w : 0.0f;
    const float cterm = Col<NC>::cterm(v1, t0, dt, t.fa);
    const float d0 = g1.x - __builtin_fmaf(t.fa, tdg.x, tg.x);
    const float d1 = g1.y - __builtin_fmaf(t.fa, tdg.y, tg.y);
    const float d2 = g1.z - __builtin_fmaf(t.fa, tdg.z, tg.z);
    const float d3 = g1.w - __builtin_fmaf(t.fa, tdg.w, tg.w);
    const float gs = ((fminf(fabsf(d0), 0.01f) + fminf(fabsf(d1), 0.01f)) + fminf(fabsf(d2), 0.01f)) +
                     fminf(fabsf(d3), 0.01f);
    t.C = __builtin_fmaf(0.95f, gs, cterm);
    t.dC_dg1 = make_float4(kink_grad(d0, 0.01f, 0.95f), kink_grad(d1, 0.01f, 0.95f),
                           kink_grad(d2, 0.01f, 0.95f), kink_grad(d3, 0.01f, 0.95f));
    if constexpr (NC == 1) {
        const float dv = v1 - __builtin_fmaf(t.fa, dt, t0);
        t.v1 = v1;
        t.dC_dv1 = kink_grad(dv, 0.04f, 0.15f);
        t.dC_dx = -(t.dC_dv1 * dt + t.dC_dg1.x * tdg.x + t.dC_dg1.y * tdg.y +
                    t.dC_dg1.z * tdg.z + t.dC_dg1.w * tdg.w);
    } else {
        const float gdx = t.dC_dg1.x * tdg.x + t.dC_dg1.y * tdg.y + t.dC_dg1.z * tdg.z +
                          t.dC_dg1.w * tdg.w;
        // d/dx' of 0.05 sum_c min(|v1_c - (t0_c + x' dt_c)|, 0.04)
        const float dr = v1.x - __builtin_fmaf(t.fa, dt.x, t0.x);
        const float dg = v1.y - __builtin_fmaf(t.fa, dt.y, t0.y);
        const float db = v1.z - __builtin_fmaf(t.fa, dt.z, t0.z);
        const float cx = (kink_sel(dr, 0.04f, dt.x) + kink_sel(dg, 0.04f, dt.y)) +
                         kink_sel(db, 0.04f, dt.z);
        t.v1 = 0.0f;
        t.dC_dv1 = 0.0f;
        t.v1c = v1;
        t.dCv = make_float4(kink_grad(dr, 0.04f, 0.05f), kink_grad(dg, 0.04f, 0.05f),
                            kink_grad(db, 0.04f, 0.05f), 0.0f);
        t.dC_dx = -(0.05f * cx + gdx);
    }
    return t;
}

// One tap at column qx of window row rowo.  FETCH(i0, valid, G, dG, I, dI)
// returns the other view's bilinear cell.
template <int NC, class FETCH>
__device__ __forceinline__ Tap eval_tap_f(const CostCtx& cc, int x, bool active, int qx,
                                          unsigned rowo, float rb, float a,
                                          typename Col<NC>::V c0, float Fw, FETCH fetch) {
    using V = typename Col<NC>::V;
    const int W = cc.W;
    const bool inq = (unsigned)qx < (unsigned)W;
    const unsigned q = rowo + (unsigned)(inq ? qx : x);
    const V v1 = Col<NC>::load(cc.i1, q);
    const float4 g1 = load16(cc.g1, q * 16u);
    const float xp = __builtin_fmaf(a, (float)qx, rb);
    const bool valid = active && (xp >= 0.0f) && (xp < (float)W);
    const float xs = valid ? xp : 0.0f;
    float4 tg, tdg;
    V t0, dt;
    fetch((int)xs, valid, tg, tdg, t0, dt);
    return tap_math<NC>(c0, v1, g1, __builtin_amdgcn_fractf(xs), (int)xs, tg, tdg, t0, dt, Fw,
                        valid && inq);
}

template <int NC>
__device__ __forceinline__ Tap eval_tap(const CostCtx& cc, int x, int qx, unsigned rowo, float rb,
                                        float a, typename Col<NC>::V c0, float Fw) {
    using V = typename Col<NC>::V;
    return eval_tap_f<NC>(cc, x, true, qx, rowo, rb, a, c0, Fw,
                          [&](int i0, bool, float4& tg, float4& tdg, V& t0, V& dt) {
                              gather_texel<NC>(cc, (unsigned)i0 + rowo, tg, tdg, t0, dt);
                          });
}

// Cost of one plane and the window sums of its gradient: ge = sum w dC/dx',
// gx = sum w dC/dx' (qx - x), gy = sum w dC/dx' (qy - y); ws = sum w.
struct CostGrad {
    float cost, ws, ge, gx, gy;
    bool ok;  // finite plane and ws > 1.1 (otherwise cost = 1e5, no gradient)
};

template <int NC>
__device__ __forceinline__ CostGrad plane_cost_grad(const CostCtx& cc, int x, int y,
                                                    typename Col<NC>::V c0, const float4 pl) {
    const int P = cc.P, S = 2 * P + 1;
    float ws = 0.0f, rt = 0.0f, ge = 0.0f, gx = 0.0f, gy = 0.0f;
    for (int dy = -P; dy <= P; ++dy) {
        const int qy = y + dy;
        if (qy < 0 || qy >= cc.H) continue;
        const unsigned rowo = (unsigned)(qy * cc.W);
        const float rb = __builtin_fmaf(pl.y, (float)qy, pl.z);
        CFloat* Fwr = cc.Fw + (dy + P) * S + P;
        float gyr = 0.0f;
#pragma unroll 3
        for (int dx = -P; dx <= P; ++dx) {
            const Tap t = eval_tap<NC>(cc, x, x + dx, rowo, rb, pl.x, c0, Fwr[dx]);
            ws += t.wv;
            rt = __builtin_fmaf(t.wv, t.C, rt);
            const float e = t.wv * t.dC_dx;
            ge += e;
            gx = __builtin_fmaf(e, (float)dx, gx);
            gyr += e;
        }
        gy = __builtin_fmaf(gyr, (float)dy, gy);
    }
    CostGrad r;
    r.ok = isfinite(pl.x) && isfinite(pl.y) && isfinite(pl.z) && ws > 1.1f;
    r.cost = r.ok ? rt / ws : 1e5f;
    r.ws = ws;
    r.ge = ge;
    r.gx = gx;
    r.gy = gy;
    return r;
}

// plane_cost_grad for a whole wave with the other view's rows staged in LDS
// (as plane_cost does); same arithmetic in the same order, so the same bits.
template <int kS, int NC>
__device__ __forceinline__ CostGrad plane_cost_grad_staged(const CostCtx& cc, int x, bool active,
                                                           int y, typename Col<NC>::V c0,
                                                           const float4 pl,
                                                           GatherStageT<kS, NC>& st) {
    using V = typename Col<NC>::V;
    const int P = cc.P, S = 2 * P + 1;
    float ws = 0.0f, rt = 0.0f, ge = 0.0f, gx = 0.0f, gy = 0.0f;
    const bool interior = __ballot(active && x - P >= 0 && x + P < cc.W) == ~0ull;
    for (int dy = -P; dy <= P; ++dy) {
        const int qy = y + dy;
        if (qy < 0 || qy >= cc.H) continue;
        const unsigned rowo = (unsigned)(qy * cc.W);
        const float rb = __builtin_fmaf(pl.y, (float)qy, pl.z);
        CFloat* Fwr = cc.Fw + (dy + P) * S + P;
        int s0;
        bool staged, whole;
        if (!stage_row(cc, x, active, pl.x, rb, rowo, st, s0, staged, whole)) continue;
        float gyr = 0.0f;
        auto acc = [&](const Tap& t, int dx) {
            ws += t.wv;
            rt = __builtin_fmaf(t.wv, t.C, rt);
            const float e = t.wv * t.dC_dx;
            ge += e;
            gx = __builtin_fmaf(e, (float)dx, gx);
            gyr += e;
        };
        auto taps = [&](auto fetch) {
#pragma unroll 3
            for (int dx = -P; dx <= P; ++dx)
                acc(eval_tap_f<NC>(cc, x, active, x + dx, rowo, rb, pl.x, c0, Fwr[dx], fetch), dx);
        };
        if (staged && interior) {
            // as row_taps_inner: this view's loads at a per-lane base plus a
            // wave-uniform offset; with `whole` every tap counts (no selects)
            const auto base = st.row_base(s0);
            const unsigned q0 = rowo + (unsigned)(x - P);
            const unsigned vi = q0 * (4u * NC), vg = q0 * 16u;
            const float Wf = (float)cc.W, park = (float)s0;
            auto inner = [&](auto all_valid) {
                float qf = (float)(x - P);
#pragma unroll 3
                for (int k = 0; k <= 2 * P; ++k, qf += 1.0f) {
                    const V v1 = Col<NC>::load_s(cc.i1, vi, (unsigned)k * (4u * NC));
                    const float4 g1 = load16s(cc.g1, vg, (unsigned)k * 16u);
                    const float xp = __builtin_fmaf(pl.x, qf, rb);
                    const bool valid = decltype(all_valid)::value || ((xp >= 0.0f) && (xp < Wf));
                    const float xs = valid ? xp : park;
                    float4 tg, tdg;
                    V t0, dt;
                    st.get_texel(base, (int)xs, tg, tdg, t0, dt);
                    acc(tap_math<NC>(c0, v1, g1, __builtin_amdgcn_fractf(xs), (int)xs, tg, tdg, t0,
                                     dt, Fwr[k - P], valid),
                        k - P);
                }
            };
            if (whole)
                inner(std::true_type{});
            else
                inner(std::false_type{});
        } else if (staged) {
            taps([&](int i0, bool valid, float4& tg, float4& tdg, V& t0, V& dt) {
                st.get(valid ? i0 - s0 : 0, tg, tdg, t0, dt);
            });
        } else {
            taps([&](int i0, bool, float4& tg, float4& tdg, V& t0, V& dt) {
                gather_texel<NC>(cc, (unsigned)i0 + rowo, tg, tdg, t0, dt);
            });
        }
        gy = __builtin_fmaf(gyr, (float)dy, gy);
    }
    CostGrad r;
    r.ok = isfinite(pl.x) && isfinite(pl.y) && isfinite(pl.z) && ws > 1.1f;
    r.cost = r.ok ? rt / ws : 1e5f;
    r.ws = ws;
    r.ge = ge;
    r.gx = gx;
    r.gy = gy;
    return r;
}

// ----------------------------------------------------------- cost backward --
// dstates[p] = gout[p] * d cost_p / d (a, b, c).  When image gradients are
// requested, a second sweep spreads gout[p] * d cost_p / d (image texel) over
// the texels (the texels a window reads depend on the plane, so the other
// view has no gather form).  The wave works one window row at a time:
//  - this view: at tap k lane l adds to column l + k of the row (counted from
//    x0 - P), so a register that moves down one lane per tap carries the
//    column sums (lane l holds column l + k); after each tap lane 0's column
//    k is final and is parked in lane k of `done` (k < 64) or `done2`.  The
//    row then goes to the image with one global atomic per texel and
//    component;
//  - the other view: a lane's taps walk its texels monotonically, so it
//    keeps the bilinear pair (i, i + 1) it is adding to in registers and
//    hands a texel on only when its taps leave it, into the wave's LDS row
//    of the span between the window ends of every lane (wider spans than
//    kAdjTex texels go to global atomics directly); the LDS row then goes to
//    the image like this view's.
constexpr int kAdjTex = 384;
constexpr int kAdjPad = 16;  // row pitch = 16 mod 64 banks: the flush's 4 components hit apart
// Colour channels of the window cost: gray counts its one channel three times
// (folded into the factors), RGBA sums R, G and B once each (alpha unused).
template <int NC>
constexpr int adj_colours() { return NC == kGray ? 1 : 3; }
template <int NC>
struct AdjRow {
    static constexpr int K = adj_colours<NC>() + 4;  // colour channels, then G.x .. G.w
    float v[K][kAdjTex + kAdjPad];  // the other view's sums by texel - s0
    int claim[kAdjTex];             // lane that adds to the texel without an atomic
};

__device__ __forceinline__ void lds_add(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// lane l <- lane l + 1, lane 63 <- 0 (DPP wave_shl:1)
__device__ __forceinline__ float lane_down(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
template <int NC, bool kImages>
__global__ __launch_bounds__(kWave) void cost_backward_kernel(
    const float* __restrict__ I1, const float4* __restrict__ G1, Gather g2,
    const float* __restrict__ F, int W, int H, int P, const float4* __restrict__ st,
    const float* __restrict__ gout, float4* __restrict__ dst, float* __restrict__ dI1,
    float4* __restrict__ dG1, float* __restrict__ dI2, float4* __restrict__ dG2, int nbx,
    int nb) {
    using V = typename Col<NC>::V;
    constexpr int NCC = adj_colours<NC>(), K = NCC + 4;
    const int t = xcd_tile(nb);
    if (t >= nb) return;
    const int y = __builtin_amdgcn_readfirstlane(t / nbx);
    const int bx = t % nbx;
    const Seg sg = row_segment(bx, y, W, 0);
    if (!kImages && !sg.active) return;
    const int x = sg.x;
    const int idx = y * W + x;
    const CostCtx cc = cost_ctx<NC>(I1, G1, g2, F, W, H, P);
    const V c0 = own_colour<NC>(I1, idx);
    const float4 pl = st[idx];
    const CostGrad cg = plane_cost_grad<NC>(cc, x, y, c0, pl);
    const bool ok = sg.active && cg.ok;
    const float k = ok ? gout[idx] / cg.ws : 0.0f;
    // x' = a (x + dx) + b (y + dy) + c
    if (sg.active)
        dst[idx] = ok ? make_float4(k * __builtin_fmaf((float)x, cg.ge, cg.gx),
                                    k * __builtin_fmaf((float)y, cg.ge, cg.gy), k * cg.ge, 0.0f)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (kImages) {
        if (!__ballot(ok)) return;
        __shared__ AdjRow<NC> acc;
        const int S = 2 * P + 1, lane = lane_id();
        const int c1 = bx * kWave - P;  // this view's first column of the row
        const bool want1 = dI1 || dG1, want2 = dI2 || dG2;
        float* const dG1f = reinterpret_cast<float*>(dG1);
        float* const dG2f = reinterpret_cast<float*>(dG2);
        const float Wf = (float)W;
        float dcenter[NCC];
#pragma unroll
        for (int c = 0; c < NCC; ++c) dcenter[c] = 0.0f;
        for (int dy = -P; dy <= P; ++dy) {
            const int qy = y + dy;
            if (qy < 0 || qy >= H) continue;
            const unsigned rowo = (unsigned)(qy * W);
            const float rb = __builtin_fmaf(pl.y, (float)qy, pl.z);
            // the other view's texels: x' = a qx + rb is monotone between the
            // window ends (the same fma as the taps), i1 = i0 + 1 clamped
            const float xa = __builtin_fmaf(pl.x, (float)(x - P), rb);
            const float xb = __builtin_fmaf(pl.x, (float)(x + P), rb);
            const float lo = fminf(xa, xb), hi = fmaxf(xa, xb);
            const bool any = ok && hi >= 0.0f && lo < Wf;
            const int s0 = wave_min_i(any ? (int)fmaxf(lo, 0.0f) : 0x7fffffff);
            const int s1 = wave_max_i(any ? min((int)fminf(hi, Wf) + 1, W - 1) : -1);
            if (s1 < s0) continue;  // no counted tap in the wave
            const int n2 = s1 - s0 + 1;
            const bool lds2 = want2 && n2 <= kAdjTex;
            if (lds2) {
                wave_lds_sync();  // the previous row's flush has read the sums
                for (int i = lane; i < n2; i += kWave)
#pragma unroll
                    for (int c = 0; c < K; ++c) acc.v[c][i] = 0.0f;
                wave_lds_sync();
            }
            // adds the other view's pending sums of texel tx (lanes in `on`).
            // Each lane writes its id to the texel's claim slot and reads it
            // back: one lane per texel finds its own id and adds with a plain
            // read-add-write (the winners' texels are distinct); lanes that
            // share a texel with the winner then add with LDS atomics, after
            // the winners' writes (a wave's LDS operations are performed in
            // program order).  LDS float atomics cost about two cycles per
            // lane, plain accesses a few per instruction.
            auto put2 = [&](bool on, int tx, const float (&v)[K]) {
                if (!__ballot(on)) return;
                if (lds2) {
                    const int j = on ? tx - s0 : 0;  // in [0, n2) for a lane in `on`
                    // atomic accesses: the read must see the wave's write, not
                    // this lane's own value forwarded by the compiler
                    if (on)
                        __hip_atomic_store(&acc.claim[j], lane, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    const bool win = on && __hip_atomic_load(&acc.claim[j], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP) == lane;
                    if (win) {
                        if (dI2)
                            for (int c = 0; c < NCC; ++c) acc.v[c][j] += v[c];
                        if (dG2)
                            for (int c = NCC; c < K; ++c) acc.v[c][j] += v[c];
                    }
                    if (__ballot(on && !win)) {
                        wave_lds_sync();
                        if (on && !win) {
                            if (dI2)
                                for (int c = 0; c < NCC; ++c) lds_add(&acc.v[c][j], v[c]);
                            if (dG2)
                                for (int c = NCC; c < K; ++c) lds_add(&acc.v[c][j], v[c]);
                        }
                    }
                } else if (on) {
                    if (dI2)
                        for (int c = 0; c < NCC; ++c)
                            unsafeAtomicAdd(&dI2[(rowo + tx) * (unsigned)NC + c], v[c]);
                    if (dG2)
                        for (int c = 0; c < 4; ++c)
                            unsafeAtomicAdd(&dG2f[(rowo + tx) * 4u + c], v[NCC + c]);
                }
            };
            float run[K], done[K], done2[K], p0[K], p1[K];
#pragma unroll
            for (int c = 0; c < K; ++c) run[c] = done[c] = done2[c] = p0[c] = p1[c] = 0.0f;
            int pi = -1;  // the other view's pending texels pi, min(pi + 1, W - 1)
            CFloat* Fr = cc.F + (dy + P) * S + P;
            CFloat* Fwr = cc.Fw + (dy + P) * S + P;
            auto eval = [&](int dx) {
                return eval_tap_f<NC>(cc, x, ok, x + dx, rowo, rb, pl.x, c0, Fwr[dx],
                                      [&](int i0, bool, float4& tg, float4& tdg, V& t0, V& dt) {
                                          gather_texel<NC>(cc, (unsigned)i0 + rowo, tg, tdg, t0,
                                                           dt);
                                      });
            };
            // adds one tap to the row sums (taps in order)
            auto step = [&](const Tap& tp, int dx) {
                const bool on = tp.wv != 0.0f;      // a counted tap
                const float kC = k * tp.wv;         // d L / d C_q (0 if not counted)
                const float kw = on ? k * (tp.C - cg.cost) : 0.0f;  // d L / d w_q
                // w = exp(-sum_rgb |c0 - v1| F): d w / d v1_c = F w sign(c0_c - v1_c)
                // = - d w / d c0_c (gray: 3 F w sign(c0 - v1))
                float dCc[NCC], dw[NCC];
                if constexpr (NC == kGray) {
                    dCc[0] = tp.dC_dv1;
                    dw[0] = kw * (3.0f * Fr[dx] * tp.w * sgnf(c0 - tp.v1));
                } else {
                    const float fw = Fr[dx] * tp.w;
                    dCc[0] = tp.dCv.x;
                    dCc[1] = tp.dCv.y;
                    dCc[2] = tp.dCv.z;
                    dw[0] = kw * (fw * sgnf(c0.x - tp.v1c.x));
                    dw[1] = kw * (fw * sgnf(c0.y - tp.v1c.y));
                    dw[2] = kw * (fw * sgnf(c0.z - tp.v1c.z));
                }
                const float dg[4] = {tp.dC_dg1.x, tp.dC_dg1.y, tp.dC_dg1.z, tp.dC_dg1.w};
                if (want1) {
                    float u[K];
#pragma unroll
                    for (int c = 0; c < NCC; ++c) {
                        u[c] = kC * dCc[c] + dw[c];
                        dcenter[c] -= dw[c];
                    }
#pragma unroll
                    for (int c = 0; c < 4; ++c) u[NCC + c] = kC * dg[c];
                    const int kk = dx + P;
                    const bool park = lane == kk, park2 = lane + kWave == kk;
#pragma unroll
                    for (int c = 0; c < K; ++c) {
                        if ((c < NCC && !dI1) || (c >= NCC && !dG1)) continue;
                        const float tsum = run[c] + u[c];
                        if (dx < P) {
                            const float sv =
                                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tsum), 0));
                            done[c] = park ? sv : done[c];
                            done2[c] = park2 ? sv : done2[c];
                            run[c] = lane_down(tsum);
                        } else {
                            run[c] = tsum;
                        }
                    }
                }
                if (want2) {
                    float e[K];
#pragma unroll
                    for (int c = 0; c < NCC; ++c) e[c] = -kC * dCc[c];
#pragma unroll
                    for (int c = 0; c < 4; ++c) e[NCC + c] = -kC * dg[c];
                    const float w0 = 1.0f - tp.fa, w1 = tp.fa;
                    const int i0 = tp.i0;
                    const bool same = on && i0 == pi, up = on && i0 == pi + 1;
                    const bool down = on && i0 == pi - 1, jump = on && !same && !up && !down;
                    put2(pi >= 0 && (up || jump), pi, p0);
                    put2(pi >= 0 && (down || jump), min(pi + 1, W - 1), p1);
#pragma unroll
                    for (int c = 0; c < K; ++c) {
                        const float u0 = e[c] * w0, u1 = e[c] * w1;
                        const float n0 = same ? p0[c] + u0 : up ? p1[c] + u0 : u0;
                        const float n1 = same ? p1[c] + u1 : down ? p0[c] + u1 : u1;
                        p0[c] = on ? n0 : p0[c];
                        p1[c] = on ? n1 : p1[c];
                    }
                    pi = on ? i0 : pi;
                }
            };
            // the loop holds wave-wide operations, so the compiler does not
            // unroll it: two taps per trip let one tap's gathers overlap the
            // other's LDS round trip
            step(eval(-P), -P);
            for (int dx = -P + 1; dx < P; dx += 2) {
                const Tap ta = eval(dx), tb = eval(dx + 1);
                step(ta, dx);
                step(tb, dx + 1);
            }
            if (want2) {
                put2(pi >= 0, pi, p0);
                put2(pi >= 0, min(pi + 1, W - 1), p1);
            }
            // this view's row: columns l and 64 + l (< 2P) parked, 2P + l running
            if (want1) {
                auto put1 = [&](bool on, int col, const float (&v)[K]) {
                    const int qx = c1 + col;
                    if (!on || (unsigned)qx >= (unsigned)W) return;
                    if (dI1)
                        for (int c = 0; c < NCC; ++c)
                            if (v[c] != 0.0f)
                                unsafeAtomicAdd(&dI1[(rowo + qx) * (unsigned)NC + c], v[c]);
                    if (dG1)
                        for (int c = 0; c < 4; ++c)
                            if (v[NCC + c] != 0.0f)
                                unsafeAtomicAdd(&dG1f[(rowo + qx) * 4u + c], v[NCC + c]);
                };
                put1(lane < 2 * P, lane, done);
                put1(kWave + lane < 2 * P, kWave + lane, done2);
                put1(true, 2 * P + lane, run);
            }
            if (lds2) {
                wave_lds_sync();
                // one atomic per texel and component
                if (dI2)
                    for (int c = 0; c < NCC; ++c)
                        for (int i = lane; i < n2; i += kWave) {
                            const float v = acc.v[c][i];
                            if (v != 0.0f)
                                unsafeAtomicAdd(&dI2[(rowo + s0 + i) * (unsigned)NC + c], v);
                        }
                if (dG2)  // 64 lanes add 64 consecutive floats
                    for (int f = lane; f < 4 * n2; f += kWave) {
                        const int i = f >> 2;
                        const float v = acc.v[NCC + (f & 3)][i];
                        if (v != 0.0f) unsafeAtomicAdd(&dG2f[(rowo + s0 + i) * 4u + (f & 3)], v);
                    }
            }
        }
        if (dI1 && ok)
            for (int c = 0; c < NCC; ++c)
                unsafeAtomicAdd(&dI1[(unsigned)idx * (unsigned)NC + c], dcenter[c]);
    }
}

// ------------------------------------------------------- gradxy backward ----
// dI += gradxy^T dG for the stencil of pmhstereo.cl:817-822 (gather form).
// The gray conversion of an intensity image has derivative 0.2126 + 0.7152 +
// 0.0722 = 1 and the zero border passes no gradient.
__global__ void gradxy_backward_kernel(const float4* __restrict__ dG, int W, int H,
                                       float* __restrict__ dI) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    // contribution of G(p) to g(q) for p = q - o, o = (ox, oy)
    auto at = [&](int px, int py) -> float4 {
        if (px < 0 || py < 0 || px >= W || py >= H) return make_float4(0.f, 0.f, 0.f, 0.f);
        return dG[py * W + px];
    };
    float acc = 0.0f;
    // o = (+1, 0): dx 0.5 ; (-1, 0): dx -0.5
    acc += 0.5f * at(x - 1, y).x - 0.5f * at(x + 1, y).x;
    // o = (0, +1): dy 0.5 ; (0, -1): dy -0.5
    acc += 0.5f * at(x, y - 1).y - 0.5f * at(x, y + 1).y;
    // o = (+1, +1): dx .25, dy .25, dxy 1 ; o = (-1, -1): dx -.25, dy -.25, dxy -1
    {
        const float4 a = at(x - 1, y - 1), b = at(x + 1, y + 1);
        acc += 0.25f * a.x + 0.25f * a.y + a.z;
        acc += -0.25f * b.x - 0.25f * b.y - b.z;
    }
    // o = (+1, -1): dx .25, dy -.25, dyx -1 ; o = (-1, +1): dx -.25, dy .25, dyx 1
    {
        const float4 a = at(x - 1, y + 1), b = at(x + 1, y - 1);
        acc += 0.25f * a.x - 0.25f * a.y - a.w;
        acc += -0.25f * b.x + 0.25f * b.y + b.w;
    }
    dI[y * W + x] += acc;
}

// ------------------------------------------------ gradient plane refinement --
// Directed refinement of each pixel's plane (used beside, or instead of, the
// random refinement of pmhstereo.cl:383-389).  In the local parameters
// (e = x'(x, y), a, b) the window gradient is well conditioned; a step moves
// x' by at most `radius` px anywhere in the window (|de| + p(|da| + |db|) =
// radius), and is kept with the reference's acceptance rule (cost plus theta
// times the smoothness distance, `<=`).  The radius halves every step.  The
// trial sweep also yields the gradient at the trial plane for the next step.
// Every lane of the wave runs the same number of sweeps (the window rows are
// staged for the whole wave); a lane whose plane has no gradient rides along
// on its own plane and discards the result, a lane without a pixel counts no
// taps.
template <int NC>
__global__ __launch_bounds__(kWave) void grad_refine_kernel(const RefineParams* __restrict__ pp,
                                                            Tiles T, const float* __restrict__ F,
                                                            int P, float theta, int steps,
                                                            float radius) {
    __shared__ GatherStageT<stage_narrow(NC), NC> st;
    for_tiles(T, [&](int item, int lt) {
        const RefineParams p = pp[item];
        const int W = p.W, H = p.H, lr = p.lr;
        const float depthmax = p.depthmax;
        const int nbx = (W + kWave - 1) / kWave;
        const int y = __builtin_amdgcn_readfirstlane(lt / nbx);
        const Seg sg = row_segment(lt % nbx, y, W, 0);
        const int x = sg.x;
        const int idx = y * W + x;
        const float cx = (float)x, cy = (float)y;
        const CostCtx cc = cost_ctx<NC>(p.I1, p.G1, p.g2, F, W, H, P);
        const auto c0 = own_colour<NC>(p.I1, idx);
        float4 self = p.states[idx];
        const float4 sp = p.smooth[idx];
        float sx = sp.x, sy = sp.y;
        const float nf = fmaxf(__builtin_sqrtf(sx * sx + sy * sy) + 0.001f, 1.0f);
        sx = sx / nf;
        sy = sy / nf;
        const float4 sm =
            make_float4(sx, sy, __builtin_sqrtf((1.0f - sx * sx) - sy * sy), sp.z * depthmax);
        const float dw = 1.0f / (depthmax * depthmax);
        CostGrad cg = plane_cost_grad_staged(cc, x, sg.active, y, c0, self, st);
        bool live = sg.active;
        float r = radius;
        for (int s = 0; s < steps; ++s, r *= 0.5f) {
            live = live && cg.ok;
            const float ge = cg.ge / cg.ws, ga = cg.gx / cg.ws, gb = cg.gy / cg.ws;
            const float norm = fabsf(ge) + (float)P * (fabsf(ga) + fabsf(gb));
            live = live && norm > 0.0f;
            if (!__ballot(live)) break;  // wave-uniform
            const float k = live ? r / norm : 0.0f;
            const float de = -k * ge, da = -k * ga, db = -k * gb;
            float4 trial = live ? make_float4(self.x + da, self.y + db,
                                              self.z + de - da * cx - db * cy, 0.0f)
                                : self;
            // lanes that are not live cost their own plane and discard it: counting
            // their taps keeps the wave on the fast path and changes no live lane.
            // The last step's trial needs no gradient: the value-only sweep gives
            // the same cost bits (same taps, same order).
            if (s == steps - 1) {
                trial.w = plane_cost(cc, x, sg.active, y, c0, trial, st);
                if (live && accept(trial, self, sm, cx, cy, dw, lr, theta)) self = trial;
                break;
            }
            const CostGrad tg = plane_cost_grad_staged(cc, x, sg.active, y, c0, trial, st);
            trial.w = tg.cost;
            if (live && accept(trial, self, sm, cx, cy, dw, lr, theta)) {
                self = trial;
                cg = tg;
            }
        }
        if (sg.active) p.states[idx] = self;
    });
}

// ============================================================== launchers ====
void launch_cost_backward(const float* I1, const float4* G1, Gather g2, const float* F, int W,
                          int H, int P, const float4* states, const float* gout, float4* dst,
                          float* dI1, float4* dG1, float* dI2, float4* dG2, int nc,
                          hipStream_t s) {
    const int nbx = (W + kWave - 1) / kWave, nb = nbx * H;
    auto go = [&](auto kernel) {
        kernel<<<xcd_grid(nb), kWave, 0, s>>>(I1, G1, g2, F, W, H, P, states, gout, dst, dI1, dG1,
                                               dI2, dG2, nbx, nb);
    };
    const bool images = dI1 || dG1 || dI2 || dG2;
    if (nc == kRGBA) {
        if (images) go(cost_backward_kernel<kRGBA, true>);
        else go(cost_backward_kernel<kRGBA, false>);
    } else {
        if (images) go(cost_backward_kernel<kGray, true>);
        else go(cost_backward_kernel<kGray, false>);
    }
    PMH_HIP(hipGetLastError());
}

void launch_gradxy_backward(const float4* dG, int W, int H, float* dI, hipStream_t s) {
    gradxy_backward_kernel<<<dim3((W + 63) / 64, (H + 3) / 4), dim3(64, 4), 0, s>>>(dG, W, H, dI);
    PMH_HIP(hipGetLastError());
}

void launch_grad_refine(const RefineParams* pp, Tiles T, const float* F, int P, float theta,
                        int steps, float radius, int nc, hipStream_t s) {
    if (nc == kRGBA)
        grad_refine_kernel<kRGBA><<<grid_of(grad_refine_kernel<kRGBA>, kWave, T), kWave, 0, s>>>(
            pp, T, F, P, theta, steps, radius);
    else
        grad_refine_kernel<kGray><<<grid_of(grad_refine_kernel<kGray>, kWave, T), kWave, 0, s>>>(
            pp, T, F, P, theta, steps, radius);
    PMH_HIP(hipGetLastError());
}

}  // namespace pmh
