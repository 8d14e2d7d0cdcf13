// pmh_kernels.hip -- gfx950 kernels of the slanted-plane PatchMatch stereo path.
//
// Kernel map (reference kernels in cvt/cl/kernel/pmhstereo.cl unless noted):
//   prep_kernel      -- pmhstereo_gradxy :785-831 + pmhstereo_weight :737-783
//                       + (new) the gather planes of an image as the other view
//   zero_kernel      -- (new) fill kernel, keeps launch sequences capturable
//   pyr_down_kernel, plane_up_kernel -- (new) coarse-to-fine start
//   init_kernel      -- pmhstereo_init          :230-248
//   cost_kernel      -- patch_eval_color_grad_weighted :170-228 over a plane field
//   propagate_kernel -- pmhstereo_propagate_view :286-404 (Jacobi, red or black)
//   viewbuf_kernel   -- the view scatter of :391-401, deterministic slot order
//   consistency_kernel -- pmhstereo_consistency :428-464
//   fill_state_kernel  -- pmhstereo_fill_state  :466-510 (row scan)
//   pd_step_kernel   -- PDHuberWeighted (PDHuberWeighted.cl:34-104) and
//                       PDHuberWeightedInpaint (PDHuberWeightedInpaint.cl:34-104)
//   occmap_kernel, normal_depth_kernel -- pmhstereo.cl:682-712, :663-680
//   fill_depth_kernel  -- pmhstereo_fill_depthmap / fill_normalmap :512-603
//
// Layout: every image is dense row-major (pitch = width).  A plane field is
// float4 (a, b, c, cost) with x' = a*x + b*y + c.  One lane owns one pixel;
// a wave covers 64 pixels of one row, so window rows are wave-uniform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

#include "pmh_device.hpp"
#include "pmh_internal.hpp"
#include "pmh_window.hpp"

namespace pmh {

// Every kernel of the depthMap pipeline takes an array of per-item parameters
// and a work list (Tiles) over all items, so one launch serves a whole batch
// of pairs; a single pair is a batch of one.

// ------------------------------------------------------------------ prep ----
// gray value dot(read_imagef(img), (0.2126, 0.7152, 0.0722, 0)) of
// pmhstereo.cl:758,807 with the zero border of CLK_ADDRESS_CLAMP, for a
// CL_INTENSITY (nc = 1) or CL_RGBA (nc = 4) image
__device__ __forceinline__ float gray_at(const float* __restrict__ img, int W, int H, int nc,
                                         int x, int y) {
    if (x < 0 || y < 0 || x >= W || y >= H) return 0.0f;
    if (nc == 1) {
        const float I = img[y * W + x];
        return ((I * 0.2126f + I * 0.7152f) + I * 0.0722f) + I * 0.0f;
    }
    const float4 v = reinterpret_cast<const float4*>(img)[y * W + x];
    return ((v.x * 0.2126f + v.y * 0.7152f) + v.z * 0.0722f) + v.w * 0.0f;
}

// pmhstereo_gradxy, pmhstereo.cl:785-831
__device__ __forceinline__ float4 gradxy_at(const float* __restrict__ img, int W, int H, int nc,
                                            int x, int y) {
    const float gm0 = gray_at(img, W, H, nc, x - 1, y), gp0 = gray_at(img, W, H, nc, x + 1, y);
    const float gmm = gray_at(img, W, H, nc, x - 1, y - 1);
    const float gpm = gray_at(img, W, H, nc, x + 1, y - 1);
    const float gmp = gray_at(img, W, H, nc, x - 1, y + 1);
    const float gpp = gray_at(img, W, H, nc, x + 1, y + 1);
    const float g0m = gray_at(img, W, H, nc, x, y - 1), g0p = gray_at(img, W, H, nc, x, y + 1);
    const float dx = ((gp0 - gm0) * 0.5f + (gpm - gmm) * 0.25f) + (gpp - gmp) * 0.25f;
    const float dy = ((g0p - g0m) * 0.5f + (gmp - gmm) * 0.25f) + (gpp - gpm) * 0.25f;
    return make_float4(dx, dy, gpp - gmm, gmp - gpm);
}

// pmhstereo_weight, pmhstereo.cl:737-783 (dy = g(x,y-1) - g(x,y-1) at :769)
__device__ __forceinline__ float weight_at(const float* __restrict__ img, int W, int H, int nc,
                                           int x, int y) {
    const float dx = gray_at(img, W, H, nc, x + 1, y) - gray_at(img, W, H, nc, x - 1, y);
    const float g = gray_at(img, W, H, nc, x, y - 1);
    const float dy = g - g;
    const float s = __builtin_sqrtf(dx * dx + dy * dy);
    const float t = (float)pow((double)s, (double)0.8f);
    const float m = -3.0f * t;
    const float e = (float)exp((double)m);
    return e + 0.0001f;
}

// 64x4-pixel tile of a 2-D kernel: thread -> pixel
__device__ __forceinline__ bool tile_pixel(int lt, int W, int H, int& x, int& y) {
    const int tx = (W + 63) / 64;
    x = (lt % tx) * 64 + ((int)threadIdx.x & 63);
    y = (lt / tx) * 4 + ((int)threadIdx.x >> 6);
    return x < W && y < H;
}

// Gradient image, Huber weight and the gather planes of one image (the planes
// hold G and its forward difference, recomputed at x+1 rather than read back).
__global__ __launch_bounds__(256) void prep_kernel(const PrepParams* __restrict__ pp, Tiles T) {
    for_tiles(T, [&](int i, int lt) {
        const PrepParams p = pp[i];
        int x, y;
        if (!tile_pixel(lt, p.W, p.H, x, y)) return;
        const int W = p.W, H = p.H, nc = p.nc, j = y * W + x;
        const float4 g0 = (p.grad || !p.gin) ? gradxy_at(p.img, W, H, nc, x, y) : float4{};
        if (p.grad) p.grad[j] = g0;
        if (p.weight) p.weight[j] = weight_at(p.img, W, H, nc, x, y);
        if (p.rec) {
            const size_t n = (size_t)W * H;
            const int x1 = x + 1 < W ? x + 1 : W - 1;
            const int j1 = y * W + x1;
            const float4 g0p = p.gin ? p.gin[j] : g0;
            const float4 g1 = p.gin ? p.gin[j1] : gradxy_at(p.img, W, H, nc, x1, y);
            p.rec[j] = g0p;
            p.rec[n + j] = make_float4(g1.x - g0p.x, g1.y - g0p.y, g1.z - g0p.z, g1.w - g0p.w);
            if (nc == 1) {
                const float t0 = p.img[j];
                reinterpret_cast<float2*>(p.rec + 2 * n)[j] = make_float2(t0, p.img[j1] - t0);
            } else {
                const float4* im = reinterpret_cast<const float4*>(p.img);
                const float4 t0 = im[j], t1 = im[j1];
                p.rec[2 * n + 2 * (size_t)j] = t0;
                p.rec[2 * n + 2 * (size_t)j + 1] =
                    make_float4(t1.x - t0.x, t1.y - t0.y, t1.z - t0.z, t1.w - t0.w);
            }
        }
    });
}

// ------------------------------------------------------------------ zero ----
// a fill kernel (not hipMemsetAsync) so that launch sequences stay capturable
__global__ __launch_bounds__(256) void zero_kernel(const ZeroParams* __restrict__ pp, Tiles T) {
    for_tiles(T, [&](int i, int lt) {
        const ZeroParams p = pp[i];
        const unsigned b = (unsigned)lt * kZeroTileWords;
        for (unsigned k = threadIdx.x; k < kZeroTileWords; k += 256)
            if (b + k < p.words) p.p[b + k] = 0u;
    });
}

// ------------------------------------------------------------- pyramid ----
// (new) coarse-to-fine start: 2x2 box reduction of the input images, and the
// plane field of a half-resolution level carried up.  A level-(l+1) pixel i
// covers level-l pixels 2i, 2i+1 (centre 2i + 1/2) and disparities double, so
// x' = a x + b y + c becomes x' = a x + b y + (2c + (1 - a - b) / 2).
__global__ __launch_bounds__(256) void pyr_down_kernel(const PyrParams* __restrict__ pp, Tiles T) {
    for_tiles(T, [&](int i, int lt) {
        const PyrParams p = pp[i];
        const int w = (p.W + 1) / 2, h = (p.H + 1) / 2;
        int x, y;
        if (!tile_pixel(lt, w, h, x, y)) return;
        const int x0 = 2 * x, x1 = min(2 * x + 1, p.W - 1);
        const int y0 = 2 * y, y1 = min(2 * y + 1, p.H - 1);
        for (int c = 0; c < p.nc; ++c) {
            const float a = p.in[(y0 * p.W + x0) * p.nc + c], b = p.in[(y0 * p.W + x1) * p.nc + c];
            const float e = p.in[(y1 * p.W + x0) * p.nc + c], d = p.in[(y1 * p.W + x1) * p.nc + c];
            p.out[(y * w + x) * p.nc + c] = 0.25f * ((a + b) + (e + d));
        }
    });
}

__global__ __launch_bounds__(256) void plane_up_kernel(const UpParams* __restrict__ pp, Tiles T) {
    for_tiles(T, [&](int i, int lt) {
        const UpParams p = pp[i];
        int x, y;
        if (!tile_pixel(lt, p.W, p.H, x, y)) return;
        const float4 s = p.in[min(y >> 1, p.h - 1) * p.w + min(x >> 1, p.w - 1)];
        p.out[y * p.W + x] = make_float4(s.x, s.y, 2.0f * s.z + 0.5f * ((1.0f - s.x) - s.y), 0.0f);
    });
}

// ------------------------------------------------------------------ init ----
// pmhstereo_init, pmhstereo.cl:230-248; with p.src the planes are given (a
// field carried up from a coarser level) and only their costs are computed.
template <int NC>
__global__ __launch_bounds__(kWave) void init_kernel(const InitParams* __restrict__ pp, Tiles T,
                                                     const float* __restrict__ F, int P,
                                                     SkipTable tab) {
    __shared__ GatherStageT<stage_narrow(NC), NC> st;
    for_tiles(T, [&](int i, int lt) {
        const InitParams p = pp[i];
        const int W = p.W, H = p.H;
        const int nbx = (W + kWave - 1) / kWave;
        const int y = __builtin_amdgcn_readfirstlane(lt / nbx);
        const Seg sg = row_segment(lt % nbx, y, W, 0);
        const int x = sg.x;
        const CostCtx cc = cost_ctx<NC>(p.I1, p.G1, p.g2, F, W, H, P);
        uint32_t rx, rc;
        mwc_seed(tab, (uint64_t)y * W + x, rx, rc);
        const float u0 = mwc_next_float(rx, rc);
        const float u1 = mwc_next_float(rx, rc);
        const float u2 = mwc_next_float(rx, rc);
        float4 s = p.src ? p.src[y * W + x]
                         : state_init(u0, u1, u2, (float)x, (float)y, p.lr, 1.0f, p.depthmax);
        s.w = plane_cost(cc, x, sg.active, y, own_colour<NC>(p.I1, y * W + x), s, st);
        if (sg.active) p.out[y * W + x] = s;
    });
}

// ------------------------------------------------------------------ cost ----
template <int NC>
__global__ __launch_bounds__(kWave) void cost_kernel(const float* __restrict__ I1,
                                                     const float4* __restrict__ G1, Gather g2,
                                                     const float* __restrict__ F, int W, int H,
                                                     int P, const float4* __restrict__ st,
                                                     float* __restrict__ cost_out, int nbx,
                                                     int nb) {
    __shared__ GatherStageT<stage_narrow(NC), NC> stage;
    const int t = xcd_tile(nb);
    if (t >= nb) return;
    const int y = __builtin_amdgcn_readfirstlane(t / nbx);
    const Seg sg = row_segment(t % nbx, y, W, 0);
    const CostCtx cc = cost_ctx<NC>(I1, G1, g2, F, W, H, P);
    const float c = plane_cost(cc, sg.x, sg.active, y, own_colour<NC>(I1, y * W + sg.x),
                               st[y * W + sg.x], stage);
    if (sg.active) cost_out[y * W + sg.x] = c;
}

// ------------------------------------------------------------- propagate ----
// smooth prior of pmhstereo.cl:320-323 -> (nx, ny, nz, depth)
__device__ __forceinline__ float4 smooth_prior(const float4 sp, float depthmax) {
    float sx = sp.x, sy = sp.y;
    const float nf = fmaxf(__builtin_sqrtf(sx * sx + sy * sy) + 0.001f, 1.0f);
    sx = sx / nf;
    sy = sy / nf;
    return make_float4(sx, sy, __builtin_sqrtf((1.0f - sx * sx) - sy * sy), sp.z * depthmax);
}

// Per-pixel state of one pass of pmhstereo_propagate_view (pmhstereo.cl:286-404)
// without the scatter: the pixel of this lane, its draws (:327, or supplied),
// current plane, smoothness prior and view-propagation slots.  kv (the wave's
// largest slot count) is wave-uniform.
struct PropPixel {
    int x, y, idx, nview, kv;
    int xt;       // the lane's column, outside [0, W) for a lane without a pixel (x: clamped)
    float cx, cy, dw;
    bool active;  // the lane has a pixel (costs count)
    bool owned;   // ... and the tile owns it (writes it)
    float u[kPropDraws];
    float4 self, sm;
};

// Tile t of a propagate launch (prop_tiles_x segments per row).  kShared:
// segments of prop_seg_pixels pixels whose first nv = 64 - prop_seg_pixels
// lanes re-hold the previous segment's last pixels (active, not owned), so that
// lane L + nv holds the pixel two columns right of lane L's; otherwise the
// 64-pixel segments of row_segment (a tile past the last one is empty).
template <bool kShared>
__device__ __forceinline__ PropPixel prop_pixel(const PropParams& p, int t,
                                                const SkipTable& tab) {
    PropPixel q;
    const int nbx = prop_tiles_x(p.W, p.sweep);
    q.y = __builtin_amdgcn_readfirstlane(t / nbx);
    const int bx = t % nbx;
    if constexpr (kShared) {
        const int seg = prop_seg_pixels(p.sweep);
        const int lx = bx * seg + lane_id() - (kWave - seg);
        q.xt = p.sweep == 0 ? lx : 2 * lx + ((q.y + (p.sweep - 1)) & 1);
        q.active = q.xt >= 0 && q.xt < p.W;
        q.owned = q.active && lane_id() >= kWave - seg;
        q.x = clampi(q.xt, 0, p.W - 1);
    } else {
        const Seg sg = row_segment(bx, q.y, p.W, p.sweep);
        q.x = sg.x;
        q.xt = sg.x;
        q.active = sg.active;
        q.owned = sg.active;
    }
    q.cx = (float)q.x;
    q.cy = (float)q.y;
    q.idx = q.y * p.W + q.x;
    if (p.draws) {
        const float* d = p.draws + (size_t)q.idx * kPropDraws;
#pragma unroll
        for (int i = 0; i < kPropDraws; ++i) q.u[i] = d[i];
    } else {
        uint32_t rx, rc;
        mwc_seed(tab, (uint64_t)q.idx + (uint64_t)p.iter, rx, rc);
#pragma unroll
        for (int i = 0; i < kPropDraws; ++i) q.u[i] = mwc_next_float(rx, rc);
    }
    q.self = p.old[q.idx];
    q.sm = smooth_prior(p.smooth[q.idx], p.depthmax);
    q.dw = 1.0f / (p.depthmax * p.depthmax);
    q.nview = 0;
    if (p.viewin) q.nview = min(p.viewin[q.idx].n, kViewSamples);
    if (!q.active) q.nview = 0;
    int kv = 0;
#pragma unroll
    for (int j = 1; j <= kViewSamples; ++j)
        if (__ballot(q.nview >= j)) kv = j;
    q.kv = __builtin_amdgcn_readfirstlane(kv);
    return q;
}

// Adds the pixels' candidate counts of the algorithm to p.cands (one atomic
// per wave): 16 (10 without the random neighbourhood) + the pixel's view
// candidates + the randomized refinements; lanes without a pixel count 0.
__device__ __forceinline__ void count_candidates(const PropParams& p, const PropPixel& q) {
    if (!p.cands) return;
    unsigned long long n = (unsigned long long)__builtin_popcountll(__ballot(q.owned)) *
                           (unsigned long long)((p.skip_nbhd ? 10 : 16) + p.random_refine);
#pragma unroll
    for (int j = 1; j <= kViewSamples; ++j)
        n += (unsigned long long)__builtin_popcountll(__ballot(q.owned && q.nview >= j));
    if (lane_id() == 0) atomicAdd(p.cands, n);
}

// The plane of neighbour (nx, ny) clamped to the image (:330-342, :355-361):
// in the black pass the red neighbours are already new.
__device__ __forceinline__ float4 prop_neighbour(const PropParams& p, int nx, int ny) {
    nx = clampi(nx, 0, p.W - 1);
    ny = clampi(ny, 0, p.H - 1);
    const int j = ny * p.W + nx;
    if (p.sweep == 2 && (((nx + ny) & 1) == 0)) return p.out[j];
    return p.old[j];
}

// Candidate k < 16 + kv, none of which depends on the folds before it;
// use = false for a view slot this lane lacks (costed, never folded).
__device__ __forceinline__ float4 prop_candidate(const PropParams& p, const PropPixel& q, int k,
                                                 bool& use) {
    const int W = p.W;
    auto cur = [&](int nx, int ny) -> float4 { return prop_neighbour(p, nx, ny); };
    use = true;
    if (k < 8) {  // spatial neighbours, :330-342 (row-major, centre skipped)
        const int n = k < 4 ? k : k + 1;
        return cur(q.x + n % 3 - 1, q.y + n / 3 - 1);
    }
    if (k == 8) {  // smooth plane, :346-348
        const float4 sm = q.sm;
        const float4 s4 = make_float4(1.0f - (-sm.x / sm.z), 0.0f - (-sm.y / sm.z),
                                      0.0f - ((sm.x * q.cx + sm.y * q.cy) / sm.z + sm.w), 0.0f);
        return p.lr ? s4 : viewprop(s4);
    }
    if (k < 15) {  // random neighbourhood, :355-361 (x offset drawn first)
        const int i = k - 9;
        float ux = 0.0f, uy = 0.0f;
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j == i) { ux = q.u[2 * j]; uy = q.u[2 * j + 1]; }
        const int ox = (int)(ux * 7.0f + 0.5f);
        const int oy = (int)(uy * 7.0f + 0.5f);
        return cur(q.x + ox, q.y + oy);
    }
    if (k == 15)  // random plane, :364
        return state_init(q.u[12], q.u[13], q.u[14], q.cx, q.cy, p.lr, 2.0f, p.depthmax);
    // view propagation, :372-379
    const int i = k - 16;
    use = i < q.nview;
    return use ? viewprop(p.viewsrc[q.y * W + p.viewin[q.idx].src[i]])
               : make_float4(1.0f, 0.0f, 0.0f, 0.0f);
}

// The random-neighbourhood candidates (k = 9 .. 14) that PropParams::skip_nbhd
// leaves out.
constexpr int kNbhd0 = 9, kNbhd1 = 15;

// Randomized refinement i (0 or 1) of plane s, :383-389.
__device__ __forceinline__ float4 prop_refine(const PropParams& p, const PropPixel& q, int i,
                                              const float4& s) {
    const float r0 = i ? q.u[18] : q.u[15], r1 = i ? q.u[19] : q.u[16], r2 = i ? q.u[20] : q.u[17];
    return refine(r0, r1, r2, s, q.cx, q.cy, p.depthmax, p.lr);
}

// The candidates are built, costed and folded one at a time in the reference
// order.  The stage loop is wave-uniform; a lane with fewer view candidates
// than the wave's maximum costs a placeholder it does not fold.
template <int kS, int NC>
__device__ __forceinline__ void propagate_tile(const PropParams& p, int t,
                                               const SkipTable& tab, GatherStageT<kS, NC>& st) {
    const PropPixel q = prop_pixel<false>(p, t, tab);
    if (!__ballot(q.active)) return;  // a tile past the row's last segment
    count_candidates(p, q);
    const CostCtx cc = cost_ctx<NC>(p.I1, p.G1, p.g2, p.F, p.W, p.H, p.P);
    const auto c0 = own_colour<NC>(p.I1, q.idx);
    float4 self = q.self;
    const int nind = 16 + q.kv;
    for (int k = 0; k < nind + p.random_refine; ++k) {
        if (p.skip_nbhd && k >= kNbhd0 && k < kNbhd1) continue;
        bool use = true;
        float4 c = k < nind ? prop_candidate(p, q, k, use) : prop_refine(p, q, k - nind, self);
        c.w = plane_cost(cc, q.x, q.active, q.y, c0, c, st);
        if (use && accept(c, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta)) self = c;
    }
    if (q.active) p.out[q.idx] = self;
}

template <int kS, int NC>
__global__ __launch_bounds__(kWave) void propagate_kernel(const PropParams* __restrict__ pp,
                                                          Tiles T, SkipTable tab) {
    __shared__ GatherStageT<kS, NC> st;
    for_tiles(T, [&](int i, int lt) {
        propagate_tile(pp[i], lt, tab, st);
    });
}

// The same pass with the candidates of one row segment spread over the four
// waves of a block (for small launches, where one wave per segment leaves
// most of the GPU idle): the 16 + kv candidates that do not depend on each
// other are costed four at a time, wave w taking k = w, w + 4, ...; wave 0
// then folds them in the reference order.  Of the randomized refinements the
// second depends on whether the first is taken, so waves 0-2 cost the first
// and both possible seconds at once and wave 0 folds.  Every cost is the one
// plane_cost gives the sequential pass and the folds run in the same order,
// so the result is the same bit for bit; the critical path is about 6 window
// sweeps instead of 20.  kMulti: a wave costs its 4 or 5 independent
// candidates together (plane_costs: one staged row, one load of this view's
// texels and one support weight per tap for all of them) instead of one
// window sweep each.
constexpr int kSplitWaves = 4;
template <int kS, int NC, bool kMulti>
__device__ __forceinline__ void propagate_split(const PropParams* __restrict__ pp, const Tiles& T,
                                                const SkipTable& tab) {
    __shared__ GatherStageT<kS, NC> st[kSplitWaves];
    __shared__ float cost[16 + kViewSamples][kWave];
    __shared__ float4 base[kWave];    // the plane after the folds of phase 1
    __shared__ float rcost[3][kWave];  // refinement costs: first, second if taken, second if not
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = lane_id();
    for_tiles(T, [&](int i, int lt) {
        const PropParams p = pp[i];
        const PropPixel q = prop_pixel<false>(p, lt, tab);
        if (!__ballot(q.active)) return;  // past the row's last segment: every wave of the block
        if (w == 0) count_candidates(p, q);
        const CostCtx cc = cost_ctx<NC>(p.I1, p.G1, p.g2, p.F, p.W, p.H, p.P);
        const auto c0 = own_colour<NC>(p.I1, q.idx);
        const int nind = 16 + q.kv;
        if constexpr (kMulti) {
            // candidates k = w + 4 j for the j in the wave-uniform mask J (N bits),
            // together; one call site per N keeps the kernel's code small
            auto costs = [&](auto nn, unsigned J) {
                constexpr int N = decltype(nn)::value;
                float4 c[N];
                float r[N];
                int ks[N];
#pragma unroll
                for (int m = 0; m < N; ++m) {
                    ks[m] = w + kSplitWaves * __builtin_ctz(J);
                    J &= J - 1u;
                    bool use;
                    c[m] = prop_candidate(p, q, ks[m], use);
                }
                plane_costs<kS, NC, N>(cc, q.x, q.active, q.y, c0, c, st[w], r);
#pragma unroll
                for (int m = 0; m < N; ++m) cost[ks[m]][lane] = r[m];
            };
            // 16 + kv candidates: wave w has a fifth (k = w + 16, j = 4) when w < kv.
            // The random plane (k = 15, wave 3) is often steep or far from the
            // others (for the right view viewprop makes a near-0 slope huge), which
            // would overflow the stage of the union, so it takes its own sweep.
            // Without the random neighbourhood (k = 9 .. 14) waves 1 and 2 drop
            // j = 2 and 3, wave 0 j = 3 and wave 3 j = 2.
            unsigned J = p.skip_nbhd ? (w == 0 ? 0x07u : 0x03u) : (w == 3 ? 0x07u : 0x0fu);
            if (w < q.kv) J |= 0x10u;
            switch (__builtin_popcount(J)) {
            case 2: costs(std::integral_constant<int, 2>{}, J); break;
            case 3: costs(std::integral_constant<int, 3>{}, J); break;
            case 4: costs(std::integral_constant<int, 4>{}, J); break;
            default: costs(std::integral_constant<int, 5>{}, J); break;
            }
            if (w == 3) {
                bool use;
                const float4 c = prop_candidate(p, q, 15, use);
                cost[15][lane] = plane_cost(cc, q.x, q.active, q.y, c0, c, st[w]);
            }
        } else {
            for (int k = w; k < nind; k += kSplitWaves) {
                if (p.skip_nbhd && k >= kNbhd0 && k < kNbhd1) continue;
                bool use;
                const float4 c = prop_candidate(p, q, k, use);
                cost[k][lane] = plane_cost(cc, q.x, q.active, q.y, c0, c, st[w]);
            }
        }
        __syncthreads();
        float4 self = q.self;
        if (w == 0) {
            for (int k = 0; k < nind; ++k) {
                if (p.skip_nbhd && k >= kNbhd0 && k < kNbhd1) continue;
                bool use;
                float4 c = prop_candidate(p, q, k, use);
                c.w = cost[k][lane];
                if (use && accept(c, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta)) self = c;
            }
            base[lane] = self;
        }
        __syncthreads();
        const int nref = p.random_refine;
        if (nref > 0 && w < 1 + 2 * (nref - 1)) {
            const float4 b = base[lane];
            const float4 r1 = prop_refine(p, q, 0, b);
            // wave 0: first refinement; wave 1: second after taking the first;
            // wave 2: second after keeping b
            const float4 c = w == 0 ? r1 : prop_refine(p, q, 1, w == 1 ? r1 : b);
            rcost[w][lane] = plane_cost(cc, q.x, q.active, q.y, c0, c, st[w]);
        }
        __syncthreads();
        if (w == 0) {
            if (nref > 0) {
                float4 c = prop_refine(p, q, 0, self);
                c.w = rcost[0][lane];
                const bool took = accept(c, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta);
                if (took) self = c;
                if (nref > 1) {
                    float4 c2 = prop_refine(p, q, 1, self);
                    c2.w = rcost[took ? 1 : 2][lane];
                    if (accept(c2, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta)) self = c2;
                }
            }
            if (q.active) p.out[q.idx] = self;
        }
        __syncthreads();  // the next tile reuses cost, base and rcost
    });
}

template <int kS, int NC, bool kMulti>
__global__ __launch_bounds__(kSplitWaves * kWave) void propagate_split_kernel(
    const PropParams* __restrict__ pp, Tiles T, SkipTable tab) {
    propagate_split<kS, NC, kMulti>(pp, T, tab);
}

// The shared-sweep four-wave kernel at a narrower stage (576 texels gray, 352
// RGBA: 53 KB per block, three blocks per CU) compiled for three waves per SIMD
// (the register allocator keeps it within 168 VGPRs, no spills).  Faster on
// launches with many row segments (64 uneven pairs at full resolution: 10%
// over the best one-wave shape); more rows overflow the narrower stage, so
// the autotuner chooses per shape class and view.
template <int kS, int NC>
__global__ __launch_bounds__(kSplitWaves * kWave) __attribute__((amdgpu_waves_per_eu(3, 3)))
void propagate_split3_kernel(const PropParams* __restrict__ pp, Tiles T, SkipTable tab) {
    propagate_split<kS, NC, true>(pp, T, tab);
}

// The four-wave pass with one LDS stage per block (block_costs, gray inputs):
// the waves sweep the window rows in lockstep on the union of their spans,
// staged once per row as records that carry the bilinear differences.
//
// A tile is a row segment of prop_seg_pixels pixels on 64 lanes: the first nv
// lanes re-hold the previous segment's last pixels (costed, not written), so
// that lane L + nv holds the pixel two columns right of lane L's.  Wave 3
// costs each lane's right, up-right and down-right neighbours (k = 4, 2, 7)
// over the columns of both windows (block_costs<kExt>): they are also the
// left, up-left and down-left neighbours (k = 3, 0, 5) of the pixel two
// columns right, whose costs go to lane L + nv; the first nv lanes reject
// those three (cost 1e5).  The other candidates -- k = 1, 6, 8, 9 .. 14 (not
// without the random neighbourhood), 15 and the kv view candidates -- are
// dealt round robin over waves 0-2 (at most 4 each; a 13th to wave 3 and a
// 14th to wave 0), so the waves stay balanced row by row.  The refinements
// are costed as in propagate_split (the first and both possible seconds on
// waves 0-2); folds and costs are those of plane_cost in the reference order:
// the same bits as every other shape.
// kShare false: no shared planes; the 16 + kv candidates are dealt round
// robin over the four waves (k = w, w + 4, ...; the 3-waves-per-SIMD build,
// whose 168 VGPRs do not hold the shared sweep's second set of sums without
// spilling).
constexpr int kExtWave = 3;
template <int kS, bool kShare>
__device__ __forceinline__ void propagate_block(const PropParams* __restrict__ pp, const Tiles& T,
                                                const SkipTable& tab) {
    static_assert(kSplitWaves == kSplitWavesStage, "one stage per four-wave block");
    static_assert(kExtWave < kSplitWaves, "the shared planes' wave is one of the block's");
    __shared__ BlockStage<kS> st;
    __shared__ int spans[2 * 2 * kSplitWaves];
    __shared__ float cost[16 + kViewSamples][kWave];
    __shared__ float4 base[kWave];    // the plane after the folds of the candidates
    __shared__ float rcost[3][kWave];  // refinement costs: first, second if taken, second if not
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = lane_id();
    for_tiles(T, [&](int i, int lt) {
        const PropParams p = pp[i];
        const PropPixel q = prop_pixel<true>(p, lt, tab);
        if (w == 0) count_candidates(p, q);
        const CostCtx cc = cost_ctx<kGray>(p.I1, p.G1, p.g2, p.F, p.W, p.H, p.P);
        const float c0 = own_colour<kGray>(p.I1, q.idx);
        const int nind = 16 + q.kv;
        const int nv = kWave - prop_seg_pixels(p.sweep);
        ExtLane e;
        e.xt = q.xt;
        e.own = q.active;
        e.nb = lane + nv < kWave && q.xt + 2 < p.W;
        e.c0n = own_colour<kGray>(p.I1, q.y * p.W + clampi(q.xt + 2, 0, p.W - 1));
        // m-th candidate costed at the lane's own pixel
        const int nnorm = (p.skip_nbhd ? 4 : 10) + q.kv;
        auto knorm = [&](int m) {
            if (m < 3) return m == 0 ? 1 : (m == 1 ? 6 : 8);
            return p.skip_nbhd ? (m == 3 ? 15 : m + 12) : m + 6;
        };
        // Without the random neighbourhood and with at most two view slots the
        // candidates dealt evenly (three per wave) keep the waves' rows shorter
        // than the shared planes' wave would (tile-uniform choice, same bits).
        const bool share = kShare && !(p.skip_nbhd && q.kv <= 2);
        // the wave with the shared planes rotates with the tile, so that the
        // blocks on a CU do not all put their heaviest wave on one SIMD
        const int role = share ? (w + lt) & (kSplitWaves - 1) : w;
        float4 pl[kBlockMaxN];
        int ks[kBlockMaxN];
        int n = 0;
        // no sharing: entry m of the candidate list k = 0 .. 15 (0 .. 8 and 15
        // without the random neighbourhood), then the view candidates
        const int nlist = (p.skip_nbhd ? 10 : 16) + q.kv;
        auto kof = [&](int m) {
            return p.skip_nbhd ? (m < 9 ? m : (m == 9 ? 15 : m + 6)) : m;
        };
#pragma unroll
        for (int j = 0; j < kBlockMaxN; ++j) {
            int m;
            bool has;
            if (share) {
                if (role == kExtWave) m = j == 3 ? 12 : -1;
                else m = j < 4 ? role + 3 * j : (role == 0 ? 13 : -1);
                has = m >= 0 && m < nnorm && (j == 4 || role == kExtWave || m < 12);
                ks[j] = has ? knorm(m) : 0;
            } else {
                m = w + kSplitWaves * j;
                has = m < nlist;
                ks[j] = has ? kof(m) : 0;
            }
            bool use;
            pl[j] = has ? prop_candidate(p, q, ks[j], use) : make_float4(1.f, 0.f, 0.f, 0.f);
            if (has) n = j + 1;
        }
        float r[kBlockMaxN], rn[kBlockMaxN];
        if (share && role == kExtWave) {
            pl[0] = prop_neighbour(p, q.xt + 1, q.y - 1);  // k = 2, up-right
            pl[1] = prop_neighbour(p, q.xt + 1, q.y);      // k = 4, right
            pl[2] = prop_neighbour(p, q.xt + 1, q.y + 1);  // k = 7, down-right
            ks[0] = 2;
            ks[1] = 4;
            ks[2] = 7;
            n = max(n, kExtPlanes);
            block_costs<kS, true>(cc, q.x, q.active, q.y, c0, pl, n, e, st, spans, r, rn);
            const int dst = (lane + nv) & (kWave - 1);
            const bool mine = lane + nv < kWave;
            cost[0][dst] = mine ? rn[0] : 1e5f;
            cost[3][dst] = mine ? rn[1] : 1e5f;
            cost[5][dst] = mine ? rn[2] : 1e5f;
        } else {
            block_costs<kS, false>(cc, q.x, q.active, q.y, c0, pl, n, e, st, spans, r, rn);
        }
#pragma unroll
        for (int j = 0; j < kBlockMaxN; ++j)
            if (j < n) cost[ks[j]][lane] = r[j];
        __syncthreads();
        float4 self = q.self;
        if (w == 0) {
            for (int k = 0; k < nind; ++k) {
                if (p.skip_nbhd && k >= kNbhd0 && k < kNbhd1) continue;
                bool use;
                float4 c = prop_candidate(p, q, k, use);
                c.w = cost[k][lane];
                if (use && accept(c, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta)) self = c;
            }
            base[lane] = self;
        }
        __syncthreads();
        const int nref = p.random_refine;
        if (nref > 0) {  // the same for every wave of the block
            float4 rp[kBlockMaxN];
#pragma unroll
            for (int j = 0; j < kBlockMaxN; ++j) rp[j] = make_float4(1.f, 0.f, 0.f, 0.f);
            const int rnum = w < 1 + 2 * (nref - 1) ? 1 : 0;
            if (rnum) {
                const float4 b = base[lane];
                const float4 r1 = prop_refine(p, q, 0, b);
                // wave 0: first refinement; wave 1: second after taking the first;
                // wave 2: second after keeping b
                rp[0] = w == 0 ? r1 : prop_refine(p, q, 1, w == 1 ? r1 : b);
            }
            float rr[kBlockMaxN], rrn[kBlockMaxN];
            block_costs<kS, false>(cc, q.x, q.active, q.y, c0, rp, rnum, e, st, spans, rr, rrn);
            if (rnum) rcost[w][lane] = rr[0];
        }
        __syncthreads();
        if (w == 0) {
            if (nref > 0) {
                float4 c = prop_refine(p, q, 0, self);
                c.w = rcost[0][lane];
                const bool took = accept(c, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta);
                if (took) self = c;
                if (nref > 1) {
                    float4 c2 = prop_refine(p, q, 1, self);
                    c2.w = rcost[took ? 1 : 2][lane];
                    if (accept(c2, self, q.sm, q.cx, q.cy, q.dw, p.lr, p.theta)) self = c2;
                }
            }
            if (q.owned) p.out[q.idx] = self;
        }
        __syncthreads();  // the next tile reuses cost, base, rcost, the stage and spans
    });
}

// Stage codes kStageBlock / kStageBlock3: one stage of 1792 records (70 KB,
// two blocks per CU, two waves per SIMD) or of 1152 records (45 KB, three
// blocks per CU, compiled for three waves per SIMD).
template <int kS>
__global__ __launch_bounds__(kSplitWaves * kWave) void propagate_block_kernel(
    const PropParams* __restrict__ pp, Tiles T, SkipTable tab) {
    propagate_block<kS, true>(pp, T, tab);
}
template <int kS>
__global__ __launch_bounds__(kSplitWaves * kWave) __attribute__((amdgpu_waves_per_eu(3, 3)))
void propagate_block3_kernel(const PropParams* __restrict__ pp, Tiles T, SkipTable tab) {
    propagate_block<kS, false>(pp, T, tab);
}

// --------------------------------------------------------------- viewbuf ----
// The scatter of pmhstereo.cl:391-401 with the slot order of a sequential run:
// per target pixel the count of sources and the first four sources by
// ascending source x.  One block per row; four rounds of LDS atomicMin pick
// the next-smallest unplaced source of every target.
__global__ __launch_bounds__(256) void viewbuf_kernel(const ViewParams* __restrict__ pp,
                                                      Tiles T) {
    __shared__ int cnt[kMaxRowWidth];
    __shared__ int mins[kMaxRowWidth];
    __shared__ int tgt[kMaxRowWidth];
    for_tiles(T, [&](int i, int y) {
        const ViewParams p = pp[i];
        const int W = p.W;
        __syncthreads();  // the previous row's LDS is no longer read
        for (int x = threadIdx.x; x < W; x += blockDim.x) {
            cnt[x] = 0;
            mins[x] = 0x7fffffff;
            const float xp = transform_x(p.states[y * W + x], (float)x, (float)y);
            const float tt = xp + 0.5f;
            tgt[x] = (tt > -1.0f && tt < (float)W) ? (int)tt : -1;
        }
        __syncthreads();
        for (int x = threadIdx.x; x < W; x += blockDim.x)
            if (tgt[x] >= 0) atomicAdd(&cnt[tgt[x]], 1);
        __syncthreads();
        for (int r = 0; r < kViewSamples; ++r) {
            for (int x = threadIdx.x; x < W; x += blockDim.x)
                if (tgt[x] >= 0) atomicMin(&mins[tgt[x]], x);
            __syncthreads();
            for (int x = threadIdx.x; x < W; x += blockDim.x) {
                const int tg = tgt[x];
                if (tg >= 0 && mins[tg] == x) {
                    p.out[y * W + tg].src[r] = x;
                    tgt[x] = -1 - tg;  // placed
                }
            }
            __syncthreads();
            for (int x = threadIdx.x; x < W; x += blockDim.x) mins[x] = 0x7fffffff;
            __syncthreads();
        }
        for (int x = threadIdx.x; x < W; x += blockDim.x) p.out[y * W + x].n = cnt[x];
    });
}

// ----------------------------------------------------------- consistency ----
// The L/R test of pmhstereo_consistency (pmhstereo.cl:428-464) and
// pmhstereo_occmap (:682-712) at (x, y): true if the plane passes.
__device__ __forceinline__ bool lr_check(const float4* __restrict__ left,
                                         const float4* __restrict__ right, int W, int x, int y,
                                         float maxdisp, float maxdeg, int lr) {
    const float4 sl = left[y * W + x];
    const float xp = transform_x(sl, (float)x, (float)y);
    float4 sr;
    if (!(xp >= 0.0f && xp < (float)W)) {
        sr = make_float4(1e5f, 1e5f, 1e5f, 1e5f);
    } else {
        const int xr = clampi((int)roundf(xp), 0, W - 1);
        sr = right[y * W + xr];
    }
    float3 n1, n2;
    if (lr) {
        n1 = to_normal(sl);
        n2 = to_normal(viewprop(sr));
    } else {
        n1 = to_normal(sr);
        n2 = to_normal(viewprop(sl));
    }
    const float ndiff = (n1.x * n2.x + n1.y * n2.y) + n1.z * n2.z;
    const float dmax = fabsf((float)x - transform_x(sr, xp, (float)y));
    const float ang = (float)(acos((double)ndiff) / 3.14159265358979323846);
    return !(dmax >= maxdisp || ang >= maxdeg / 180.0f);
}

// pmhstereo_consistency, pmhstereo.cl:428-464
__global__ __launch_bounds__(256) void consistency_kernel(const ConsParams* __restrict__ pp,
                                                          Tiles T, float maxdisp, float maxdeg) {
    for_tiles(T, [&](int i, int lt) {
        const ConsParams p = pp[i];
        int x, y;
        if (!tile_pixel(lt, p.W, p.H, x, y)) return;
        const int j = y * p.W + x;
        p.out[j] = lr_check(p.left, p.right, p.W, x, y, maxdisp, maxdeg, p.lr)
                       ? p.left[j]
                       : make_float4(0.f, 0.f, 0.f, 0.f);
    });
}

// pmhstereo_occmap, pmhstereo.cl:682-712: the consistency test (normals
// within maxdeg; the kernel hard-codes 5 degrees) written as (v, v, v, 1).
__global__ __launch_bounds__(256) void occmap_kernel(const ConsParams* __restrict__ pp, Tiles T,
                                                     float maxdisp, float maxdeg) {
    for_tiles(T, [&](int i, int lt) {
        const ConsParams p = pp[i];
        int x, y;
        if (!tile_pixel(lt, p.W, p.H, x, y)) return;
        const float v = lr_check(p.left, p.right, p.W, x, y, maxdisp, maxdeg, p.lr) ? 1.0f : 0.0f;
        p.out[y * p.W + x] = make_float4(v, v, v, 1.0f);
    });
}

// pmhstereo_normal_depth, pmhstereo.cl:663-680: (nx, ny, d / depthmax, 1)
__global__ __launch_bounds__(256) void normal_depth_kernel(const FillParams* __restrict__ pp,
                                                           Tiles T) {
    for_tiles(T, [&](int i, int lt) {
        const FillParams p = pp[i];
        int x, y;
        if (!tile_pixel(lt, p.W, p.H, x, y)) return;
        const float4 nd = to_ref_nd(p.in[y * p.W + x], (float)x, (float)y, p.lr);
        p.out[y * p.W + x] = make_float4(nd.x, nd.y, nd.w / p.depthmax, 1.0f);
    });
}

// ------------------------------------------------------------ row scans -----
// Per row: nearest valid pixel strictly left / right of every pixel, by a
// block-wide max / min scan over the row kept in LDS.
template <bool kLen4>
__device__ __forceinline__ void row_nearest_valid(const float4* __restrict__ in, int W, int y,
                                                  int* lidx, int* ridx) {
    __shared__ int partl[256], partr[256];
    const int tid = threadIdx.x;
    const int chunk = (W + 255) / 256;
    const int x0 = tid * chunk, x1 = min(W, x0 + chunk);
    int run = -1;
    for (int x = x0; x < x1; ++x) {
        const float4 v = in[y * W + x];
        const bool ok = (kLen4 ? len4(v) : len3(v)) >= 1e-1f;
        lidx[x] = ok ? x : -1;
        ridx[x] = ok ? x : 0x7fffffff;
        run = ok ? x : run;
    }
    partl[tid] = run;
    int rrun = 0x7fffffff;
    for (int x = x1 - 1; x >= x0; --x) rrun = min(rrun, ridx[x]);
    partr[tid] = rrun;
    __syncthreads();
    // inclusive max-scan of partl (left to right), min-scan of partr (right to left)
    for (int off = 1; off < 256; off <<= 1) {
        const int a = tid >= off ? partl[tid - off] : -1;
        const int b = tid + off < 256 ? partr[tid + off] : 0x7fffffff;
        __syncthreads();
        partl[tid] = max(partl[tid], a);
        partr[tid] = min(partr[tid], b);
        __syncthreads();
    }
    int carry = tid > 0 ? partl[tid - 1] : -1;
    for (int x = x0; x < x1; ++x) {
        carry = max(carry, lidx[x]);
        lidx[x] = carry;
    }
    int rcarry = tid < 255 ? partr[tid + 1] : 0x7fffffff;
    for (int x = x1 - 1; x >= x0; --x) {
        rcarry = min(rcarry, ridx[x]);
        ridx[x] = rcarry;
    }
    __syncthreads();
}

// pmhstereo_fill_state, pmhstereo.cl:466-510
__global__ __launch_bounds__(256) void fill_state_kernel(const FillParams* __restrict__ pp,
                                                         Tiles T) {
    __shared__ int lidx[kMaxRowWidth], ridx[kMaxRowWidth];
    for_tiles(T, [&](int i, int y) {
        const FillParams p = pp[i];
        const int W = p.W;
        row_nearest_valid<true>(p.in, W, y, lidx, ridx);
        for (int x = threadIdx.x; x < W; x += blockDim.x) {
            const float4 s = p.in[y * W + x];
            float4 val;
            if (len4(s) < 1e-1f) {
                const int li = x > 0 ? lidx[x - 1] : -1;
                const int ri = x + 1 < W ? ridx[x + 1] : 0x7fffffff;
                float4 l = li >= 0 ? p.in[y * W + li] : make_float4(1e5f, 1e5f, 1e5f, 1e5f);
                float4 r = ri < W ? p.in[y * W + ri] : make_float4(1e5f, 1e5f, 1e5f, 1e5f);
                l.w = fabsf(transform_x(l, (float)x, (float)y) - (float)x);
                r.w = fabsf(transform_x(r, (float)x, (float)y) - (float)x);
                val = l.w < r.w ? l : r;
            } else {
                val = s;
            }
            const float4 nd = to_ref_nd(val, (float)x, (float)y, p.lr);
            p.out[y * W + x] = make_float4(nd.x, nd.y, nd.w / p.depthmax, 1.0f);
        }
        __syncthreads();  // row_nearest_valid's LDS is reused by the next row
    });
}

// pmhstereo_fill_depthmap / pmhstereo_fill_normalmap, pmhstereo.cl:512-603
__global__ __launch_bounds__(256) void fill_depth_kernel(const FillDepthParams* __restrict__ pp,
                                                         Tiles T, int normals) {
    __shared__ int lidx[kMaxRowWidth], ridx[kMaxRowWidth];
    for_tiles(T, [&](int i, int y) {
        const FillDepthParams p = pp[i];
        const int W = p.W;
        row_nearest_valid<false>(p.in, W, y, lidx, ridx);
        for (int x = threadIdx.x; x < W; x += blockDim.x) {
            const float4 s = p.in[y * W + x];
            const float fx = (float)x, fy = (float)y;
            if (len3(s) < 1e-1f) {
                const int li = x > 0 ? lidx[x - 1] : -1;
                const int ri = x + 1 < W ? ridx[x + 1] : 0x7fffffff;
                float4 l = li >= 0 ? p.in[y * W + li] : make_float4(-1e5f, -1e5f, -1e5f, 0.f);
                float4 r = ri < W ? p.in[y * W + ri] : make_float4(-1e5f, -1e5f, -1e5f, 0.f);
                l.w = -(transform_x(l, fx, fy) - fx);
                r.w = -(transform_x(r, fx, fy) - fx);
                if (normals) {
                    const float3 n = l.w < r.w ? to_normal(l) : to_normal(r);
                    p.nmap[y * W + x] = make_float4(n.x, n.y, n.z, 0.0f);
                } else {
                    p.dmap[y * W + x] = fminf(l.w, r.w) * p.scale;
                }
            } else {
                if (normals) {
                    const float3 n = to_normal(s);
                    p.nmap[y * W + x] = make_float4(n.x, n.y, n.z, 0.0f);
                } else {
                    p.dmap[y * W + x] = fabsf(transform_x(s, fx, fy) - fx) * p.scale;
                }
            }
        }
        __syncthreads();
    });
}

// --------------------------------------------------------------- PD step ----
// PDHuberWeighted.cl:34-104: dual ascent at the pixel and at its left and top
// neighbours (recomputed, as the reference recomputes its tile halo), then
// the primal step with over-relaxation.  Channel 3 of the primal is 1; the
// dual's channel 3 feeds only that channel and is kept at 0.
struct Dual {
    float lo0, lo1, lo2, hi0, hi1, hi2;
};

// The dual ascent at one pixel from its primal uc, right and lower
// neighbours ux, uy (clamped to the image), previous dual (pl, ph) and Huber
// weight.  kPerChannel: the projection of PDHuberWeightedInpaint.cl:76 (each
// channel alone) instead of PDHuberWeighted.cl:75-76 (x and y jointly, z
// alone).
template <bool kPerChannel>
__device__ __forceinline__ Dual pd_dual_v(const float3 uc, const float3 ux, const float3 uy,
                                          const float3 pl, const float3 ph, float weight) {
    Dual d;
    const float sigma = 1.0f / __builtin_sqrtf(8.0f);  // SIGMA, PDHuberWeighted.cl:29
    const float den = 1.0f + sigma * 0.001f / weight;
    d.lo0 = (pl.x + sigma * (weight * (uc.x - ux.x))) / den;
    d.lo1 = (pl.y + sigma * (weight * (uc.y - ux.y))) / den;
    d.lo2 = (pl.z + sigma * (weight * (uc.z - ux.z))) / den;
    d.hi0 = (ph.x + sigma * (weight * (uc.x - uy.x))) / den;
    d.hi1 = (ph.y + sigma * (weight * (uc.y - uy.y))) / den;
    d.hi2 = (ph.z + sigma * (weight * (uc.z - uy.z))) / den;
    const float t0 = d.lo0 * d.lo0 + d.hi0 * d.hi0;
    const float t1 = d.lo1 * d.lo1 + d.hi1 * d.hi1;
    const float t2 = d.lo2 * d.lo2 + d.hi2 * d.hi2;
    if (kPerChannel) {
        const float p0 = fmaxf(1.0f, __builtin_sqrtf(t0));
        const float p1 = fmaxf(1.0f, __builtin_sqrtf(t1));
        const float p2 = fmaxf(1.0f, __builtin_sqrtf(t2));
        d.lo0 /= p0; d.hi0 /= p0;
        d.lo1 /= p1; d.hi1 /= p1;
        d.lo2 /= p2; d.hi2 /= p2;
        return d;
    }
    const float pxy = fmaxf(1.0f, __builtin_sqrtf(t0 + t1));
    const float pz = fmaxf(1.0f, __builtin_sqrtf(t2));
    d.lo0 /= pxy; d.hi0 /= pxy;
    d.lo1 /= pxy; d.hi1 /= pxy;
    d.lo2 /= pz;  d.hi2 /= pz;
    return d;
}

__device__ __forceinline__ float3 xyz(const float4 v) { return make_float3(v.x, v.y, v.z); }

// The dual at (gx, gy) read from global memory; exactly 0 outside the image's
// left and top edges (weight 0 there, PDHuberWeighted.cl:63-73).
template <bool kPerChannel>
__device__ __forceinline__ Dual pd_dual(const float4* __restrict__ last,
                                        const float4* __restrict__ plast,
                                        const float* __restrict__ w, int W, int H, int gx,
                                        int gy) {
    if (gx < 0 || gy < 0) {
        Dual d;
        d.lo0 = d.lo1 = d.lo2 = d.hi0 = d.hi1 = d.hi2 = 0.0f;
        return d;
    }
    const int x1 = gx + 1 < W ? gx + 1 : W - 1;
    const int y1 = gy + 1 < H ? gy + 1 : H - 1;
    return pd_dual_v<kPerChannel>(xyz(last[gy * W + gx]), xyz(last[gy * W + x1]),
                                  xyz(last[y1 * W + gx]), xyz(plast[2 * (gy * W + gx)]),
                                  xyz(plast[2 * (gy * W + gx) + 1]), w[gy * W + gx]);
}

// The primal step with over-relaxation at one pixel from its previous primal
// ul, data f, weight and the duals at the pixel (c), its left (l) and top (u)
// neighbours (PDHuberWeighted.cl:78-104).
__device__ __forceinline__ float3 pd_primal(const float3 ul, const float3 fi, float wt,
                                            float lambda, const Dual& c, const Dual& l,
                                            const Dual& u) {
    const float tau = 1.0f / __builtin_sqrtf(8.0f);  // TAU, PDHuberWeighted.cl:30
    const float den = 1.0f + tau * lambda;
    const float div0 = ((c.lo0 - l.lo0) + c.hi0) - u.hi0;
    const float div1 = ((c.lo1 - l.lo1) + c.hi1) - u.hi1;
    const float div2 = ((c.lo2 - l.lo2) + c.hi2) - u.hi2;
    float n0 = (ul.x + tau * (lambda * fi.x - wt * div0)) / den;
    float n1 = (ul.y + tau * (lambda * fi.y - wt * div1)) / den;
    float n2 = (ul.z + tau * (lambda * fi.z - wt * div2)) / den;
    n0 += 0.5f * (n0 - ul.x);
    n1 += 0.5f * (n1 - ul.y);
    n2 += 0.5f * (n2 - ul.z);
    return make_float3(n0, n1, n2);
}

// kInpaint: PDHuberWeightedInpaint.cl:34-104 -- per-channel projection and
// the data term's lambda scaled by the mask's first channel at the pixel.
template <bool kInpaint>
__global__ __launch_bounds__(256) void pd_step_kernel(const PdParams* __restrict__ pp, Tiles T,
                                                      float lambda0) {
    for_tiles(T, [&](int i, int lt) {
        const PdParams p = pp[i];
        int x, y;
        if (!tile_pixel(lt, p.W, p.H, x, y)) return;
        const int W = p.W, H = p.H;
        const Dual c = pd_dual<kInpaint>(p.last, p.plast, p.w, W, H, x, y);
        const Dual l = pd_dual<kInpaint>(p.last, p.plast, p.w, W, H, x - 1, y);
        const Dual u = pd_dual<kInpaint>(p.last, p.plast, p.w, W, H, x, y - 1);
        const int j = y * W + x;
        const float lambda = kInpaint ? lambda0 * p.mask[j].x : lambda0;
        const float3 n = pd_primal(xyz(p.last[j]), xyz(p.f[j]), p.w[j], lambda, c, l, u);
        p.out[j] = make_float4(n.x, n.y, n.z, 1.0f);
        p.pout[2 * j] = make_float4(c.lo0, c.lo1, c.lo2, 0.0f);
        p.pout[2 * j + 1] = make_float4(c.hi0, c.hi1, c.hi2, 0.0f);
    });
}

// `steps` (<= kPdMaxSteps) consecutive steps of pd_step_kernel in one launch.
// A block owns a kPdTileW x kPdTileH output tile and holds the primal and
// dual of the tile plus a halo of kPdMaxSteps pixels in LDS; step s updates
// the tile grown by (steps - s) pixels (each step consumes one pixel of halo
// per side), so the tile's final values are those of `steps` single steps,
// bit for bit: every pixel's dual and primal come from the same functions on
// the same inputs.  Each dual is computed once per step (the single step
// recomputes the left and top neighbours' duals), and HBM is read and written
// once per launch.
constexpr int kPdRegionW = kPdTileW + 2 * kPdMaxSteps;
constexpr int kPdRegionH = kPdTileH + 2 * kPdMaxSteps;
constexpr int kPdRegion = kPdRegionW * kPdRegionH;
constexpr int kPdThreads = 256;
constexpr int kPdOwn = (kPdRegion + kPdThreads - 1) / kPdThreads;  // region pixels per thread

// One output tile (lt) of `steps` fused steps; su / sp: the block's LDS.
template <bool kInpaint>
__device__ __forceinline__ void pd_steps_tile(const PdParams& p, int lt, float lambda0, int steps,
                                              float (&su)[3][kPdRegion],
                                              float (&sp)[6][kPdRegion]) {
    {
        const int W = p.W, H = p.H;
        const int ntx = (W + kPdTileW - 1) / kPdTileW;
        // global coordinates of region pixel (0, 0)
        const int ox = (lt % ntx) * kPdTileW - kPdMaxSteps;
        const int oy = (lt / ntx) * kPdTileH - kPdMaxSteps;
        float wt[kPdOwn], lam[kPdOwn];
        float3 fi[kPdOwn];
        __syncthreads();  // the previous tile's LDS is no longer read
#pragma unroll
        for (int k = 0; k < kPdOwn; ++k) {
            const int r = (int)threadIdx.x + k * kPdThreads;
            const int gx = ox + r % kPdRegionW, gy = oy + r / kPdRegionW;
            wt[k] = lam[k] = 0.0f;
            fi[k] = make_float3(0.0f, 0.0f, 0.0f);
            if (r < kPdRegion && gx >= 0 && gy >= 0 && gx < W && gy < H) {
                const int j = gy * W + gx;
                const float4 u = p.last[j];
                const float4 pl = p.plast[2 * j], ph = p.plast[2 * j + 1];
                su[0][r] = u.x; su[1][r] = u.y; su[2][r] = u.z;
                sp[0][r] = pl.x; sp[1][r] = pl.y; sp[2][r] = pl.z;
                sp[3][r] = ph.x; sp[4][r] = ph.y; sp[5][r] = ph.z;
                wt[k] = p.w[j];
                fi[k] = xyz(p.f[j]);
                lam[k] = kInpaint ? lambda0 * p.mask[j].x : lambda0;
            }
        }
        __syncthreads();
        auto U = [&](int r) { return make_float3(su[0][r], su[1][r], su[2][r]); };
        auto D = [&](int r) {
            Dual d;
            d.lo0 = sp[0][r]; d.lo1 = sp[1][r]; d.lo2 = sp[2][r];
            d.hi0 = sp[3][r]; d.hi1 = sp[4][r]; d.hi2 = sp[5][r];
            return d;
        };
        for (int s = 1; s <= steps; ++s) {
            const int m = steps - s;  // this step updates the tile grown by m
            const int ax0 = kPdMaxSteps - m, ax1 = kPdMaxSteps + kPdTileW + m;  // [ax0, ax1)
            const int ay0 = kPdMaxSteps - m, ay1 = kPdMaxSteps + kPdTileH + m;
            // duals of the updated pixels and of their left and top neighbours
            Dual d[kPdOwn];
            bool dn[kPdOwn];
#pragma unroll
            for (int k = 0; k < kPdOwn; ++k) {
                const int r = (int)threadIdx.x + k * kPdThreads;
                const int lx = r % kPdRegionW, ly = r / kPdRegionW;
                const int gx = ox + lx, gy = oy + ly;
                dn[k] = r < kPdRegion && lx >= ax0 - 1 && lx < ax1 && ly >= ay0 - 1 && ly < ay1 &&
                        gx >= 0 && gy >= 0 && gx < W && gy < H;
                if (dn[k]) {
                    const int rx = gx + 1 < W ? r + 1 : r;
                    const int ry = gy + 1 < H ? r + kPdRegionW : r;
                    d[k] = pd_dual_v<kInpaint>(U(r), U(rx), U(ry),
                                               make_float3(sp[0][r], sp[1][r], sp[2][r]),
                                               make_float3(sp[3][r], sp[4][r], sp[5][r]), wt[k]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kPdOwn; ++k) {
                if (!dn[k]) continue;
                const int r = (int)threadIdx.x + k * kPdThreads;
                sp[0][r] = d[k].lo0; sp[1][r] = d[k].lo1; sp[2][r] = d[k].lo2;
                sp[3][r] = d[k].hi0; sp[4][r] = d[k].hi1; sp[5][r] = d[k].hi2;
            }
            __syncthreads();
            float3 nu[kPdOwn];
            bool un[kPdOwn];
#pragma unroll
            for (int k = 0; k < kPdOwn; ++k) {
                const int r = (int)threadIdx.x + k * kPdThreads;
                const int lx = r % kPdRegionW, ly = r / kPdRegionW;
                const int gx = ox + lx, gy = oy + ly;
                un[k] = r < kPdRegion && lx >= ax0 && lx < ax1 && ly >= ay0 && ly < ay1 &&
                        gx >= 0 && gy >= 0 && gx < W && gy < H;
                if (un[k]) {
                    Dual zero;
                    zero.lo0 = zero.lo1 = zero.lo2 = zero.hi0 = zero.hi1 = zero.hi2 = 0.0f;
                    const Dual l = gx > 0 ? D(r - 1) : zero;
                    const Dual u = gy > 0 ? D(r - kPdRegionW) : zero;
                    nu[k] = pd_primal(U(r), fi[k], wt[k], lam[k], D(r), l, u);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kPdOwn; ++k) {
                if (!un[k]) continue;
                const int r = (int)threadIdx.x + k * kPdThreads;
                su[0][r] = nu[k].x; su[1][r] = nu[k].y; su[2][r] = nu[k].z;
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kPdOwn; ++k) {
            const int r = (int)threadIdx.x + k * kPdThreads;
            const int lx = r % kPdRegionW, ly = r / kPdRegionW;
            const int gx = ox + lx, gy = oy + ly;
            if (r < kPdRegion && lx >= kPdMaxSteps && lx < kPdMaxSteps + kPdTileW &&
                ly >= kPdMaxSteps && ly < kPdMaxSteps + kPdTileH && gx < W && gy < H) {
                const int j = gy * W + gx;
                p.out[j] = make_float4(su[0][r], su[1][r], su[2][r], 1.0f);
                p.pout[2 * j] = make_float4(sp[0][r], sp[1][r], sp[2][r], 0.0f);
                p.pout[2 * j + 1] = make_float4(sp[3][r], sp[4][r], sp[5][r], 0.0f);
            }
        }
    }
}

template <bool kInpaint>
__global__ __launch_bounds__(kPdThreads) void pd_steps_kernel(const PdParams* __restrict__ pp,
                                                              Tiles T, float lambda0, int steps) {
    __shared__ float su[3][kPdRegion];  // primal
    __shared__ float sp[6][kPdRegion];  // dual (lo0..2, hi0..2)
    for_tiles(T, [&](int i, int lt) { pd_steps_tile<kInpaint>(pp[i], lt, lambda0, steps, su, sp); });
}

// The same for one image whose parameters come by value (no parameter array
// in device memory): one tile per block, XCD-contiguous.
__global__ __launch_bounds__(kPdThreads) void pd_steps_one_kernel(PdParams p, int tiles,
                                                                  float lambda0, int steps) {
    __shared__ float su[3][kPdRegion];
    __shared__ float sp[6][kPdRegion];
    const int lt = xcd_tile(tiles);
    if (lt >= tiles) return;  // the same for every thread of the block
    pd_steps_tile<false>(p, lt, lambda0, steps, su, sp);
}

// Clears `words` 32-bit words (a fill kernel, parameters by value).
__global__ __launch_bounds__(256) void zero_one_kernel(uint32_t* __restrict__ p, unsigned words) {
    for (unsigned k = blockIdx.x * 256u + threadIdx.x; k < words; k += gridDim.x * 256u) p[k] = 0u;
}

__global__ void rng_draws_kernel(int W, int H, int iter, int count, float* __restrict__ out,
                                 SkipTable tab) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    uint32_t rx, rc;
    mwc_seed(tab, (uint64_t)y * W + x + (uint64_t)iter, rx, rc);
    float* o = out + (size_t)(y * W + x) * count;
    for (int i = 0; i < count; ++i) o[i] = mwc_next_float(rx, rc);
}

// ============================================================== launchers ====
void launch_prep(const PrepParams* pp, Tiles T, hipStream_t s) {
    prep_kernel<<<grid_of(prep_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_zero(const ZeroParams* pp, Tiles T, hipStream_t s) {
    zero_kernel<<<grid_of(zero_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_pyr_down(const PyrParams* pp, Tiles T, hipStream_t s) {
    pyr_down_kernel<<<grid_of(pyr_down_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_plane_up(const UpParams* pp, Tiles T, hipStream_t s) {
    plane_up_kernel<<<grid_of(plane_up_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_init(const InitParams* pp, Tiles T, const float* F, int P, int nc, hipStream_t s) {
    if (nc == kRGBA)
        init_kernel<kRGBA><<<grid_of(init_kernel<kRGBA>, kWave, T), kWave, 0, s>>>(
            pp, T, F, P, skip_table(kInitStride));
    else
        init_kernel<kGray><<<grid_of(init_kernel<kGray>, kWave, T), kWave, 0, s>>>(
            pp, T, F, P, skip_table(kInitStride));
    PMH_HIP(hipGetLastError());
}

void launch_cost(const float* I1, const float4* G1, Gather g2, const float* F, int W, int H,
                 int P, const float4* states, float* cost, int nc, hipStream_t s) {
    const int nbx = (W + kWave - 1) / kWave, nb = nbx * H;
    if (nc == kRGBA)
        cost_kernel<kRGBA><<<xcd_grid(nb), kWave, 0, s>>>(I1, G1, g2, F, W, H, P, states, cost,
                                                          nbx, nb);
    else
        cost_kernel<kGray><<<xcd_grid(nb), kWave, 0, s>>>(I1, G1, g2, F, W, H, P, states, cost,
                                                          nbx, nb);
    PMH_HIP(hipGetLastError());
}

template <int NC>
static void propagate_stage(const PropParams* pp, Tiles T, int stage, hipStream_t s) {
    const SkipTable& tab = skip_table(kPropStride);
    auto go = [&](auto kernel, int threads) {
        kernel<<<grid_of(kernel, threads, T), threads, 0, s>>>(pp, T, tab);
    };
    if (stage == 0) stage = propagate_stage_default(NC);
    if (stage == kStageSplitMulti) {
        go(propagate_split_kernel<stage_default(NC), NC, true>, kSplitWaves * kWave);
        return;
    }
    if (stage == kStageSplit) {
        go(propagate_split_kernel<stage_default(NC), NC, false>, kSplitWaves * kWave);
        return;
    }
    if (stage == kStageSplit3) {
        go(propagate_split3_kernel<stage_narrow(NC), NC>, kSplitWaves * kWave);
        return;
    }
    if (stage == kStageBlock || stage == kStageBlock3) {
        if constexpr (NC == kGray) {
            if (stage == kStageBlock)
                go(propagate_block_kernel<kBlockStage>, kSplitWaves * kWave);
            else
                go(propagate_block3_kernel<kBlockStage3>, kSplitWaves * kWave);
        } else {  // gray only: RGBA takes the default shape
            go(propagate_split_kernel<stage_default(NC), NC, true>, kSplitWaves * kWave);
        }
        return;
    }
    switch (stage) {
    case 512: go(propagate_kernel<512, NC>, kWave); break;
    case 704: go(propagate_kernel<704, NC>, kWave); break;
    case 896: go(propagate_kernel<896, NC>, kWave); break;
    default: throw ArgError("unsupported propagate stage width " + std::to_string(stage));
    }
}

void launch_propagate(const PropParams* pp, Tiles T, int stage, int nc, hipStream_t s) {
    if (nc == kRGBA)
        propagate_stage<kRGBA>(pp, T, stage, s);
    else
        propagate_stage<kGray>(pp, T, stage, s);
    PMH_HIP(hipGetLastError());
}

void launch_viewbuf(const ViewParams* pp, Tiles T, hipStream_t s) {
    viewbuf_kernel<<<grid_of(viewbuf_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_consistency(const ConsParams* pp, Tiles T, float maxdisp, float maxdeg,
                        hipStream_t s) {
    consistency_kernel<<<grid_of(consistency_kernel, 256, T), 256, 0, s>>>(pp, T, maxdisp, maxdeg);
    PMH_HIP(hipGetLastError());
}

void launch_occmap(const ConsParams* pp, Tiles T, float maxdisp, float maxdeg, hipStream_t s) {
    occmap_kernel<<<grid_of(occmap_kernel, 256, T), 256, 0, s>>>(pp, T, maxdisp, maxdeg);
    PMH_HIP(hipGetLastError());
}

void launch_normal_depth(const FillParams* pp, Tiles T, hipStream_t s) {
    normal_depth_kernel<<<grid_of(normal_depth_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_fill_state(const FillParams* pp, Tiles T, hipStream_t s) {
    fill_state_kernel<<<grid_of(fill_state_kernel, 256, T), 256, 0, s>>>(pp, T);
    PMH_HIP(hipGetLastError());
}

void launch_pd_step(const PdParams* pp, Tiles T, float lambda, bool inpaint, hipStream_t s) {
    if (inpaint)
        pd_step_kernel<true><<<grid_of(pd_step_kernel<true>, 256, T), 256, 0, s>>>(pp, T, lambda);
    else
        pd_step_kernel<false><<<grid_of(pd_step_kernel<false>, 256, T), 256, 0, s>>>(pp, T, lambda);
    PMH_HIP(hipGetLastError());
}

void launch_pd_steps(const PdParams* pp, Tiles T, float lambda, bool inpaint, int steps,
                     hipStream_t s) {
    if (steps < 1 || steps > kPdMaxSteps) throw ArgError("pd_steps: 1..kPdMaxSteps steps");
    if (inpaint)
        pd_steps_kernel<true><<<grid_of(pd_steps_kernel<true>, kPdThreads, T), kPdThreads, 0, s>>>(pp, T, lambda,
                                                                                   steps);
    else
        pd_steps_kernel<false><<<grid_of(pd_steps_kernel<false>, kPdThreads, T), kPdThreads, 0, s>>>(pp, T, lambda,
                                                                                    steps);
    PMH_HIP(hipGetLastError());
}

// ------------------------------------------------------- no-alloc PDROF ----
// Scratch of pmh_pdrof_noalloc: the temporary primal image and both duals of
// PDROF::apply (80 bytes per pixel) for images of up to kNoAllocPixels
// pixels, one copy per GPU, held from module load to process end.  Host code
// never uses the name `scratch` as an address: the device address comes from
// hipGetSymbolAddress, looked up once per device by noalloc_scratch_init.
static __device__ __align__(256) float4 scratch[kNoAllocScratchFloat4];

static std::atomic<float4*> g_scratch[kMaxDevices];
static std::mutex g_scratch_mutex;

void noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw ArgError("device index beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device].load()) return;
    void* ptr = nullptr;  // the current device's copy of `scratch`
    PMH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device].store(static_cast<float4*>(ptr));
}

void launch_pdrof_noalloc(int device, float4* u, const float4* f, const float* w, int W, int H,
                          float lambda, int iters, hipStream_t s) {
    const size_t n = (size_t)W * H;
    if (n > (size_t)kNoAllocPixels)
        throw ArgError("pmh_pdrof_noalloc: " + std::to_string(W) + "x" + std::to_string(H) +
                       " exceeds the no-alloc scratch (" + std::to_string(kNoAllocPixels) +
                       " pixels)");
    float4* tmp = device >= 0 && device < kMaxDevices ? g_scratch[device].load() : nullptr;
    if (!tmp)
        throw ArgError("pmh_pdrof_noalloc: no scratch address for this device (pmh_create sets it)");
    float4* p0 = tmp + n;
    float4* p1 = p0 + 2 * n;
    // PDROF.h:62-64: the dual read by the first step starts at zero; the other
    // buffers are written before they are read
    const unsigned words = (unsigned)(2 * n * 4);
    zero_one_kernel<<<std::min<unsigned>(1024u, (words + 255u) / 256u), 256, 0, s>>>(
        reinterpret_cast<uint32_t*>(p1), words);
    PMH_HIP(hipGetLastError());
    // the ping-pong of add_pdrof (PDROF.h:66-95): an even number of steps in
    // an even number of groups of up to kPdMaxSteps, so the result lands in u
    const int steps = iters + (iters & 1);
    int groups = (steps + kPdMaxSteps - 1) / kPdMaxSteps;
    groups += groups & 1;
    const int tiles = tiles_pd(W, H);
    float4* img[2] = {tmp, u};
    float4* p[2] = {p0, p1};
    for (int g = 0; g < groups; ++g) {
        const int to = g & 1 ? 1 : 0;
        const int k = steps / groups + (g < steps % groups ? 1 : 0);
        const PdParams q{img[1 - to], p[1 - to], f, w, img[to], p[to], W, H, nullptr};
        pd_steps_one_kernel<<<xcd_grid(tiles), kPdThreads, 0, s>>>(q, tiles, lambda, k);
        PMH_HIP(hipGetLastError());
    }
}

void launch_fill_depth(const FillDepthParams* pp, Tiles T, bool normals, hipStream_t s) {
    fill_depth_kernel<<<grid_of(fill_depth_kernel, 256, T), 256, 0, s>>>(pp, T, normals ? 1 : 0);
    PMH_HIP(hipGetLastError());
}

void launch_rng_draws(int W, int H, int iter, uint32_t stride, int count, float* out,
                      hipStream_t s) {
    rng_draws_kernel<<<dim3((W + 255) / 256, H), 256, 0, s>>>(W, H, iter, count, out,
                                                              skip_table(stride));
    PMH_HIP(hipGetLastError());
}

// ------------------------------------------------------- persistent grids ----
int persistent_grid(const void* kernel, int threads, int total) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, int> cache;  // (device, kernel) -> blocks
    int dev = 0;
    PMH_HIP(hipGetDevice(&dev));
    int blocks = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find({dev, kernel});
        if (it != cache.end()) {
            blocks = it->second;
        } else {
            int per_cu = 0, cus = 0;
            PMH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
            PMH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            blocks = std::max(1, per_cu) * std::max(1, cus);
            cache[{dev, kernel}] = blocks;
        }
    }
    return std::min(((blocks + 7) >> 3) << 3, tiles_grid(total, 1));
}

// ------------------------------------------------------------ host tables ----
static uint64_t host_mulmod(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) % kMwcM);
}

static SkipTable make_skip_table(uint64_t stride) {
    SkipTable t;
    uint64_t base = 1, sq = kMwcA, e = stride;  // A^stride mod M
    while (e) {
        if (e & 1) base = host_mulmod(base, sq);
        sq = host_mulmod(sq, sq);
        e >>= 1;
    }
    t.pow[0] = base;
    for (int j = 1; j < 40; ++j) t.pow[j] = host_mulmod(t.pow[j - 1], t.pow[j - 1]);
    return t;
}

const SkipTable& skip_table(uint32_t stride) {
    static const SkipTable t45 = make_skip_table(kPropStride);
    static const SkipTable t3 = make_skip_table(kInitStride);
    static const SkipTable t33 = make_skip_table(kMVStride);
    if (stride == kPropStride) return t45;
    if (stride == kInitStride) return t3;
    if (stride == kMVStride) return t33;
    throw ArgError("unsupported RNG stride");
}

void factor_table(int P, float* table) {
    const int S = 2 * P + 1;
    for (int dy = -P; dy <= P; ++dy)
        for (int dx = -P; dx <= P; ++dx) {
            const float fx = (float)dx * 1.0f, fy = (float)dy * 1.0f;
            const float len = std::sqrt(fx * fx + fy * fy);
            float t = (len - 0.0f) / (26.0f - 0.0f);
            t = std::fmin(std::fmax(t, 0.0f), 1.0f);
            const float ss = t * t * (3.0f - 2.0f * t);
            const float f = ss * 1.5f * 26.0f + 5.0f;
            const size_t j = (size_t)(dy + P) * S + (dx + P);
            const double log2e = 1.4426950408889634;
            table[j] = f;
            table[(size_t)S * S + j] = (float)(-3.0 * (double)f * log2e);
            table[2 * (size_t)S * S + j] = (float)(-(double)f * log2e);
        }
}

}  // namespace pmh
