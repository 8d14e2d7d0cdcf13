// pmh_mv.hip -- gfx950 kernels of the multi-view slanted-plane PatchMatch:
// texel records of the source views, random initialisation
// (pmmv3_init, cvt/cl/kernel/pmmvstereo.cl:170-195), one propagation pass
// (pmmv3_iteration, pmmvstereo.cl:197-304), the window cost
// (pmmv3_patch_eval, pmmvstereo.cl:97-167), the plane-induced homography
// (planef_homography, cvt/cl/kernel/Plane.cl:60-86) and the depth of a
// particle field.  Conventions: pmh_mv.hpp.
//
// Execution shape: one wave owns a row segment of 64 pixels (every pixel, or
// every other pixel of one colour in a red / black pass), so a window row is
// wave-uniform, and every lane costs its own candidate plane, tap by tap.  At
// a projected point a lane needs two runs of 10 floats (two texel records of
// 20 bytes on each of two rows) of each source view.  The init and cost
// kernels read them through L1/L2; the propagate kernel runs three waves per
// segment on different candidates and stages, per window row, the source
// texels their planes reach in LDS (see mv_cost_block).  A tap's projection
// and bilinear sample follow the restatement's float operations one for one
// (the library is built with -ffp-contract=off and IEEE division):
// homographies, projections, samples and particles are bit-exact; the support
// weight uses the hardware exp2 of a folded exponent (a few ulp from libm's
// exp).
#include "pmh_mv.hpp"
#include "pmh_window.hpp"

#include <climits>
#include <cmath>

namespace pmh {

// ------------------------------------------------------------ plane algebra --
__device__ __forceinline__ float mv_nz(float nx, float ny) {
    return __builtin_sqrtf((1.0f - nx * nx) - ny * ny);
}

// n . K_r^-1 (x, y, 1)
__device__ __forceinline__ float mv_ray_dot(const MVConst& k, float nx, float ny, float nz,
                                            float x, float y) {
    const float rx = __builtin_fmaf(k.kinv[0], x, __builtin_fmaf(k.kinv[1], y, k.kinv[2]));
    const float ry = __builtin_fmaf(k.kinv[3], x, __builtin_fmaf(k.kinv[4], y, k.kinv[5]));
    const float rz = __builtin_fmaf(k.kinv[6], x, __builtin_fmaf(k.kinv[7], y, k.kinv[8]));
    return __builtin_fmaf(nx, rx, __builtin_fmaf(ny, ry, nz * rz));
}

// q = K_r^-T n / d
__device__ __forceinline__ void mv_plane_q(const MVConst& k, const float4& p, float q[3]) {
    const float nz = mv_nz(p.x, p.y);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float m = __builtin_fmaf(p.x, k.kinv[j],
                                       __builtin_fmaf(p.y, k.kinv[3 + j], nz * k.kinv[6 + j]));
        q[j] = m / p.z;
    }
}

// Normal components clamped to +-0.95 and scaled back into the unit disc
// (nd_state_init, pmhstereo.cl:58-62), then the particle through depth z at
// pixel (x, y).
__device__ __forceinline__ float4 mv_particle_at(const MVConst& k, float nx, float ny, float z,
                                                 float x, float y) {
    nx = clampf(nx, -0.95f, 0.95f);
    ny = clampf(ny, -0.95f, 0.95f);
    const float nf = fmaxf(__builtin_sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nf;
    ny = ny / nf;
    const float nz = mv_nz(nx, ny);
    return make_float4(nx, ny, z * mv_ray_dot(k, nx, ny, nz, x, y), 0.0f);
}

// pmmv_particle_init (pmmvstereo.cl:52-63) with depth range [dmin, dmax]
__device__ __forceinline__ float4 mv_particle_init(const MVConst& k, float u0, float u1, float u2,
                                                   float x, float y, float nm, float dmin,
                                                   float dmax) {
    const float z = __builtin_fmaf(u0, dmax - dmin, dmin);
    const float nx = ((u1 - 0.5f) * nm) * 0.95f;
    const float ny = ((u2 - 0.5f) * nm) * 0.95f;
    return mv_particle_at(k, nx, ny, z, x, y);
}

// pmmv_particle_refine (pmmvstereo.cl:65-77): depth at the pixel +- (dmax -
// dmin) / 64, normal components +- 0.1 (NORMALREFINEMUL 0.2)
__device__ __forceinline__ float4 mv_particle_refine(const MVConst& k, float u0, float u1,
                                                     float u2, const float4& p, float x, float y,
                                                     float dmin, float dmax) {
    const float z0 = p.z / mv_ray_dot(k, p.x, p.y, mv_nz(p.x, p.y), x, y);
    const float z = clampf(__builtin_fmaf(u0 - 0.5f, (dmax - dmin) / 32.0f, z0), dmin, dmax);
    const float nx = __builtin_fmaf(u1 - 0.5f, 0.2f, p.x);
    const float ny = __builtin_fmaf(u2 - 0.5f, 0.2f, p.y);
    return mv_particle_at(k, nx, ny, z, x, y);
}

// ------------------------------------------------------------- window cost --
// Bilinear sample of channel c from the corner runs r0 (row j0: texels i0,
// i0 + 1) and r1 (row j0 + 1): the cell (t00, dx, dy, dxy) of the restatement,
// then fmaf(b, fmaf(a, dxy, dy), fmaf(a, dx, t00)).
__device__ __forceinline__ float mv_sample(const float* r0, const float* r1, int c, float a,
                                           float b) {
    const float t00 = r0[c], t10 = r0[kMVTexel + c], t01 = r1[c], t11 = r1[kMVTexel + c];
    const float dx = t10 - t00;
    return __builtin_fmaf(b, __builtin_fmaf(a, (t11 - t01) - dx, t01 - t00),
                          __builtin_fmaf(a, dx, t00));
}

// Window cost of particle p at pixel (x, y) over NS source views, summed in
// dy-outer, dx-inner, view-innermost order.  Lanes whose pixel lies outside
// the image pass a clamped x and discard the result.
template <int NS>
__device__ __forceinline__ float mv_cost(const MVImages& im, const MVConst& k,
                                         const float* __restrict__ Fw, int P, int x, int y,
                                         float c0, const float4& p) {
    const int W = im.W, H = im.H;
    const float Wf = (float)W, Hf = (float)H;
    const int S = 2 * P + 1;
    float q[3];
    mv_plane_q(k, p, q);
    float h[NS][9];
#pragma unroll
    for (int v = 0; v < NS; ++v)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                h[v][i * 3 + j] = __builtin_fmaf(k.e[v][i], q[j], k.M[v][i * 3 + j]);
    float wsum = 0.0f, ret = 0.0f;
    const int y0 = max(y - P, 0), y1 = min(y + P, H - 1);
    const int x0 = max(x - P, 0), x1 = min(x + P, W - 1);
    for (int qy = y0; qy <= y1; ++qy) {
        const float fqy = (float)qy;
        const float rq = __builtin_fmaf(q[1], fqy, q[2]);
        float rX[NS], rY[NS], rZ[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            rX[v] = __builtin_fmaf(h[v][1], fqy, h[v][2]);
            rY[v] = __builtin_fmaf(h[v][4], fqy, h[v][5]);
            rZ[v] = __builtin_fmaf(h[v][7], fqy, h[v][8]);
        }
        const float* __restrict__ Fr = Fw + (qy - y + P) * S + (P - x);
        const float* __restrict__ Irow = im.Ir + (size_t)qy * W;
        const float4* __restrict__ Grow = im.Gr + (size_t)qy * W;
        for (int qx = x0; qx <= x1; ++qx) {
            const float fqx = (float)qx;
            if (!(__builtin_fmaf(q[0], fqx, rq) > 0.0f)) continue;
            const float v1 = Irow[qx];
            const float wt = __builtin_amdgcn_exp2f(fabsf(c0 - v1) * Fr[qx]);
            const float4 g1 = Grow[qx];
#pragma unroll
            for (int v = 0; v < NS; ++v) {
                const float X = __builtin_fmaf(h[v][0], fqx, rX[v]);
                const float Y = __builtin_fmaf(h[v][3], fqx, rY[v]);
                const float Z = __builtin_fmaf(h[v][6], fqx, rZ[v]);
                if (!(Z > 0.0f)) continue;
                const float r = 1.0f / Z;
                const float xp = X * r, yp = Y * r;
                if (!(xp >= 0.0f && xp < Wf && yp >= 0.0f && yp < Hf)) continue;
                wsum += wt;
                const int i0 = (int)xp, j0 = (int)yp;
                const float a = xp - (float)i0, b = yp - (float)j0;
                const float* __restrict__ p0 =
                    im.tex + (size_t)v * im.tstride +
                    ((unsigned)j0 * (unsigned)(W + 1) + (unsigned)i0) * (unsigned)kMVTexel;
                const float* __restrict__ p1 = p0 + (W + 1) * kMVTexel;
                float r0[2 * kMVTexel], r1[2 * kMVTexel];
#pragma unroll
                for (int e = 0; e < 2 * kMVTexel; ++e) {
                    r0[e] = p0[e];
                    r1[e] = p1[e];
                }
                const float v2 = mv_sample(r0, r1, 0, a, b);
                const float cterm = 3.0f * fminf(fabsf(v1 - v2), 0.04f);
                const float m0 = fminf(fabsf(g1.x - mv_sample(r0, r1, 1, a, b)), 0.01f);
                const float m1 = fminf(fabsf(g1.y - mv_sample(r0, r1, 2, a, b)), 0.01f);
                const float m2 = fminf(fabsf(g1.z - mv_sample(r0, r1, 3, a, b)), 0.01f);
                const float m3 = fminf(fabsf(g1.w - mv_sample(r0, r1, 4, a, b)), 0.01f);
                const float gs = ((m0 + m1) + m2) + m3;
                ret = __builtin_fmaf(wt, __builtin_fmaf(0.95f, gs, 0.05f * cterm), ret);
            }
        }
    }
    if (!finite3(p) || wsum <= 1.1f) return 1e5f;
    return ret / wsum;
}

// ------------------------------------------------------------------ kernels --
// Texel records (I, gx, gy, gxy, gyx) of one source view, (W + 1) x (H + 1):
// the last column and row repeated once (the clamp to edge of the samples).
__global__ __launch_bounds__(256) void mv_texels_kernel(const float* __restrict__ img,
                                                        const float4* __restrict__ grad, int W,
                                                        int H, float* __restrict__ tex) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x > W) return;
    const int i = min(y, H - 1) * W + min(x, W - 1);
    const float4 g = grad[i];
    float* o = tex + ((size_t)y * (W + 1) + x) * kMVTexel;
    o[0] = img[i];
    o[1] = g.x;
    o[2] = g.y;
    o[3] = g.z;
    o[4] = g.w;
}

// pmmv3_init (pmmvstereo.cl:170-195): RNG_init(y W + x, 3), a random particle
// (normal spread 1) and its cost.
template <int NS>
__global__ __launch_bounds__(kWave) void mv_init_kernel(MVImages im, MVConst k,
                                                        const float* __restrict__ Fw, int P,
                                                        float dmin, float dmax,
                                                        float4* __restrict__ out, int nbx, int nb,
                                                        SkipTable tab) {
    const int t = xcd_tile(nb);
    if (t >= nb) return;
    const int y = __builtin_amdgcn_readfirstlane(t / nbx);
    const Seg sg = row_segment(t % nbx, y, im.W, 0);
    const int x = sg.x;
    uint32_t rx, rc;
    mwc_seed(tab, (uint64_t)y * im.W + x, rx, rc);
    const float u0 = mwc_next_float(rx, rc);
    const float u1 = mwc_next_float(rx, rc);
    const float u2 = mwc_next_float(rx, rc);
    float4 p = mv_particle_init(k, u0, u1, u2, (float)x, (float)y, 1.0f, dmin, dmax);
    p.w = mv_cost<NS>(im, k, Fw, P, x, y, im.Ir[y * im.W + x], p);
    if (sg.active) out[y * im.W + x] = p;
}

template <int NS>
__global__ __launch_bounds__(kWave) void mv_cost_kernel(MVImages im, MVConst k,
                                                        const float* __restrict__ Fw, int P,
                                                        const float4* __restrict__ particles,
                                                        float* __restrict__ cost, int nbx,
                                                        int nb) {
    const int t = xcd_tile(nb);
    if (t >= nb) return;
    const int y = __builtin_amdgcn_readfirstlane(t / nbx);
    const Seg sg = row_segment(t % nbx, y, im.W, 0);
    const int i = y * im.W + sg.x;
    const float c = mv_cost<NS>(im, k, Fw, P, sg.x, y, im.Ir[i], particles[i]);
    if (sg.active) cost[i] = c;
}

// Draws of one pixel's pass: the supplied ones (kMVDraws per pixel) or the
// MWC64X stream RNG_init(y W + x + iter, 33) (pmmvstereo.cl:237), in order.
struct MVDraws {
    const float* d;
    uint32_t rx, rc;
    __device__ __forceinline__ float next() {
        if (d) return *d++;
        return mwc_next_float(rx, rc);
    }
};

// ------------------------------------------------------------- propagate ----
// pmmv3_iteration (pmmvstereo.cl:197-304): the 8 neighbours, kMVNbhd random
// neighbours at offsets (int)(u 7 + 0.5) in [0, 7]^2, one random particle
// (normal spread 2), nref randomized refinements; a candidate replaces the
// pixel's particle if its cost is strictly lower (:245).  sweep 0 reads `old`
// only; sweep 1 writes the red pixels ((x + y) even), sweep 2 the black ones,
// reading the red neighbours from `out`.
//
// A row segment's candidates are spread over the waves of a block: wave w
// costs candidates w, w + kMVWaves, ... (four rounds of three), then the
// refinements two at a time (wave 0 the first, waves 1 and 2 the second after
// keeping or after taking the first).  In a round the waves
// sweep the window rows in lockstep and share one LDS stage per view: for
// each row the block stages the source texels that the waves' planes reach,
// each wave's box taken from the plane of its middle lane (the window corners
// of the wave's pixels projected once per round, the rows of each window row
// interpolated between the top and bottom corners, a texel of margin); boxes
// join in wave order while the union fits the stage.  A tap reads LDS when its
// four texels lie in the staged box and global memory otherwise, so the
// boxes change only where data comes from: the costs are those of mv_cost,
// tap for tap.  The pass is the sequential fold of pmmv3_iteration: the
// per-wave folds (strict <, in candidate order) combine by lowest cost, then
// lowest candidate index, and replace the pixel's particle if strictly lower.
constexpr int kMVWaves = 3;  // 3: 15 cost slots for 14 candidates; 4 waves: 16 and 1% slower
constexpr int kMVStageFloats = 1536 * kMVTexel;  // all views' stages of one row (31 KB)

template <int NS>
struct MVBlockLds {
    static constexpr int kCap = (kMVStageFloats / NS) & ~3;  // floats per view
    __align__(16) float t[NS][kCap];
    int box[kMVWaves][NS][4];  // per wave and view: texel columns x0..x1, rows y0..y1
    float bc[kMVWaves][kWave];  // per-wave fold: cost, candidate, particle
    int bk[kMVWaves][kWave];
    float4 bp[kMVWaves][kWave];
    float rc[3][kWave];  // refinement costs
};

__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// floor of a projected coordinate, clamped to a little outside [0, hi]
__device__ __forceinline__ int mv_floor_clamped(float v, float hi) {
    return (int)floorf(fminf(fmaxf(v, -4.0f), hi + 4.0f));
}

template <int NS>
__device__ __forceinline__ float mv_cost_block(const MVImages& im, const MVConst& k,
                                               const float* __restrict__ Fw, int P, int x,
                                               int y, float c0, const float4& p, bool has,
                                               MVBlockLds<NS>& L, int w) {
    constexpr int kCap = MVBlockLds<NS>::kCap;
    const int W = im.W, H = im.H;
    const float Wf = (float)W, Hf = (float)H;
    const int S = 2 * P + 1;
    const int tid = (int)threadIdx.x;
    const unsigned rowf = (unsigned)(W + 1) * (unsigned)kMVTexel;  // floats per texel row
    float q[3];
    mv_plane_q(k, p, q);
    float h[NS][9];
#pragma unroll
    for (int v = 0; v < NS; ++v)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                h[v][i * 3 + j] = __builtin_fmaf(k.e[v][i], q[j], k.M[v][i * 3 + j]);
    float wsum = 0.0f, ret = 0.0f;
    const int y0 = max(y - P, 0), y1 = min(y + P, H - 1);
    const int x0 = max(x - P, 0), x1 = min(x + P, W - 1);

    // the wave's box parameters per view, from its middle lane's plane
    const float4 ap = make_float4(__shfl(p.x, kWave / 2), __shfl(p.y, kWave / 2),
                                  __shfl(p.z, kWave / 2), 0.0f);
    const float cxa = (float)__shfl(x0, 0), cxb = (float)__shfl(x1, kWave - 1);
    const float cya = (float)y0, cyb = (float)y1;
    bool bok[NS];
    int bc0[NS], bc1[NS];
    float ylt[NS], yht[NS], ylb[NS], yhb[NS];
    {
        float aq[3];
        mv_plane_q(k, ap, aq);
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            float hv[9];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    hv[i * 3 + j] = __builtin_fmaf(k.e[v][i], aq[j], k.M[v][i * 3 + j]);
            float xs[4], ys[4];
            bool ok = has && finite3(ap);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float cx = c & 1 ? cxb : cxa, cy = c & 2 ? cyb : cya;
                const float X = __builtin_fmaf(hv[0], cx, __builtin_fmaf(hv[1], cy, hv[2]));
                const float Y = __builtin_fmaf(hv[3], cx, __builtin_fmaf(hv[4], cy, hv[5]));
                const float Z = __builtin_fmaf(hv[6], cx, __builtin_fmaf(hv[7], cy, hv[8]));
                const float r = __builtin_amdgcn_rcpf(Z);
                xs[c] = X * r;
                ys[c] = Y * r;
                ok = ok && Z > 0.0f && isfinite(xs[c]) && isfinite(ys[c]);
            }
            bok[v] = __builtin_amdgcn_readfirstlane(ok ? 1 : 0) != 0;
            bc0[v] = __builtin_amdgcn_readfirstlane(
                mv_floor_clamped(fminf(fminf(xs[0], xs[1]), fminf(xs[2], xs[3])), Wf) - 1);
            bc1[v] = __builtin_amdgcn_readfirstlane(
                mv_floor_clamped(fmaxf(fmaxf(xs[0], xs[1]), fmaxf(xs[2], xs[3])), Wf) + 2);
            ylt[v] = uniform_f(fminf(ys[0], ys[1]));
            yht[v] = uniform_f(fmaxf(ys[0], ys[1]));
            ylb[v] = uniform_f(fminf(ys[2], ys[3]));
            yhb[v] = uniform_f(fmaxf(ys[2], ys[3]));
        }
    }

    for (int qy = y0; qy <= y1; ++qy) {
        const float fqy = (float)qy;
        const float rq = __builtin_fmaf(q[1], fqy, q[2]);
        float rX[NS], rY[NS], rZ[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            rX[v] = __builtin_fmaf(h[v][1], fqy, h[v][2]);
            rY[v] = __builtin_fmaf(h[v][4], fqy, h[v][5]);
            rZ[v] = __builtin_fmaf(h[v][7], fqy, h[v][8]);
        }
        // 1. this wave's box per view for this window row
        const float tr = y1 > y0 ? (float)(qy - y0) / (float)(y1 - y0) : 0.0f;
        if (lane_id() == 0) {
#pragma unroll
            for (int v = 0; v < NS; ++v) {
                int b[4] = {1, 0, 1, 0};  // empty
                if (bok[v]) {
                    const float lo = __builtin_fmaf(tr, ylb[v] - ylt[v], ylt[v]);
                    const float hi = __builtin_fmaf(tr, yhb[v] - yht[v], yht[v]);
                    b[0] = max(bc0[v], 0);
                    b[1] = min(bc1[v], W);
                    b[2] = max(mv_floor_clamped(lo, Hf) - 1, 0);
                    b[3] = min(mv_floor_clamped(hi, Hf) + 2, H);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) L.box[w][v][e] = b[e];
            }
        }
        __syncthreads();
        // 2. the waves' boxes joined in wave order while the union fits, staged
        //    with aligned 16-byte copies (rows of s4 floats from the aligned
        //    float below each row's first texel)
        int ux0[NS], ux1[NS], uy0[NS], uy1[NS], s4[NS];
        unsigned g0[NS];
        bool staged[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            int a0 = INT_MAX, a1 = INT_MIN, b0 = INT_MAX, b1 = INT_MIN;
#pragma unroll
            for (int ww = 0; ww < kMVWaves; ++ww) {
                const int c0w = __builtin_amdgcn_readfirstlane(L.box[ww][v][0]);
                const int c1w = __builtin_amdgcn_readfirstlane(L.box[ww][v][1]);
                const int r0w = __builtin_amdgcn_readfirstlane(L.box[ww][v][2]);
                const int r1w = __builtin_amdgcn_readfirstlane(L.box[ww][v][3]);
                if (c0w > c1w || r0w > r1w) continue;
                const int n0 = min(a0, c0w), n1 = max(a1, c1w);
                const int m0 = min(b0, r0w), m1 = max(b1, r1w);
                const int st = ((n1 - n0 + 1) * kMVTexel + 6) & ~3;
                if ((long long)st * (m1 - m0 + 1) > kCap) continue;
                a0 = n0;
                a1 = n1;
                b0 = m0;
                b1 = m1;
            }
            staged[v] = a0 <= a1;
            ux0[v] = a0;
            ux1[v] = a1;
            uy0[v] = b0;
            uy1[v] = b1;
            s4[v] = ((a1 - a0 + 1) * kMVTexel + 6) & ~3;
            // the low bits of the absolute float index of the box's first texel
            // (the alignment offsets of the staged rows; wraps harmlessly)
            g0[v] = (unsigned)v * im.tstride + ((unsigned)b0 * (unsigned)(W + 1) + (unsigned)a0) *
                                                   (unsigned)kMVTexel;
            if (staged[v]) {
                const int n4 = s4[v] >> 2;
                const size_t gv = (size_t)v * im.tstride +
                                  ((size_t)b0 * (W + 1) + (size_t)a0) * kMVTexel;
                for (int r = 0; r <= b1 - b0; ++r) {
                    const size_t gr = gv + (size_t)r * rowf;
                    const float4* __restrict__ src =
                        reinterpret_cast<const float4*>(im.tex + (gr & ~(size_t)3));
                    float4* __restrict__ dst = reinterpret_cast<float4*>(L.t[v] + r * s4[v]);
                    for (int c = tid; c < n4; c += kMVWaves * kWave) dst[c] = src[c];
                }
            }
        }
        __syncthreads();
        // 3. the taps of this row
        if (has) {
            const float* __restrict__ Fr = Fw + (qy - y + P) * S + (P - x);
            const float* __restrict__ Irow = im.Ir + (size_t)qy * W;
            const float4* __restrict__ Grow = im.Gr + (size_t)qy * W;
            for (int qx = x0; qx <= x1; ++qx) {
                const float fqx = (float)qx;
                if (!(__builtin_fmaf(q[0], fqx, rq) > 0.0f)) continue;
                const float v1 = Irow[qx];
                const float wt = __builtin_amdgcn_exp2f(fabsf(c0 - v1) * Fr[qx]);
                const float4 g1 = Grow[qx];
#pragma unroll
                for (int v = 0; v < NS; ++v) {
                    const float X = __builtin_fmaf(h[v][0], fqx, rX[v]);
                    const float Y = __builtin_fmaf(h[v][3], fqx, rY[v]);
                    const float Z = __builtin_fmaf(h[v][6], fqx, rZ[v]);
                    if (!(Z > 0.0f)) continue;
                    const float r = 1.0f / Z;
                    const float xp = X * r, yp = Y * r;
                    if (!(xp >= 0.0f && xp < Wf && yp >= 0.0f && yp < Hf)) continue;
                    wsum += wt;
                    const int i0 = (int)xp, j0 = (int)yp;
                    const float a = xp - (float)i0, b = yp - (float)j0;
                    float r0[2 * kMVTexel], r1[2 * kMVTexel];
                    if (staged[v] && i0 >= ux0[v] && i0 + 1 <= ux1[v] && j0 >= uy0[v] &&
                        j0 + 1 <= uy1[v]) {
                        const int lj = j0 - uy0[v];
                        const unsigned ga = g0[v] + (unsigned)lj * rowf;
                        const unsigned gb = ga + rowf;
                        const float* __restrict__ s0 = L.t[v] + lj * s4[v] + (int)(ga & 3u) +
                                                       (i0 - ux0[v]) * kMVTexel;
                        const float* __restrict__ s1 = L.t[v] + (lj + 1) * s4[v] +
                                                       (int)(gb & 3u) + (i0 - ux0[v]) * kMVTexel;
#pragma unroll
                        for (int e = 0; e < 2 * kMVTexel; ++e) {
                            r0[e] = s0[e];
                            r1[e] = s1[e];
                        }
                    } else {
                        const float* __restrict__ p0 =
                            im.tex + (size_t)v * im.tstride +
                            ((unsigned)j0 * (unsigned)(W + 1) + (unsigned)i0) * (unsigned)kMVTexel;
                        const float* __restrict__ p1 = p0 + (W + 1) * kMVTexel;
#pragma unroll
                        for (int e = 0; e < 2 * kMVTexel; ++e) {
                            r0[e] = p0[e];
                            r1[e] = p1[e];
                        }
                    }
                    const float v2 = mv_sample(r0, r1, 0, a, b);
                    const float cterm = 3.0f * fminf(fabsf(v1 - v2), 0.04f);
                    const float m0 = fminf(fabsf(g1.x - mv_sample(r0, r1, 1, a, b)), 0.01f);
                    const float m1 = fminf(fabsf(g1.y - mv_sample(r0, r1, 2, a, b)), 0.01f);
                    const float m2 = fminf(fabsf(g1.z - mv_sample(r0, r1, 3, a, b)), 0.01f);
                    const float m3 = fminf(fabsf(g1.w - mv_sample(r0, r1, 4, a, b)), 0.01f);
                    const float gs = ((m0 + m1) + m2) + m3;
                    ret = __builtin_fmaf(wt, __builtin_fmaf(0.95f, gs, 0.05f * cterm), ret);
                }
            }
        }
    }
    if (!finite3(p) || wsum <= 1.1f) return 1e5f;
    return ret / wsum;
}

__device__ __forceinline__ float4 sel4(bool c, const float4& a, const float4& b) {
    return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

template <int NS>
__global__ __launch_bounds__(kMVWaves* kWave) void mv_propagate_block_kernel(
    MVImages im, MVConst k, const float* __restrict__ Fw, int P, const float4* __restrict__ old,
    float4* __restrict__ out, float dmin, float dmax, int iter, const float* __restrict__ draws,
    int sweep, int nref, int nbx, int nb, SkipTable tab) {
    __shared__ MVBlockLds<NS> L;
    const int t = xcd_tile(nb);
    if (t >= nb) return;  // the same for every thread of the block
    const int W = im.W, H = im.H;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = lane_id();
    const int y = __builtin_amdgcn_readfirstlane(t / nbx);
    const Seg sg = row_segment(t % nbx, y, W, sweep);
    const int x = sg.x;
    const float cx = (float)x, cy = (float)y;
    const float c0 = im.Ir[y * W + x];
    MVDraws dr;
    dr.d = draws ? draws + (size_t)(y * W + x) * kMVDraws : nullptr;
    if (!draws) mwc_seed(tab, (uint64_t)y * W + x + (uint64_t)iter, dr.rx, dr.rc);
    float u[2 * kMVNbhd + 3];
#pragma unroll
    for (int i = 0; i < 2 * kMVNbhd + 3; ++i) u[i] = dr.next();
    const float4 self0 = old[y * W + x];

    // candidates 0..11 of pmmv3_iteration, three per wave
    float bestc = INFINITY;
    int bestk = -1;
    float4 bestp = self0;
#pragma unroll 1
    for (int r = 0; r < (8 + kMVNbhd + 1 + kMVWaves - 1) / kMVWaves; ++r) {
        const int kc = w + kMVWaves * r;
        const bool has = kc < 8 + kMVNbhd + 1;
        float4 cand;
        if (kc < 8 + kMVNbhd) {
            int nx, ny;
            if (kc < 8) {
                const int n = kc < 4 ? kc : kc + 1;
                nx = x + n % 3 - 1;
                ny = y + n / 3 - 1;
            } else {
                const int i = kc - 8;
                const float ux = i == 0 ? u[0] : (i == 1 ? u[2] : u[4]);
                const float uy = i == 0 ? u[1] : (i == 1 ? u[3] : u[5]);
                nx = x + (int)(ux * 7.0f + 0.5f);
                ny = y + (int)(uy * 7.0f + 0.5f);
            }
            nx = clampi(nx, 0, W - 1);
            ny = clampi(ny, 0, H - 1);
            const float4* src = (sweep == 2 && ((nx + ny) & 1) == 0) ? out : old;
            cand = src[ny * W + nx];
        } else {
            cand = mv_particle_init(k, u[6], u[7], u[8], cx, cy, 2.0f, dmin, dmax);
        }
        cand.w = mv_cost_block<NS>(im, k, Fw, P, x, y, c0, cand, has, L, w);
        const bool better = has && cand.w < bestc;
        bestc = better ? cand.w : bestc;
        bestk = better ? kc : bestk;
        bestp = sel4(better, cand, bestp);
    }
    L.bc[w][lane] = bestc;
    L.bk[w][lane] = bestk;
    L.bp[w][lane] = bestp;
    __syncthreads();
    float4 s = self0;
    {
        float mc = INFINITY;
        int mk = INT_MAX, win = -1;
#pragma unroll
        for (int ww = 0; ww < kMVWaves; ++ww) {
            const float c = L.bc[ww][lane];
            const int kk = L.bk[ww][lane];
            const bool better = kk >= 0 && (c < mc || (c == mc && kk < mk));
            mc = better ? c : mc;
            mk = better ? kk : mk;
            win = better ? ww : win;
        }
        if (win >= 0 && mc < s.w) s = L.bp[win][lane];
    }

    // randomized refinements, two per round
#pragma unroll 1
    for (int j = 0; j < nref; j += 2) {
        const bool two = j + 1 < nref;
        const float a0 = dr.next(), a1 = dr.next(), a2 = dr.next();
        float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
        if (two) {
            b0 = dr.next();
            b1 = dr.next();
            b2 = dr.next();
        }
        const float4 A = mv_particle_refine(k, a0, a1, a2, s, cx, cy, dmin, dmax);
        float4 mine = A;
        bool has = w == 0;
        if (w == 1 && two) {
            mine = mv_particle_refine(k, b0, b1, b2, s, cx, cy, dmin, dmax);
            has = true;
        } else if (w == 2 && two) {
            mine = mv_particle_refine(k, b0, b1, b2, A, cx, cy, dmin, dmax);
            has = true;
        }
        const float cst = mv_cost_block<NS>(im, k, Fw, P, x, y, c0, mine, has, L, w);
        if (has) L.rc[w][lane] = cst;
        __syncthreads();
        float4 a = A;
        a.w = L.rc[0][lane];
        const bool tookA = a.w < s.w;
        const float4 s1 = sel4(tookA, a, s);
        if (two) {
            float4 second = mv_particle_refine(k, b0, b1, b2, s1, cx, cy, dmin, dmax);
            second.w = tookA ? L.rc[2][lane] : L.rc[1][lane];
            s = sel4(second.w < s1.w, second, s1);
        } else {
            s = s1;
        }
    }
    if (w == 0 && sg.active) out[y * W + x] = s;
}

__global__ __launch_bounds__(256) void mv_depth_kernel(const float4* __restrict__ particles,
                                                       MVConst k, int W, int H,
                                                       float* __restrict__ depth) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const float4 p = particles[y * W + x];
    const float z = p.z / mv_ray_dot(k, p.x, p.y, mv_nz(p.x, p.y), (float)x, (float)y);
    depth[y * W + x] = (isfinite(z) && z > 0.0f) ? z : 0.0f;
}

__global__ __launch_bounds__(256) void mv_homography_kernel(const float4* __restrict__ particles,
                                                            int count, MVConst k, int nsrc,
                                                            float* __restrict__ H) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    float q[3];
    mv_plane_q(k, particles[i], q);
    for (int v = 0; v < nsrc; ++v)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                H[((size_t)i * nsrc + v) * 9 + a * 3 + b] =
                    __builtin_fmaf(k.e[v][a], q[b], k.M[v][a * 3 + b]);
}

// --------------------------------------------------------------- launchers --
void launch_mv_texels(const float* img, const float4* grad, int W, int H, float* tex,
                      hipStream_t s) {
    mv_texels_kernel<<<dim3((W + 1 + 255) / 256, H + 1), 256, 0, s>>>(img, grad, W, H, tex);
    PMH_HIP(hipGetLastError());
}

template <template <int> class Launch, class... A>
static void by_views(int nsrc, A&&... a) {
    switch (nsrc) {
        case 1: Launch<1>::run(a...); break;
        case 2: Launch<2>::run(a...); break;
        case 3: Launch<3>::run(a...); break;
        case 4: Launch<4>::run(a...); break;
        default: throw ArgError("1 to 4 source views");
    }
}

template <int NS>
struct InitL {
    static void run(const MVImages& im, const MVConst& k, const float* Fw, int P, float dmin,
                    float dmax, float4* out, hipStream_t s) {
        const int nbx = (im.W + kWave - 1) / kWave, nb = nbx * im.H;
        mv_init_kernel<NS><<<xcd_grid(nb), kWave, 0, s>>>(im, k, Fw, P, dmin, dmax, out, nbx, nb,
                                                          skip_table(kInitStride));
    }
};
template <int NS>
struct CostL {
    static void run(const MVImages& im, const MVConst& k, const float* Fw, int P,
                    const float4* particles, float* cost, hipStream_t s) {
        const int nbx = (im.W + kWave - 1) / kWave, nb = nbx * im.H;
        mv_cost_kernel<NS><<<xcd_grid(nb), kWave, 0, s>>>(im, k, Fw, P, particles, cost, nbx,
                                                          nb);
    }
};
template <int NS>
struct PropL {
    static void run(const MVImages& im, const MVConst& k, const float* Fw, int P,
                    const float4* old, float4* out, float dmin, float dmax, int iter,
                    const float* draws, int sweep, int nref, hipStream_t s) {
        const int lanes = sweep == 0 ? im.W : (im.W + 1) / 2;
        const int nbx = (lanes + kWave - 1) / kWave, nb = nbx * im.H;
        mv_propagate_block_kernel<NS><<<xcd_grid(nb), kMVWaves * kWave, 0, s>>>(
            im, k, Fw, P, old, out, dmin, dmax, iter, draws, sweep, nref, nbx, nb,
            skip_table(kMVStride));
    }
};

void launch_mv_init(const MVImages& im, const MVConst& k, int nsrc, const float* Fw, int P,
                    float dmin, float dmax, float4* out, hipStream_t s) {
    by_views<InitL>(nsrc, im, k, Fw, P, dmin, dmax, out, s);
    PMH_HIP(hipGetLastError());
}

void launch_mv_cost(const MVImages& im, const MVConst& k, int nsrc, const float* Fw, int P,
                    const float4* particles, float* cost, hipStream_t s) {
    by_views<CostL>(nsrc, im, k, Fw, P, particles, cost, s);
    PMH_HIP(hipGetLastError());
}

void launch_mv_propagate(const MVImages& im, const MVConst& k, int nsrc, const float* Fw, int P,
                         const float4* old, float4* out, float dmin, float dmax, int iter,
                         const float* draws, int sweep, int nref, hipStream_t s) {
    by_views<PropL>(nsrc, im, k, Fw, P, old, out, dmin, dmax, iter, draws, sweep, nref, s);
    PMH_HIP(hipGetLastError());
}

void launch_mv_depth(const float4* particles, const MVConst& k, int W, int H, float* depth,
                     hipStream_t s) {
    mv_depth_kernel<<<dim3((W + 255) / 256, H), 256, 0, s>>>(particles, k, W, H, depth);
    PMH_HIP(hipGetLastError());
}

void launch_mv_homography(const float4* particles, int count, const MVConst& k, int nsrc,
                          float* H, hipStream_t s) {
    if (count <= 0) return;
    mv_homography_kernel<<<(count + 255) / 256, 256, 0, s>>>(particles, count, k, nsrc, H);
    PMH_HIP(hipGetLastError());
}

// ------------------------------------------------------------ host tables --
static void inv3(const double* a, double* r) {
    const double c00 = a[4] * a[8] - a[5] * a[7];
    const double c01 = a[5] * a[6] - a[3] * a[8];
    const double c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
        throw ArgError("camera intrinsics K are singular");
    r[0] = c00 / det;
    r[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    r[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    r[3] = c01 / det;
    r[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    r[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    r[6] = c02 / det;
    r[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    r[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

static void mul3(const double* a, const double* b, double* r) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

void mv_constants(const float* Kr, const float* Tr, int nsrc, const float* Ks, const float* Ts,
                  MVConst& out) {
    if (nsrc < 1 || nsrc > kMVMaxViews) throw ArgError("1 to 4 source views");
    out = MVConst{};
    double K[9], Ki[9], Rr[9], tr[3];
    for (int i = 0; i < 9; ++i) K[i] = Kr[i];
    inv3(K, Ki);
    for (int i = 0; i < 9; ++i) out.kinv[i] = (float)Ki[i];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) Rr[i * 3 + j] = Tr[i * 4 + j];
        tr[i] = Tr[i * 4 + 3];
    }
    for (int v = 0; v < nsrc; ++v) {
        const float* Kv = Ks + 9 * v;
        const float* Tv = Ts + 12 * v;
        double Rs[9], RrT[9], R[9], t[3], KvD[9], KvI[9], KR[9], Mv[9];
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                Rs[i * 3 + j] = Tv[i * 4 + j];
                RrT[i * 3 + j] = Rr[j * 3 + i];
            }
        }
        for (int i = 0; i < 9; ++i) KvD[i] = Kv[i];
        inv3(KvD, KvI);  // rejects a singular source K as well
        mul3(Rs, RrT, R);
        for (int i = 0; i < 3; ++i)
            t[i] = (double)Tv[i * 4 + 3] -
                   (R[i * 3] * tr[0] + R[i * 3 + 1] * tr[1] + R[i * 3 + 2] * tr[2]);
        mul3(KvD, R, KR);
        mul3(KR, Ki, Mv);
        for (int i = 0; i < 9; ++i) out.M[v][i] = (float)Mv[i];
        for (int i = 0; i < 3; ++i)
            out.e[v][i] =
                (float)(KvD[i * 3] * t[0] + KvD[i * 3 + 1] * t[1] + KvD[i * 3 + 2] * t[2]);
    }
}

void mv_factor_table(int P, float* table) {
    const int S = 2 * P + 1;
    for (int dy = -P; dy <= P; ++dy)
        for (int dx = -P; dx <= P; ++dx) {
            const float fx = (float)dx, fy = (float)dy;
            const float len = std::sqrt(fx * fx + fy * fy);
            float t = len / 28.0f;
            t = std::fmin(std::fmax(t, 0.0f), 1.0f);
            const float ss = t * t * (3.0f - 2.0f * t);
            const float f = ss * 1.5f * 26.0f + 5.0f;
            table[(size_t)(dy + P) * S + (dx + P)] =
                (float)(-3.0 * (double)f * 1.4426950408889634);
        }
}

}  // namespace pmh
