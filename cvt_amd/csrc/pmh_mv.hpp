// pmh_mv.hpp -- host-side declarations of the multi-view slanted-plane
// PatchMatch (kernels in pmh_mv.hip, C-ABI in pmh_capi.cpp).
//
// The reference holds a draft of this method that does not compile
// (cvt/cl/kernel/pmmvstereo.cl:97-304, cvt/cl/kernel/Plane.cl:60-86).  This
// build fixes its conventions:
//   - a particle is float4 (nx, ny, d, cost): the plane n.X = d in the
//     reference camera frame, n = (nx, ny, sqrt(1 - nx^2 - ny^2));
//   - cameras are K (3x3) and T = [R | t] (3x4, world -> camera), row-major;
//     for source view s, R = R_s R_r^T, t = t_s - R t_r, and the plane-induced
//     homography H = K_s (R + t n^T / d) K_r^-1 = M_s + e_s q^T with the view
//     constants M_s = K_s R K_r^-1, e_s = K_s t (computed in double, rounded to
//     float) and q = K_r^-T n / d;
//   - source views are sampled bilinearly from texel records (I, gx, gy, gxy,
//     gyx) whose rows and columns carry one repeated texel past the last, so
//     that the four corners of a sample are two runs of 10 floats; the cell
//     (t, dx, dy, dxy) of each channel is formed from them as the
//     restatement forms it.
#pragma once

#include <hip/hip_runtime.h>

#include "pmh_internal.hpp"

namespace pmh {

constexpr int kMVMaxViews = 4;   // source views per scene
constexpr int kMVDraws = 33;     // RNG_init stride of pmmvstereo.cl:237
constexpr int kMVNbhd = 3;       // NUMRNDSAMPLE, pmmvstereo.cl:35
constexpr int kMVMaxRefine = 8;  // 6 + 3 + 3 * 8 = 33 draws
constexpr int kMVTexel = 5;      // floats per texel record: intensity + 4 gradient channels

struct MVConst {
    float kinv[9];                // K_r^-1
    float M[kMVMaxViews][9];      // K_s R K_r^-1
    float e[kMVMaxViews][3];      // K_s t
};

// Device images of a scene.
struct MVImages {
    const float* Ir;      // reference gray image
    const float4* Gr;     // its gradient image (dx, dy, dxy, dyx)
    const float* tex;     // source view v: tex + v * tstride, (W + 1) x (H + 1) records
    int W, H;
    unsigned tstride;     // floats per source view, (W + 1) * (H + 1) * kMVTexel
    unsigned pad_ = 0;
};
inline size_t mv_texel_floats(int W, int H) { return (size_t)(W + 1) * (H + 1) * kMVTexel; }

// View constants from the cameras (double, rounded to float).
void mv_constants(const float* Kr, const float* Tr, int nsrc, const float* Ks, const float* Ts,
                  MVConst& out);

// Weight exponents of the multi-view support factors (smoothstep(0, 28, |o|)
// * 1.5 * 26 + 5, pmmvstereo.cl:119) folded as -3 F log2(e): (2P+1)^2 floats.
void mv_factor_table(int P, float* table);

// Texel records of one source view from its image and gradient image.
void launch_mv_texels(const float* img, const float4* grad, int W, int H, float* tex,
                      hipStream_t s);
void launch_mv_init(const MVImages& im, const MVConst& k, int nsrc, const float* Fw, int P,
                    float dmin, float dmax, float4* out, hipStream_t s);
void launch_mv_cost(const MVImages& im, const MVConst& k, int nsrc, const float* Fw, int P,
                    const float4* particles, float* cost, hipStream_t s);
void launch_mv_propagate(const MVImages& im, const MVConst& k, int nsrc, const float* Fw, int P,
                         const float4* old, float4* out, float dmin, float dmax, int iter,
                         const float* draws, int sweep, int nref, hipStream_t s);
void launch_mv_depth(const float4* particles, const MVConst& k, int W, int H, float* depth,
                     hipStream_t s);
void launch_mv_homography(const float4* particles, int count, const MVConst& k, int nsrc,
                          float* H, hipStream_t s);

}  // namespace pmh
