/*
 * pmh_oracle.h -- CPU restatement of CVT's slanted-plane PatchMatch stereo path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (cvt_amd/, libpmh_amd.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference paths
 * are relative to the mherb/cvt source tree:
 *   cvt/cl/kernel/pmhstereo.cl, cvt/cl/kernel/RNG.cl,
 *   cvt/cl/kernel/PDHuberWeighted.cl, cvt/gfx/PDROF.h,
 *   cvt/vision/PMHuberStereo.cpp, cvt/math/Math.h; the multi-view functions
 *   (pmh_oracle_mv.c) follow the draft cvt/cl/kernel/pmmvstereo.cl with
 *   cvt/cl/kernel/Plane.cl.
 *
 * Arithmetic conventions (fixed here because OpenCL leaves them open):
 *   - compiled with -ffp-contract=off; every fused multiply-add is an explicit
 *     fmaf() so that the device code can follow the same rounding;
 *   - nd_state_transform(s, (x, y)) = fmaf(s.a, x, fmaf(s.b, y, s.c));
 *   - native_sqrt -> sqrtf, native_rsqrt -> 1/sqrtf, native_exp -> expf;
 *   - bilinear sampling of the other view at x' is T0 + a*(T1 - T0) with
 *     i0 = floor(x'), a = x' - i0, T1 clamped to the last column;
 *   - exp/pow of the weight image and acospi of the consistency check are
 *     evaluated in double and rounded to float;
 *   - a non-finite transformed coordinate counts as outside the image;
 *   - denormals are kept: the reference builds its kernels with
 *     -cl-denorms-are-zero (cvt/cl/CLKernel.h:97), which flushes them.  The
 *     window cost divides sums with sum(w) > 1.1 and the weights are
 *     exp(-sad F) with F <= 44, so a flushed denormal changes a cost by less
 *     than one ulp of the sums it is added to; the plane algebra and the
 *     smoother never produce denormals from finite non-denormal inputs except
 *     in the last bits of near-zero differences;
 *   - the other view is sampled at x' itself: the reference samples at
 *     (x' + 0.5, y + 0.5) with a linear filter (pmhstereo.cl:214-215), whose
 *     OpenCL 1.2 definition subtracts 0.5 again; in float the round trip
 *     (x' + 0.5) - 0.5 returns x' exactly unless x' + 0.5 crosses into the
 *     next binade (then the last bit of x' is lost), and the vendor's filter
 *     hardware rounds the fraction to its own precision, which no
 *     specification fixes -- this restatement takes the exact fraction;
 *   - the random-neighbourhood offset draws x before y: the reference draws
 *     both inside one vector literal (pmhstereo.cl:356), whose evaluation
 *     order C leaves unspecified; the clang-based OpenCL C front ends of CPU
 *     OpenCL runtimes evaluate it left to right (checked on this image's
 *     clang by tests/test_oracle_conventions.py, which also shows that
 *     contracting r * 7 + 0.5 into an FMA changes no offset).
 *
 * Images are dense row-major float arrays (pitch = width).  An input image has
 * nc floats per pixel: nc = 1 for gray (CL_INTENSITY), nc = 4 for RGBA
 * (CL_RGBA); a gradient image is W*H*4 floats (dx, dy, dxy, dyx), a plane
 * field W*H*4 floats (a, b, c, cost) with x' = a*x + b*y + c.  The FP64
 * gradient functions take gray images only.
 */
#ifndef PMH_ORACLE_H
#define PMH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_VIEWSAMPLES 4
#define ORC_PROP_DRAWS 21 /* draws consumed per pixel by one propagate */

enum { ORC_SWEEP_ALL = 0, ORC_SWEEP_RED = 1, ORC_SWEEP_BLACK = 2 };

typedef struct {
    int n;                        /* number of scatters received (may exceed 4) */
    int src[ORC_VIEWSAMPLES];     /* source x of the first min(n,4), ascending */
} orc_viewbuf;

/* ---- MWC64X (RNG.cl:35-156) ---- */
uint64_t orc_mwc_mulmod(uint64_t a, uint64_t b);           /* RNG.cl:49-59 */
uint64_t orc_mwc_powmod(uint64_t a, uint64_t e);           /* RNG.cl:66-76 */
void orc_rng_init(uint64_t linpos, uint32_t numdraws, uint32_t st[2]); /* RNG.cl:151-154 */
void orc_rng_init_u128(uint64_t linpos, uint32_t numdraws, uint32_t st[2]);
float orc_rng_float(uint32_t st[2]);                       /* RNG.cl:142-147 */
void orc_rng_step(uint32_t st[2]);                         /* RNG.cl:108-118 */
void orc_rng_draws(uint64_t linpos, uint32_t numdraws, int count, float* out);

/* ---- plane algebra (pmhstereo.cl:50-153) ---- */
void orc_state_init(const float u[3], float cx, float cy, int lr, float normmul,
                    float depthmax, float out[4]);
void orc_state_viewprop(const float s[4], float out[4]);
float orc_state_transform(const float s[4], float cx, float cy);
void orc_state_to_ref_normal_depth(const float s[4], float cx, float cy, int lr, float out[4]);
void orc_state_refine(const float u[3], const float s[4], float cx, float cy,
                      float depthmax, int lr, float out[4]);
void orc_state_to_normal(const float s[4], float out[3]);

/* ---- pre-processing (pmhstereo.cl:737-831) ---- */
void orc_gradxy(const float* img, int W, int H, int nc, float* grad);
void orc_weight(const float* img, int W, int H, int nc, float* weight);
void orc_factor_table(int P, float* table);

/* ---- window cost (pmhstereo.cl:170-228) ---- */
float orc_cost(const float* I1, const float* G1, const float* I2, const float* G2,
               int W, int H, int nc, int x, int y, const float s[4], int P);
void orc_cost_field(const float* I1, const float* G1, const float* I2, const float* G2,
                    int W, int H, int nc, const float* states, int P, float* cost);
/* Same window cost in double precision with its exact gradient w.r.t. the
 * plane (a, b, c) and both images (adjoint reference); nc = 1 or 4 as
 * orc_cost, dI1 / dI2 nc values per pixel (RGBA alpha gets none). */
double orc_cost_grad_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                         int W, int H, int nc, int x, int y, const double s[3], int P,
                         double dplane[3], double* dI1, double* dG1, double* dI2, double* dG2);
void orc_cost_grad_field_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                             int W, int H, int nc, const float* states, int P, const float* gout,
                             double* dplane, double* dI1, double* dG1, double* dI2, double* dG2);
double orc_cost_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                    int W, int H, int nc, int x, int y, const double s[3], int P);

/* ---- kernels of the iteration ---- */
void orc_init(const float* I1, const float* G1, const float* I2, const float* G2,
              int W, int H, int nc, int P, float depthmax, int lr, float* out);
void orc_propagate(const float* old, float* out,
                   const float* I1, const float* G1, const float* I2, const float* G2,
                   const float* smooth, float theta, int W, int H, int nc, int P,
                   float depthmax, int lr, int iter, const orc_viewbuf* viewin,
                   const float* viewsrc, const float* draws, int sweep);
/* The same with `nbhd` random-neighbourhood candidates (the reference's 6, or 0)
 * and `nref` randomized refinements (the reference's 2, or 0, 1): the options
 * random_neighbours and random_refine of the batched method. */
/* Gradient plane refinement of a plane field in place (the product's
 * directed refinement; no reference counterpart). */
void orc_grad_refine(float* states, const float* I1, const float* G1, const float* I2,
                     const float* G2, const float* smooth, int W, int H, int nc, int P,
                     float depthmax, int lr, float theta, int steps, float radius);
void orc_propagate_ex(const float* old, float* out,
                      const float* I1, const float* G1, const float* I2, const float* G2,
                      const float* smooth, float theta, int W, int H, int nc, int P,
                      float depthmax, int lr, int iter, const orc_viewbuf* viewin,
                      const float* viewsrc, const float* draws, int sweep, int nbhd, int nref);
void orc_viewbuf_build(const float* states, int W, int H, orc_viewbuf* out);
void orc_consistency(const float* left, const float* right, int W, int H,
                     float maxdispdiff, float maxdeg, int lr, float* out);
void orc_fill_state(const float* in, int W, int H, float depthmax, int lr, float* out);
void orc_pd_step(const float* last, const float* plast, const float* f, const float* w,
                 int W, int H, float lambda, float* out, float* pout);
void orc_pdrof(float* u, const float* f, const float* w, int W, int H, float lambda, int iters);
void orc_fill_depthmap(const float* in, int W, int H, float scale, float* dmap);
/* depthMapInpaint's kernels (pmhstereo.cl:663-712, PDHuberWeightedInpaint.cl,
 * PDROFInpaint.h:50-98) */
void orc_occmap(const float* left, const float* right, int W, int H, float maxdiff,
                float maxdeg, int lr, float* out);
void orc_normal_depth(const float* in, int W, int H, float depthmax, int lr, float* out);
void orc_pd_step_inpaint(const float* last, const float* plast, const float* f, const float* w,
                         const float* mask, int W, int H, float lambda, float* out, float* pout);
void orc_pdrof_inpaint(float* u, const float* f, const float* w, const float* mask, int W, int H,
                       float lambda, int iters);
void orc_fill_normalmap(const float* in, int W, int H, float* nmap);

/* ---- PMHuberStereo::depthMap (PMHuberStereo.cpp:64-291, minus :153-185) ---- */
typedef struct {
    float maxdispdiff;   /* 0.5 in the reference (PMHuberStereo.cpp:70) */
    float maxanglediff;  /* 10 degrees (PMHuberStereo.cpp:71) */
    int pd_iters;        /* 250 (PMHuberStereo.cpp:243) */
    int sweep;           /* 0 = Jacobi over all pixels (reference), 1 = red-black */
    int random_neighbours; /* random-neighbourhood candidates per pixel: 6 (reference) or 0 */
    int random_refine;     /* randomized refinements per pixel: 2 (reference), 1 or 0 */
} orc_opts;

void orc_depthmap(const float* left, const float* right, int W, int H, int nc, int P,
                  float depthmax, int iterations, float dscale, const orc_opts* opts,
                  float* dmap, float* normalmap, float* lstate_out, float* rstate_out);

/* ---- PMHuberStereo::depthMapInpaint (PMHuberStereo.cpp:293-507) ----
 * opts->maxdispdiff / maxanglediff: occlusion-map tolerances (0.5, 5);
 * opts->pd_iters: smoother steps (100); opts->sweep as for orc_depthmap. */
void orc_depthmap_inpaint(const float* left, const float* right, int W, int H, int nc, int P,
                          float depthmax, int iterations, const orc_opts* opts, float* dmap,
                          float* lstate_out, float* rstate_out);

/* ---- Multi-view PatchMatch (pmh_oracle_mv.c; the reference's draft is
 * pmmvstereo.cl:97-304 with Plane.cl:60-86, which does not compile; the
 * conventions fixed here are listed at the top of pmh_oracle_mv.c) ----
 * Cameras: K row-major 3x3, T = [R | t] row-major 3x4 (world -> camera).
 * A particle is (nx, ny, d, cost): the plane n.X = d in the reference camera.
 * Source views are stacked: Is = nsrc gray images, Gs = nsrc gradient images;
 * kinv[9] = K_r^-1, M[9 nsrc] = K_s R K_r^-1, e[3 nsrc] = K_s t. */
#define ORC_MV_DRAWS 33 /* RNG_init stride of pmmvstereo.cl:237 */
void orc_mv_constants(const float* Kr, const float* Tr, int nsrc, const float* Ks,
                      const float* Ts, float* kinv, float* M, float* e);
void orc_mv_homography(const float p[4], const float* kinv, const float* M, const float* e,
                       float H[9]);
float orc_mv_depth(const float p[4], const float* kinv, float x, float y);
void orc_mv_cells(const float* img, int W, int H, int nch, int c, float* cells);
float orc_mv_bilinear(const float* img, int W, int H, int nch, int c, float x, float y);
void orc_mv_factor_table(int P, float* table);
float orc_mv_cost(const float* Ir, const float* Gr, int nsrc, const float* Is, const float* Gs,
                  int W, int H, int x, int y, const float p[4], int P, const float* kinv,
                  const float* M, const float* e);
void orc_mv_particle_init(const float u[3], float x, float y, float nm, float dmin, float dmax,
                          const float* kinv, float out[4]);
void orc_mv_particle_refine(const float u[3], const float p[4], float x, float y, float dmin,
                            float dmax, const float* kinv, float out[4]);
void orc_mv_init(const float* Ir, const float* Gr, int nsrc, const float* Is, const float* Gs,
                 int W, int H, int P, float dmin, float dmax, const float* kinv, const float* M,
                 const float* e, float* out);
void orc_mv_propagate(const float* old, float* out, const float* Ir, const float* Gr, int nsrc,
                      const float* Is, const float* Gs, int W, int H, int P, float dmin,
                      float dmax, int iter, const float* kinv, const float* M, const float* e,
                      const float* draws, int sweep, int nref);
void orc_mv_depthmap(const float* Ir, const float* Gr, int nsrc, const float* Is,
                     const float* Gs, int W, int H, int P, float dmin, float dmax,
                     int iterations, int rb, int nref, const float* kinv, const float* M,
                     const float* e, float* depth, float* particles);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
