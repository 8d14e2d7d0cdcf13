/*
 * pmh_capi.h -- C-ABI of the MI355X (gfx950) slanted-plane PatchMatch stereo
 * path.  Drop-in boundary for cvt::PMHuberStereo (cvt/vision/PMHuberStereo.h)
 * and the OpenCL kernels it launches (cvt/cl/kernel/pmhstereo.cl,
 * cvt/cl/kernel/PDHuberWeighted.cl, cvt/gfx/PDROF.h).
 *
 * Conventions
 *   - All image and field pointers are DEVICE pointers (HBM resident), dense
 *     row-major with pitch == width.  Input images are PMH_FORMAT_GRAY
 *     (width*height floats; the reference's GRAY_FLOAT / CL_INTENSITY) or
 *     PMH_FORMAT_RGBA (width*height float4; RGBA_FLOAT / CL_RGBA, colour sums
 *     over R, G, B as in pmhstereo.cl:199,217, IFormat.h:302,312).
 *     Gradient images: float4 (dx, dy, dxy, dyx).  Plane fields: float4
 *     (a, b, c, cost), mapping pixel x to x' = a*x + b*y + c in the other view.
 *   - Work is enqueued on the context's HIP stream in order.  pmh_depthmap and
 *     pmh_batch_depthmap return after the stream has drained (as the
 *     reference's trailing runWait does, PMHuberStereo.cpp:282); the per-kernel
 *     entry points return after enqueueing (call pmh_synchronize).
 *   - Every function returns PMH_OK (0) or a negative pmh_status; the message
 *     of the last failure on the calling thread is pmh_last_error().  No C++
 *     exception crosses this boundary (the reference throws CVTException /
 *     CLException, cvt/util/Exception.h:35, cvt/cl/CLException.h:34-44; the
 *     C++ facade cvt::PMHuberStereo maps the status back to an exception).
 *   - One pmh_ctx per thread; contexts are independent.
 */
#ifndef PMH_CAPI_H
#define PMH_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmh_ctx pmh_ctx;

typedef enum {
    PMH_OK = 0,
    PMH_ERR_INVALID = -1,     /* bad argument (size mismatch, null pointer, ...) */
    PMH_ERR_HIP = -2,         /* HIP runtime error */
    PMH_ERR_NOMEM = -3,       /* device allocation failed */
    PMH_ERR_UNSUPPORTED = -4  /* valid request outside what this build supports */
} pmh_status;

enum { PMH_SWEEP_ALL = 0, PMH_SWEEP_RED = 1, PMH_SWEEP_BLACK = 2, PMH_SWEEP_REDBLACK = 3 };
enum { PMH_FORMAT_GRAY = 0, PMH_FORMAT_RGBA = 1 };

/* Options of one depth-map run.  pmh_default_options gives the reference's
 * constants (PMHuberStereo.cpp:70-72, :243, :261). */
typedef struct {
    float maxdispdiff;   /* L/R disparity tolerance in px (reference 0.5) */
    float maxanglediff;  /* L/R normal tolerance in degrees (reference 10) */
    int pd_iters;        /* Huber-ROF iterations per view per iteration (250) */
    int sweep;           /* PMH_SWEEP_ALL (reference Jacobi) or PMH_SWEEP_REDBLACK */
    int grad_steps;      /* gradient plane-refinement steps per view per iteration (0) */
    float grad_step;     /* largest move of x' in the window by the first gradient step, px (0.5) */
    int random_refine;   /* randomized refinements per pixel per pass (reference: 2) */
    int use_graph;       /* replay the whole run as one captured HIP graph (1) or launch
                            eagerly (0); the first run of a shape is eager (it tunes) */
    /* Coarse-to-fine start (no counterpart in the reference, which starts from
     * random planes): the images are reduced coarse_levels times by 2x2 box
     * averages, the coarsest level starts from random planes, every level runs
     * coarse_iters iterations and hands its planes up as the next level's
     * start; the full-resolution level then runs `iterations` iterations.
     * 0 levels (the default) is the reference's method. */
    int coarse_levels;   /* 0 .. 4 */
    int coarse_iters;    /* iterations per coarse level (4) */
    int coarse_patch;    /* half window at the coarse levels; 0: the same as patchsize */
    int coarse_pd_iters; /* Huber-ROF iterations at the coarse levels; 0: pd_iters */
    int coarse_random_refine; /* randomized refinements at the coarse levels (0 .. 2);
                                 -1 (the default): random_refine */
    int random_neighbours;    /* random-neighbourhood candidates per pixel per pass,
                                 pmhstereo.cl:355-361: 6 (the reference) or 0 (none) */
    int coarse_random_neighbours; /* the same at the coarse levels; -1 (the default):
                                     random_neighbours */
    int coarse_joint_views;   /* 1: at the coarse levels both views propagate (and refine, and
                                 scatter) in one launch per sweep, each reading the other
                                 view's buffer of the previous iteration; 0 (the default): the
                                 reference's order (left, then right reading the left's new
                                 scatter) at every level.  The full-resolution level always
                                 keeps the reference's order. */
} pmh_options;

/* One rectified pair of a batched run (pmh_batch_depthmap). */
typedef struct {
    const float* left;     /* device image of `format` */
    const float* right;
    int width, height;
    int format;            /* PMH_FORMAT_GRAY or PMH_FORMAT_RGBA, the same for all pairs */
    float depthmax;        /* disparity range of this pair */
    float dscale;          /* <= 0: 1/depthmax */
    float* dmap;           /* width*height floats (device) */
    float* normalmap;      /* NULL or width*height float4 */
    float* left_states;    /* NULL or width*height float4: final plane fields */
    float* right_states;
} pmh_pair;

/* ---- context ------------------------------------------------------------ */
/* Creates a context on the current HIP device.  stream: the hipStream_t to
 * enqueue on; NULL is the device's default (null) stream, the in-order queue
 * the reference uses for everything.  Replaces the reference's
 * CL::defaultContext/defaultQueue singletons (cvt/cl/OpenCL.cpp:36-46) and the
 * 13 runtime-compiled CLKernel members (PMHuberStereo.cpp:42-58). */
int pmh_create(pmh_ctx** ctx, void* stream);
int pmh_destroy(pmh_ctx* ctx);
int pmh_synchronize(pmh_ctx* ctx);
const char* pmh_last_error(void);
void pmh_default_options(pmh_options* opts);

/* ---- the drop-in entry point -------------------------------------------- */
/* PMHuberStereo::depthMap (PMHuberStereo.h:40, PMHuberStereo.cpp:64-291).
 * left/right: gray float images (device).  dmap: width*height floats,
 * |x' - x| * dscale (dscale <= 0 means 1/depthmax, :75-76).  normalmap: NULL
 * or width*height float4 (fill_normalmap, :285-290).  patchsize is the half
 * window (loops dy,dx in [-p, p]).  viewsamples is accepted and ignored, as in
 * the reference (pmhstereo.cl:40 hard-codes 4).  left_states/right_states:
 * NULL or width*height float4 buffers that receive the final plane fields. */
int pmh_depthmap(pmh_ctx* ctx, const float* left, const float* right, int width, int height,
                 int patchsize, float depthmax, int iterations, int viewsamples, float dscale,
                 float* dmap, float* normalmap, const pmh_options* opts, float* left_states,
                 float* right_states);

/* PMHuberStereo::depthMapInpaint (PMHuberStereo.h:41, PMHuberStereo.cpp:293-507,
 * without the debug block :376-409).  The depthMap propagation, but the smooth
 * prior comes from an occlusion map (pmhstereo_occmap, pmhstereo.cl:682-712)
 * and the masked Huber-ROF of PDROFInpaint (PDROFInpaint.h:50-98,
 * PDHuberWeightedInpaint.cl) with lambda = 20 theta + 5, theta rising from
 * iteration 4; the final L/R check (1 px, 5 degrees) reads the plane fields of
 * index 1 as the reference does.  dmap holds |x' - x| in px.  Options:
 * maxdispdiff / maxanglediff are the occlusion map's tolerances and pd_iters
 * the smoother's steps; pmh_default_inpaint_options gives the reference's
 * (0.5 px, 5 degrees, 100).  iterations >= 1. */
int pmh_depthmap_inpaint(pmh_ctx* ctx, const float* left, const float* right, int width,
                         int height, int patchsize, float depthmax, int iterations,
                         int viewsamples, float* dmap, const pmh_options* opts,
                         float* left_states, float* right_states);
void pmh_default_inpaint_options(pmh_options* opts);
/* depthMapInpaint for many pairs in one run (each pair's dscale is ignored). */
int pmh_batch_depthmap_inpaint(pmh_ctx* ctx, const pmh_pair* pairs, int npairs, int patchsize,
                               int iterations, const pmh_options* opts);

/* Many pairs of uneven size and disparity range in one run (no counterpart
 * in the reference, whose callers loop over depthMap): every launch of the
 * depthMap sequence covers all pairs, so small pairs share the GPU, and with
 * opts->use_graph the whole run is one replayed HIP graph.  Each pair's
 * result equals pmh_depthmap on that pair alone with the same options. */
int pmh_batch_depthmap(pmh_ctx* ctx, const pmh_pair* pairs, int npairs, int patchsize,
                       int iterations, const pmh_options* opts);

/* ---- launch-shape autotuner (no counterpart in the reference) ------------
 * The first batched or depthMap run of a shape class times the propagate
 * launch over its kernel shapes (1 tile per block, and the default shape with
 * persistent blocks or with 1 tile per block), and the fused Huber-ROF steps over 1, 2, 4 or 8 tiles
 * per block and persistent blocks, on its own inputs, and keeps the fastest.
 * Persistent blocks (chunk 0): the grid holds as many blocks as the GPU runs
 * at once and each block takes tiles from a queue per XCD band of the work
 * list, then from the other bands.  Propagate shapes (stage codes): one
 * wave per row segment with an LDS stage of 512, 704 or 896 texels per window
 * row; 1: a row segment's candidates spread over four waves, one window sweep
 * per candidate, at the format's stage width (896 gray, 512 RGBA); 2: the
 * same with each wave sweeping its candidates together; 3: code 2 at a 576 /
 * 352-texel stage with three waves per SIMD; 4 (gray only): the four waves share one LDS stage of
 * 1792 texels per window row that carries each texel's bilinear differences
 * and sweep the window rows in lockstep, one wave costing the right-side
 * neighbour planes for two pixels at once; 5: code 4 at 1152 texels with three
 * waves per SIMD (without the shared planes).  Propagate launches are tuned per
 * view, and the first iteration apart from the later ones.  The choices never
 * change results.
 * enable = 0 uses the defaults (stage code 4 for gray inputs, 2 for RGBA;
 * persistent blocks for propagate, 1 tile per block otherwise).  A tuning file
 * holds one "key stage chunk" line per shape class (stage 0: the default;
 * chunk 0: persistent blocks). */
int pmh_autotune_enable(pmh_ctx* ctx, int enable);
/* Largest sample of a launch's work list the autotuner times per candidate
 * (every stride-th tile of the launch); 0 restores the defaults (16384 tiles
 * for propagate launches, 65536 for the others), otherwise 8 .. 2^30.  A
 * launch of at least twice the sample is scored by its marginal time,
 * time(sample) - time(half the sample).  Smaller samples tune faster and
 * judge big launches by fewer rounds of blocks. */
int pmh_autotune_sample(pmh_ctx* ctx, int tiles);
int pmh_tuning_load(pmh_ctx* ctx, const char* path);
int pmh_tuning_save(pmh_ctx* ctx, const char* path);
/* graphs cached by the context and graph replays so far */
int pmh_graph_stats(pmh_ctx* ctx, int* graphs_cached, long long* graph_launches);

/* ---- per-kernel entry points (each replaces one reference kernel) -------- */
/* Every entry point below that reads input images takes their `format`
 * (PMH_FORMAT_GRAY or PMH_FORMAT_RGBA) as its last argument. */
/* pmhstereo_gradxy, pmhstereo.cl:785-831 */
int pmh_gradxy(pmh_ctx* ctx, const float* img, int width, int height, float* grad, int format);
/* pmhstereo_weight, pmhstereo.cl:737-783 */
int pmh_weight(pmh_ctx* ctx, const float* img, int width, int height, float* weight,
               int format);
/* RNG_init + RNG_float of RNG.cl:142-156 as used at pmhstereo.cl:241 (stride 3,
 * iter 0), :327 (stride 45) and pmmvstereo.cl:237 (stride 33): count draws per
 * pixel, linpos = y*W + x + iter */
int pmh_rng_draws(pmh_ctx* ctx, int width, int height, int iter, int stride, int count,
                  float* out);
/* pmhstereo_init, pmhstereo.cl:230-248 (lr = 1 left view, 0 right view) */
int pmh_init(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2, const float* G2,
             int width, int height, int patchsize, float depthmax, int lr, float* out,
             int format);
/* patch_eval_color_grad_weighted, pmhstereo.cl:170-228, for a supplied plane
 * field: cost[y*W+x] = window cost of states[y*W+x] at (x, y). */
int pmh_cost(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2, const float* G2,
             int width, int height, int patchsize, const float* states, float* cost,
             int format);
/* pmhstereo_propagate_view, pmhstereo.cl:286-404 minus the scatter (see
 * pmh_viewbuf_build).  viewin: NULL or width*height records of 5 int32
 * {n, src0..src3} whose sources index viewsrc (the other view's field) on the
 * same row.  draws: NULL (MWC64X draws of :327) or 21 floats per pixel.
 * sweep: PMH_SWEEP_ALL (Jacobi, the reference), RED or BLACK (a black pass
 * reads the red pixels of out).  old and out must be distinct fields. */
int pmh_propagate(pmh_ctx* ctx, const float* old, float* out, const float* I1, const float* G1,
                  const float* I2, const float* G2, const float* smooth, float theta, int width,
                  int height, int patchsize, float depthmax, int lr, int iter, const int* viewin,
                  const float* viewsrc, const float* draws, int sweep, int format);
/* The view scatter of pmhstereo.cl:391-401 with the sequential slot order
 * (first four sources by ascending x). */
int pmh_viewbuf_build(pmh_ctx* ctx, const float* states, int width, int height, int* viewbuf);
/* pmhstereo_consistency, pmhstereo.cl:428-464 */
int pmh_consistency(pmh_ctx* ctx, const float* left, const float* right, int width, int height,
                    float maxdispdiff, float maxanglediff, int lr, float* out);
/* pmhstereo_fill_state, pmhstereo.cl:466-510 */
int pmh_fill_state(pmh_ctx* ctx, const float* in, int width, int height, float depthmax, int lr,
                   float* out);
/* One PDHuberWeighted step (PDHuberWeighted.cl:34-104).  p buffers hold
 * 2*width*height float4 (lo, hi interleaved per pixel, as the reference's
 * 2W-wide p image, PDROF.h:54-55). */
int pmh_pd_step(pmh_ctx* ctx, const float* last, const float* plast, const float* f,
                const float* w, int width, int height, float lambda, float* out, float* pout);
/* PDROF::apply (PDROF.h:50-96): u holds the warm start and receives the
 * result; iters >= 1 (with 0 the reference reads an unset dual buffer). */
int pmh_pdrof(pmh_ctx* ctx, float* u, const float* f, const float* w, int width, int height,
              float lambda, int iters);
/* pmh_pdrof without device allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same steps and bits as
 * pmh_pdrof; the temporary image and both duals live in one device array per
 * GPU, sized for images of up to 2^20 pixels (80 bytes per pixel, 80 MiB, held
 * from library load to process end; larger images return PMH_ERR_INVALID).
 * Parameters go to the kernels by value: no host-to-device copy, no
 * allocation.  Calls on one GPU share that array, so calls that overlap on
 * one GPU (other streams, other contexts) race; the dual the first step reads
 * is cleared by a fill kernel on every call.  Enqueues on ctx's stream and
 * returns; ctx's device must be the current device. */
int pmh_pdrof_noalloc(pmh_ctx* ctx, float* u, const float* f, const float* w, int width,
                      int height, float lambda, int iters);
/* pmhstereo_occmap, pmhstereo.cl:682-712: (v, v, v, 1), v = 1 where the
 * consistency test passes (maxanglediff: the kernel hard-codes 5 degrees). */
int pmh_occmap(pmh_ctx* ctx, const float* left, const float* right, int width, int height,
               float maxdispdiff, float maxanglediff, int lr, float* out);
/* pmhstereo_normal_depth, pmhstereo.cl:663-680 */
int pmh_normal_depth(pmh_ctx* ctx, const float* in, int width, int height, float depthmax,
                     int lr, float* out);
/* One PDHuberWeightedInpaint step (PDHuberWeightedInpaint.cl:34-104); mask:
 * width*height float4, the data term's lambda is scaled by mask.x. */
int pmh_pd_step_inpaint(pmh_ctx* ctx, const float* last, const float* plast, const float* f,
                        const float* w, const float* mask, int width, int height, float lambda,
                        float* out, float* pout);
/* PDROFInpaint::apply (PDROFInpaint.h:50-98); u: warm start in, result out. */
int pmh_pdrof_inpaint(pmh_ctx* ctx, float* u, const float* f, const float* w, const float* mask,
                      int width, int height, float lambda, int iters);
/* pmhstereo_fill_depthmap, pmhstereo.cl:512-555 */
int pmh_fill_depthmap(pmh_ctx* ctx, const float* in, int width, int height, float scale,
                      float* dmap);
/* pmhstereo_fill_normalmap, pmhstereo.cl:557-603 */
int pmh_fill_normalmap(pmh_ctx* ctx, const float* in, int width, int height, float* nmap);

/* ---- coarse-to-fine start (no counterpart in the reference) --------------
 * pmh_pyr_down: out ((W+1)/2 x (H+1)/2) = 2x2 box average of img, the last
 * row/column repeated for odd sizes.  pmh_plane_up: a plane field of the
 * half-resolution level (w x h) carried to W x H: pixel (x, y) takes the plane
 * of (x/2, y/2) with c -> 2c + (1 - a - b)/2, cost 0. */
int pmh_pyr_down(pmh_ctx* ctx, const float* img, int width, int height, float* out, int format);
int pmh_plane_up(pmh_ctx* ctx, const float* in, int w, int h, int width, int height, float* out);

/* ---- adjoint of the window cost (no counterpart in the reference) --------
 * For L = sum_p gout[p] * cost[p], with cost the field of pmh_cost:
 * dstates[p] = (dL/da, dL/db, dL/dc, 0) of pixel p's plane; when non-NULL,
 * dI1, dI2 (W*H values, W*H float4 for PMH_FORMAT_RGBA, whose alpha gets
 * none), dG1, dG2 (W*H float4) receive dL/d(texel), ADDED to their contents
 * (zero them first).  Windows are recomputed, nothing is stored by the
 * forward call.  The derivative of min(|d|, t) is sign(d) below t. */
int pmh_cost_backward(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2,
                      const float* G2, int width, int height, int patchsize, const float* states,
                      const float* gout, float* dstates, float* dI1, float* dG1, float* dI2,
                      float* dG2, int format);
/* Adjoint of pmhstereo_gradxy: dI += gradxy^T(dG), chaining a gradient
 * image's adjoint to its intensity image. */
int pmh_gradxy_backward(pmh_ctx* ctx, const float* dG, int width, int height, float* dI);
/* Gradient-step plane refinement (no counterpart in the reference): `steps`
 * descent steps of each pixel's window cost, the first moving x' by at most
 * `radius` px anywhere in the window, the radius halving per step; a step is
 * kept under the reference's acceptance rule (pmhstereo.cl:340).  states is
 * updated in place. */
int pmh_grad_refine(pmh_ctx* ctx, float* states, const float* I1, const float* G1,
                    const float* I2, const float* G2, const float* smooth, float theta, int width,
                    int height, int patchsize, float depthmax, int lr, int steps, float radius,
                    int format);

/* ---- multi-view PatchMatch ------------------------------------------------
 * The reference holds a draft of a homography-based multi-view PatchMatch
 * (cvt/cl/kernel/pmmvstereo.cl:97-304 with cvt/cl/kernel/Plane.cl:60-86) that
 * has no host code and does not compile; these entry points stand in for its
 * kernels with its structure and constants and these conventions:
 *   - a particle field is width*height float4 (nx, ny, d, cost): the plane
 *     n.X = d in the reference camera frame, n = (nx, ny, sqrt(1-nx^2-ny^2));
 *   - cameras: K row-major 3x3, T = [R | t] row-major 3x4 (world -> camera);
 *     source view s sees the plane through H = K_s (R + t n^T / d) K_r^-1 with
 *     R = R_s R_r^T, t = t_s - R t_r (the draft's planef_homography, whose
 *     sign of d it marks FIXME, Plane.cl:29);
 *   - the window cost sums, over taps whose plane point lies in front of the
 *     reference camera and every source view that sees it in front of itself
 *     inside its image, the support weight (smoothstep(0, 28, |offset|) * 39
 *     + 5, pmmvstereo.cl:119) times the truncated colour + gradient cost of
 *     pmhstereo.cl:199-222 with 2-D bilinear samples; 1e5 if the weights sum
 *     to at most 1.1.  The draft tests the opposite of "inside"
 *     (pmmvstereo.cl:127,142); this build counts the taps inside;
 *   - depths are drawn in [dmin, dmax] at the pixel; the depth of a particle
 *     at (x, y) is d / (n . K_r^-1 (x, y, 1)), 0 where that is not a finite
 *     positive number.
 * Images are gray (width*height floats), all views of one size, 1 to 4
 * source views (the draft: 2). */
typedef struct {
    const float* image;  /* device gray image, width*height floats */
    float K[9];          /* intrinsics, row-major */
    float T[12];         /* [R | t], world -> camera, row-major 3x4 */
} pmh_mv_view;

typedef struct pmh_mv_scene pmh_mv_scene;

/* Prepares a scene on ctx's device: a copy of the reference image, its
 * gradient image (pmhstereo_gradxy) and texel records (intensity and
 * gradient, 20 bytes per pixel per view) of every source view.  Returns
 * after the stream has drained; the caller's images may then be freed. */
int pmh_mv_scene_create(pmh_ctx* ctx, const pmh_mv_view* ref, const pmh_mv_view* srcs,
                        int nsrc, int width, int height, pmh_mv_scene** scene);
int pmh_mv_scene_destroy(pmh_mv_scene* scene);
/* pmmv3_init, pmmvstereo.cl:170-195: RNG_init(y*W + x, 3), a random particle
 * and its cost.  out: width*height float4. */
int pmh_mv_init(pmh_ctx* ctx, pmh_mv_scene* scene, int patchsize, float dmin, float dmax,
                float* out);
/* pmmv3_patch_eval, pmmvstereo.cl:97-167, for a supplied particle field:
 * cost[y*W+x] = window cost of particles[y*W+x] at (x, y). */
int pmh_mv_cost(pmh_ctx* ctx, pmh_mv_scene* scene, int patchsize, const float* particles,
                float* cost);
/* pmmv3_iteration, pmmvstereo.cl:197-304: the 8 neighbours, 3 random
 * neighbours at offsets in [0, 7]^2, one random particle and `refine`
 * (0 .. 8; the draft: 0) randomized refinements; strictly lower cost wins.
 * draws: NULL (RNG_init(y*W + x + iter, 33)) or 33 floats per pixel.  sweep:
 * PMH_SWEEP_ALL (the draft), RED or BLACK (a black pass reads the red pixels
 * of out).  old and out must be distinct fields. */
int pmh_mv_propagate(pmh_ctx* ctx, pmh_mv_scene* scene, const float* old, float* out,
                     int patchsize, float dmin, float dmax, int iter, const float* draws,
                     int sweep, int refine);
/* planef_homography, Plane.cl:60-86: H (9 floats, row-major) of each of
 * `count` particles for every source view, H[(i*nsrc + v)*9 + ...]. */
int pmh_mv_homography(pmh_ctx* ctx, pmh_mv_scene* scene, const float* particles, int count,
                      float* H);
/* depth[y*W+x] of particles[y*W+x] at (x, y). */
int pmh_mv_depth(pmh_ctx* ctx, pmh_mv_scene* scene, const float* particles, float* depth);
/* init + `iterations` passes (sweep PMH_SWEEP_ALL: one Jacobi pass each, the
 * draft; PMH_SWEEP_REDBLACK: a red and a black pass) + depth.  particles: NULL
 * or width*height float4 for the final field.  Returns after the stream has
 * drained. */
int pmh_mv_depthmap(pmh_ctx* ctx, pmh_mv_scene* scene, int patchsize, float dmin, float dmax,
                    int iterations, int sweep, int refine, float* depth, float* particles);

/* ---- on-disk formats (host memory; no context needed) --------------------
 * pmh_pfm_save: PFMSaver::save (plugins/isaver/PFM/PFMSaver.cpp:117-160) --
 * "Pf" for gray, "PF" for RGBA with alpha dropped, rows bottom to top,
 * little-endian scale -1.0.  pmh_cvtraw_*: CVTRawSaver / CVTRawLoader
 * (plugins/isaver/CVTRaw/CVTRawSaver.cpp, plugins/iloader/CVTRaw/
 * CVTRawLoader.cpp) -- a 16-byte header (uint32 width, height, row stride in
 * bytes, IFormat id of cvt/gfx/IFormat.h:36-63; GRAY_FLOAT = 4, RGBA_FLOAT =
 * 12) then height x stride bytes.  pmh_cvtraw_load copies rows to the
 * caller's stride.  Failures: PMH_ERR_INVALID and pmh_io_last_error(). */
int pmh_pfm_save(const char* path, const float* img, int width, int height, int format);
int pmh_cvtraw_save(const char* path, const void* img, int width, int height, size_t stride,
                    int format_id);
int pmh_cvtraw_info(const char* path, int* width, int* height, size_t* stride, int* format_id);
int pmh_cvtraw_load(const char* path, void* img, size_t stride);
const char* pmh_io_last_error(void);

/* ---- measurement hook (no counterpart in the reference) ------------------
 * When enabled, the context brackets every eager kernel launch with HIP
 * events on its own stream and accumulates the durations per kernel name
 * ("propagate", "pd_step", "consistency", "fill_state", "viewbuf", "init",
 * "prep", "zero", "grad_refine", "fill_depthmap", "fill_normalmap",
 * "mv_init", "mv_propagate"); a
 * replayed graph is timed as a whole ("graph").  Used by the benchmark to
 * time kernels live. */
int pmh_profile_enable(pmh_ctx* ctx, int enable);
int pmh_profile_reset(pmh_ctx* ctx);
int pmh_profile_read(pmh_ctx* ctx, const char* kernel, double* total_ms, long long* launches);
/* Brackets only the launches of one kernel name (NULL or "": all), so that the
 * events of a timed run are few. */
int pmh_profile_only(pmh_ctx* ctx, const char* kernel);
/* Window-cost candidates the algorithm evaluates in the propagate launches
 * run while profiling was enabled (since the last reset): per pixel and pass
 * 16 (10 without the random neighbourhood) + the pixel's view-propagation
 * candidates (at most 4) + the randomized refinements.  The algorithmic work
 * of those launches is this count x (2p+1)^2 window taps. */
int pmh_profile_candidates(pmh_ctx* ctx, long long* candidates);

#ifdef __cplusplus
}
#endif
#endif
