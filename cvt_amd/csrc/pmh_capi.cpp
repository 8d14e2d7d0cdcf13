// pmh_capi.cpp -- C-ABI of the gfx950 PatchMatch stereo path (include/pmh_capi.h).
//
// The context holds the stream it enqueues on, an Engine (scratch pool, param
// arena, graph cache, tuned launch shapes) and the device copy of the
// support-factor table.  Every entry point builds a launch plan and runs it;
// pmh_depthmap is pmh_batch_depthmap with one pair (pmh_pipeline.cpp).
#include "pmh_capi.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "pmh_internal.hpp"
#include "pmh_mv.hpp"
#include "pmh_pipeline.hpp"

using namespace pmh;

static thread_local std::string g_last_error;

struct pmh_ctx {
    Engine eng;
    int device = 0;
    std::map<int, float*> ftab;  // support-factor tables by half window

    explicit pmh_ctx(hipStream_t s) : eng(s) {}
    const float* factors(int P) {
        auto it = ftab.find(P);
        if (it != ftab.end()) return it->second;
        const int S = 2 * P + 1;
        std::vector<float> h((size_t)kFactorTables * S * S);
        factor_table(P, h.data());
        float* d = eng.get<float>("factor_table." + std::to_string(P), h.size());
        PMH_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                               eng.stream));
        PMH_HIP(hipStreamSynchronize(eng.stream));
        ftab[P] = d;
        return d;
    }
    FactorFn factor_fn() {
        return [this](int P) { return factors(P); };
    }
    std::map<int, float*> mvtab;  // multi-view weight exponents by half window
    const float* mv_factors(int P) {
        auto it = mvtab.find(P);
        if (it != mvtab.end()) return it->second;
        const int S = 2 * P + 1;
        std::vector<float> h((size_t)S * S);
        mv_factor_table(P, h.data());
        float* d = eng.get<float>("mv_factor_table." + std::to_string(P), h.size());
        PMH_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                               eng.stream));
        PMH_HIP(hipStreamSynchronize(eng.stream));
        mvtab[P] = d;
        return d;
    }
};

// A prepared multi-view scene (pmh_mv_scene_create): one device allocation
// holding the reference image | its gradient image | the source texel records.
struct pmh_mv_scene {
    pmh::MVImages im{};
    pmh::MVConst k{};
    int nsrc = 0;
    int device = 0;
    void* mem = nullptr;
};

namespace {

template <class F>
int guarded(F&& f) {
    try {
        f();
        return PMH_OK;
    } catch (const Unsupported& e) {
        g_last_error = e.what();
        return PMH_ERR_UNSUPPORTED;
    } catch (const ArgError& e) {
        g_last_error = e.what();
        return PMH_ERR_INVALID;
    } catch (const HipError& e) {
        g_last_error = e.what();
        return e.code == (int)hipErrorOutOfMemory || e.code == (int)hipErrorMemoryAllocation
                   ? PMH_ERR_NOMEM
                   : PMH_ERR_HIP;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PMH_ERR_HIP;
    } catch (...) {
        g_last_error = "unknown error";
        return PMH_ERR_HIP;
    }
}

void check_ctx(pmh_ctx* c) {
    if (!c) throw ArgError("null pmh_ctx");
}
void check_ptr(const void* p, const char* what) {
    if (!p) throw ArgError(std::string("null pointer: ") + what);
}
void check_size(int W, int H) {
    if (W <= 0 || H <= 0) throw ArgError("image size must be positive");
    if (W > kMaxRowWidth)
        throw ArgError("image width " + std::to_string(W) + " exceeds the supported " +
                       std::to_string(kMaxRowWidth));
    if ((long long)W * H > (1ll << 28)) throw ArgError("image too large");
}
void check_patch(int P) {
    if (P < 0 || P > 64) throw ArgError("patchsize (half window) must be in [0, 64]");
}

// Adds the packing of an other-view image (I2 with its supplied gradient G2)
// into the gather planes of the window cost.
Gather pack(pmh_ctx* c, Plan& plan, const std::string& role, const float* I, const float4* G,
            int W, int H, int nc) {
    const size_t n = (size_t)W * H;
    if (n * rec_bytes(nc) > 0x7fffffffull)
        throw ArgError("image too large for 32-bit gather offsets");
    float4* rec = c->eng.get<float4>(role + ".rec", (n * rec_bytes(nc) + 15) / 16);
    Batch<PrepParams> b;
    b.push(PrepParams{I, G, nullptr, nullptr, rec, W, H, nc}, tiles_2d(W, H));
    add_batch(plan, Kern::Prep, b, "prep");
    return Gather{rec, (unsigned)n};
}

// A launch of one item.
template <class P>
Launch* one(Plan& plan, Kern k, const P& p, int tiles, const char* name) {
    Batch<P> b;
    b.push(p, tiles);
    return add_batch(plan, k, b, name);
}

// floats per pixel of an image format
int channels(int format) {
    if (format == PMH_FORMAT_GRAY) return kGray;
    if (format == PMH_FORMAT_RGBA) return kRGBA;
    throw ArgError("format must be PMH_FORMAT_GRAY or PMH_FORMAT_RGBA");
}

void check_options(const pmh_options& o) {
    if (o.sweep != PMH_SWEEP_ALL && o.sweep != PMH_SWEEP_REDBLACK)
        throw ArgError("options.sweep must be PMH_SWEEP_ALL or PMH_SWEEP_REDBLACK");
    if (o.pd_iters < 1) throw ArgError("options.pd_iters must be >= 1");
    if (o.random_refine < 0 || o.random_refine > 2)
        throw ArgError("options.random_refine must be in [0, 2]");
    if (o.grad_steps < 0 || o.grad_steps > 16)
        throw ArgError("options.grad_steps must be in [0, 16]");
    if (o.grad_steps > 0 && !(o.grad_step > 0.0f))
        throw ArgError("options.grad_step must be positive");
    if (o.coarse_levels < 0 || o.coarse_levels > 4)
        throw ArgError("options.coarse_levels must be in [0, 4]");
    if (o.coarse_levels > 0 && (o.coarse_iters < 1 || o.coarse_iters > 64))
        throw ArgError("options.coarse_iters must be in [1, 64]");
    if (o.coarse_patch < 0 || o.coarse_patch > 64)
        throw ArgError("options.coarse_patch must be in [0, 64]");
    if (o.coarse_pd_iters < 0) throw ArgError("options.coarse_pd_iters must be >= 0");
    if (o.coarse_random_refine < -1 || o.coarse_random_refine > 2)
        throw ArgError("options.coarse_random_refine must be in [-1, 2]");
    if (o.random_neighbours != 0 && o.random_neighbours != 6)
        throw ArgError("options.random_neighbours must be 0 or 6");
    if (o.coarse_random_neighbours != -1 && o.coarse_random_neighbours != 0 &&
        o.coarse_random_neighbours != 6)
        throw ArgError("options.coarse_random_neighbours must be -1, 0 or 6");
    if (o.coarse_joint_views != 0 && o.coarse_joint_views != 1)
        throw ArgError("options.coarse_joint_views must be 0 or 1");
}

}  // namespace

extern "C" {

const char* pmh_last_error(void) { return g_last_error.c_str(); }

void pmh_default_options(pmh_options* o) {
    if (!o) return;
    o->maxdispdiff = 0.5f;    // PMHuberStereo.cpp:70
    o->maxanglediff = 10.0f;  // PMHuberStereo.cpp:71
    o->pd_iters = 250;        // PMHuberStereo.cpp:243,261
    o->sweep = PMH_SWEEP_ALL;
    o->grad_steps = 0;
    o->grad_step = 0.5f;
    o->random_refine = 2;     // pmhstereo.cl:383 (NUMRNDTRIES - 1)
    o->use_graph = 0;
    o->coarse_levels = 0;
    o->coarse_iters = 4;
    o->coarse_patch = 0;
    o->coarse_pd_iters = 0;
    o->coarse_random_refine = -1;
    o->random_neighbours = 6;  // pmhstereo.cl:355-361
    o->coarse_random_neighbours = -1;
    o->coarse_joint_views = 0;
}

int pmh_create(pmh_ctx** out, void* stream) {
    return guarded([&] {
        check_ptr(out, "ctx");
        *out = nullptr;
        int dev = 0;
        PMH_HIP(hipGetDevice(&dev));
        noalloc_scratch_init(dev);  // the device's pmh_pdrof_noalloc scratch address
        auto* c = new pmh_ctx(static_cast<hipStream_t>(stream));  // NULL: the default stream
        c->device = dev;
        *out = c;
    });
}

int pmh_destroy(pmh_ctx* ctx) {
    return guarded([&] { delete ctx; });
}

int pmh_synchronize(pmh_ctx* ctx) {
    return guarded([&] {
        check_ctx(ctx);
        PMH_HIP(hipStreamSynchronize(ctx->eng.stream));
    });
}

int pmh_gradxy(pmh_ctx* ctx, const float* img, int W, int H, float* grad, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(img, "img"); check_ptr(grad, "grad"); check_size(W, H);
        const int nc = channels(format);
        Plan plan;
        one(plan, Kern::Prep,
            PrepParams{img, nullptr, reinterpret_cast<float4*>(grad), nullptr, nullptr, W, H, nc},
            tiles_2d(W, H), "prep");
        ctx->eng.run(plan, false);
    });
}

int pmh_weight(pmh_ctx* ctx, const float* img, int W, int H, float* weight, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(img, "img"); check_ptr(weight, "weight"); check_size(W, H);
        const int nc = channels(format);
        Plan plan;
        one(plan, Kern::Prep, PrepParams{img, nullptr, nullptr, weight, nullptr, W, H, nc},
            tiles_2d(W, H), "prep");
        ctx->eng.run(plan, false);
    });
}

int pmh_rng_draws(pmh_ctx* ctx, int W, int H, int iter, int stride, int count, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(out, "out"); check_size(W, H);
        if (count < 0 || count > 64) throw ArgError("count must be in [0, 64]");
        if (stride != (int)kPropStride && stride != (int)kInitStride &&
            stride != (int)kMVStride)
            throw ArgError("stride must be 45 (propagate), 3 (init) or 33 (multi-view)");
        if (iter < 0) throw ArgError("iter must be >= 0");
        launch_rng_draws(W, H, iter, (uint32_t)stride, count, out, ctx->eng.stream);
    });
}

int pmh_init(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2, const float* G2,
             int W, int H, int P, float depthmax, int lr, float* out, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(I1, "I1"); check_ptr(G1, "G1"); check_ptr(I2, "I2");
        check_ptr(G2, "G2"); check_ptr(out, "out"); check_size(W, H); check_patch(P);
        const int nc = channels(format);
        const float* F = ctx->factors(P);
        Plan plan;
        Gather g = pack(ctx, plan, "other", I2, reinterpret_cast<const float4*>(G2), W, H, nc);
        Launch* l = one(plan, Kern::Init,
                        InitParams{I1, reinterpret_cast<const float4*>(G1), g, nullptr,
                                   reinterpret_cast<float4*>(out), W, H, lr ? 1 : 0, depthmax},
                        tiles_waves(W, H, 0), "init");
        l->F = F;
        l->P = P;
        l->nc = nc;
        ctx->eng.run(plan, false);
    });
}

int pmh_cost(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2, const float* G2,
             int W, int H, int P, const float* states, float* cost, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(I1, "I1"); check_ptr(G1, "G1"); check_ptr(I2, "I2");
        check_ptr(G2, "G2"); check_ptr(states, "states"); check_ptr(cost, "cost");
        check_size(W, H); check_patch(P);
        const int nc = channels(format);
        const float* F = ctx->factors(P);
        Plan plan;
        Gather g = pack(ctx, plan, "other", I2, reinterpret_cast<const float4*>(G2), W, H, nc);
        ctx->eng.run(plan, false);
        launch_cost(I1, reinterpret_cast<const float4*>(G1), g, F, W, H, P,
                    reinterpret_cast<const float4*>(states), cost, nc, ctx->eng.stream);
    });
}

int pmh_propagate(pmh_ctx* ctx, const float* old, float* out, const float* I1, const float* G1,
                  const float* I2, const float* G2, const float* smooth, float theta, int W,
                  int H, int P, float depthmax, int lr, int iter, const int* viewin,
                  const float* viewsrc, const float* draws, int sweep, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(old, "old"); check_ptr(out, "out"); check_ptr(I1, "I1");
        check_ptr(G1, "G1"); check_ptr(I2, "I2"); check_ptr(G2, "G2");
        check_ptr(smooth, "smooth"); check_size(W, H); check_patch(P);
        if (viewin && !viewsrc) throw ArgError("viewin given without viewsrc");
        if (sweep < PMH_SWEEP_ALL || sweep > PMH_SWEEP_BLACK)
            throw ArgError("sweep must be PMH_SWEEP_ALL, RED or BLACK");
        if (old == out) throw ArgError("propagate needs distinct old and out fields");
        if (iter < 0) throw ArgError("iter must be >= 0");
        const int nc = channels(format);
        Plan plan;
        PropParams p{};
        p.old = reinterpret_cast<const float4*>(old);
        p.out = reinterpret_cast<float4*>(out);
        p.I1 = I1;
        p.G1 = reinterpret_cast<const float4*>(G1);
        p.g2 = pack(ctx, plan, "other", I2, reinterpret_cast<const float4*>(G2), W, H, nc);
        p.smooth = reinterpret_cast<const float4*>(smooth);
        p.viewin = reinterpret_cast<const ViewRec*>(viewin);
        p.viewsrc = reinterpret_cast<const float4*>(viewsrc);
        p.draws = draws;
        p.F = ctx->factors(P);
        p.W = W; p.H = H; p.P = P;
        p.theta = theta; p.depthmax = depthmax;
        p.lr = lr ? 1 : 0; p.iter = iter; p.sweep = sweep;
        p.random_refine = 2;
        Launch* l = one(plan, Kern::Propagate, p, tiles_prop(W, H, sweep), "propagate");
        l->nc = nc;
        l->chunk = 0;  // persistent blocks by default, as depthMap's
        l->tunable = true;  // tuned like depthMap's propagate launches (same keys)
        l->tune_key = propagate_tune_key(p, nc, l->total);
        ctx->eng.run(plan, false);
    });
}

int pmh_viewbuf_build(pmh_ctx* ctx, const float* states, int W, int H, int* viewbuf) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(states, "states"); check_ptr(viewbuf, "viewbuf");
        check_size(W, H);
        Plan plan;
        one(plan, Kern::Viewbuf,
            ViewParams{reinterpret_cast<const float4*>(states), reinterpret_cast<ViewRec*>(viewbuf),
                       W, H},
            tiles_rows(H), "viewbuf");
        ctx->eng.run(plan, false);
    });
}

int pmh_consistency(pmh_ctx* ctx, const float* left, const float* right, int W, int H,
                    float maxdispdiff, float maxanglediff, int lr, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(left, "left"); check_ptr(right, "right"); check_ptr(out, "out");
        check_size(W, H);
        Plan plan;
        Launch* l = one(plan, Kern::Consistency,
                        ConsParams{reinterpret_cast<const float4*>(left),
                                   reinterpret_cast<const float4*>(right),
                                   reinterpret_cast<float4*>(out), W, H, lr ? 1 : 0},
                        tiles_2d(W, H), "consistency");
        l->f0 = maxdispdiff;
        l->f1 = maxanglediff;
        ctx->eng.run(plan, false);
    });
}

int pmh_fill_state(pmh_ctx* ctx, const float* in, int W, int H, float depthmax, int lr,
                   float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(out, "out"); check_size(W, H);
        if (in == out) throw ArgError("fill_state cannot run in place");
        Plan plan;
        one(plan, Kern::FillState,
            FillParams{reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out), W, H,
                       lr ? 1 : 0, depthmax},
            tiles_rows(H), "fill_state");
        ctx->eng.run(plan, false);
    });
}

int pmh_pd_step(pmh_ctx* ctx, const float* last, const float* plast, const float* f,
                const float* w, int W, int H, float lambda, float* out, float* pout) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(last, "last"); check_ptr(plast, "plast"); check_ptr(f, "f");
        check_ptr(w, "w"); check_ptr(out, "out"); check_ptr(pout, "pout"); check_size(W, H);
        Plan plan;
        Launch* l = one(plan, Kern::PdStep,
                        PdParams{reinterpret_cast<const float4*>(last),
                                 reinterpret_cast<const float4*>(plast),
                                 reinterpret_cast<const float4*>(f), w,
                                 reinterpret_cast<float4*>(out), reinterpret_cast<float4*>(pout),
                                 W, H},
                        tiles_2d(W, H), "pd_step");
        l->f0 = lambda;
        ctx->eng.run(plan, false);
    });
}

int pmh_pdrof(pmh_ctx* ctx, float* u, const float* f, const float* w, int W, int H, float lambda,
              int iters) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(u, "u"); check_ptr(f, "f"); check_ptr(w, "w");
        check_size(W, H);
        if (iters < 1) throw ArgError("iters must be >= 1");
        const size_t n = (size_t)W * H;
        PdItem it{reinterpret_cast<float4*>(u), reinterpret_cast<const float4*>(f), w,
                  ctx->eng.get<float4>("pdrof.tmp", n), ctx->eng.get<float4>("pdrof.p0", 2 * n),
                  ctx->eng.get<float4>("pdrof.p1", 2 * n), W, H};
        Plan plan;
        add_pdrof(plan, {it}, lambda, iters, false);
        ctx->eng.run(plan, false);
    });
}

int pmh_pdrof_noalloc(pmh_ctx* ctx, float* u, const float* f, const float* w, int W, int H,
                      float lambda, int iters) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(u, "u"); check_ptr(f, "f"); check_ptr(w, "w");
        check_size(W, H);
        if (iters < 1) throw ArgError("iters must be >= 1");
        int dev = 0;
        PMH_HIP(hipGetDevice(&dev));
        if (dev != ctx->device)
            throw ArgError("pmh_pdrof_noalloc: the context's device is not the current device");
        launch_pdrof_noalloc(dev, reinterpret_cast<float4*>(u), reinterpret_cast<const float4*>(f),
                             w, W, H, lambda, iters, ctx->eng.stream);
    });
}

int pmh_fill_depthmap(pmh_ctx* ctx, const float* in, int W, int H, float scale, float* dmap) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(dmap, "dmap"); check_size(W, H);
        Plan plan;
        one(plan, Kern::FillDepth,
            FillDepthParams{reinterpret_cast<const float4*>(in), dmap, nullptr, W, H, scale},
            tiles_rows(H), "fill_depthmap");
        ctx->eng.run(plan, false);
    });
}

int pmh_fill_normalmap(pmh_ctx* ctx, const float* in, int W, int H, float* nmap) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(nmap, "nmap"); check_size(W, H);
        Plan plan;
        Launch* l = one(plan, Kern::FillDepth,
                        FillDepthParams{reinterpret_cast<const float4*>(in), nullptr,
                                        reinterpret_cast<float4*>(nmap), W, H, 1.0f},
                        tiles_rows(H), "fill_normalmap");
        l->i0 = 1;
        ctx->eng.run(plan, false);
    });
}

int pmh_depthmap(pmh_ctx* ctx, const float* left, const float* right, int W, int H, int P,
                 float depthmax, int iterations, int viewsamples, float dscale, float* dmap,
                 float* normalmap, const pmh_options* opts, float* left_states,
                 float* right_states) {
    (void)viewsamples;  // unused, as in the reference (pmhstereo.cl:40)
    return guarded([&] {
        check_ctx(ctx); check_ptr(left, "left"); check_ptr(right, "right");
        check_ptr(dmap, "dmap"); check_size(W, H); check_patch(P);
        if (iterations < 0) throw ArgError("iterations must be >= 0");
        if (!(depthmax > 0.0f)) throw ArgError("depthmax must be positive");
        pmh_options o;
        pmh_default_options(&o);
        if (opts) o = *opts;
        check_options(o);
        if (dscale <= 0.0f) dscale = 1.0f / depthmax;  // PMHuberStereo.cpp:75-76
        const PairIO io{left, right, W, H, kGray, depthmax, dscale, dmap,
                        reinterpret_cast<float4*>(normalmap),
                        reinterpret_cast<float4*>(left_states),
                        reinterpret_cast<float4*>(right_states)};
        depthmap_batch(ctx->eng, &io, 1, P, iterations, o, ctx->factor_fn());
    });
}

int pmh_batch_depthmap(pmh_ctx* ctx, const pmh_pair* pairs, int npairs, int P, int iterations,
                       const pmh_options* opts) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(pairs, "pairs"); check_patch(P);
        if (npairs < 1 || npairs > 4096) throw ArgError("npairs must be in [1, 4096]");
        if (iterations < 0) throw ArgError("iterations must be >= 0");
        pmh_options o;
        pmh_default_options(&o);
        if (opts) o = *opts;
        check_options(o);
        std::vector<PairIO> io(npairs);
        const int nc = channels(pairs[0].format);
        for (int i = 0; i < npairs; ++i) {
            const pmh_pair& q = pairs[i];
            const std::string at = " (pair " + std::to_string(i) + ")";
            if (!q.left || !q.right || !q.dmap)
                throw ArgError("null left, right or dmap" + at);
            check_size(q.width, q.height);
            if (!(q.depthmax > 0.0f)) throw ArgError("depthmax must be positive" + at);
            if (channels(q.format) != nc)
                throw ArgError("all pairs of a batch must have the same format" + at);
            if ((size_t)q.width * q.height * rec_bytes(nc) > 0x7fffffffull)
                throw ArgError("image too large for 32-bit gather offsets" + at);
            io[i] = PairIO{q.left, q.right, q.width, q.height, nc, q.depthmax,
                           q.dscale > 0.0f ? q.dscale : 1.0f / q.depthmax, q.dmap,
                           reinterpret_cast<float4*>(q.normalmap),
                           reinterpret_cast<float4*>(q.left_states),
                           reinterpret_cast<float4*>(q.right_states)};
        }
        depthmap_batch(ctx->eng, io.data(), npairs, P, iterations, o, ctx->factor_fn());
    });
}

void pmh_default_inpaint_options(pmh_options* o) {
    if (!o) return;
    pmh_default_options(o);
    o->maxdispdiff = 0.5f;    // occlusion map, PMHuberStereo.cpp:454
    o->maxanglediff = 5.0f;   // pmhstereo.cl:710
    o->pd_iters = 100;        // PMHuberStereo.cpp:466,484
}

namespace {

void inpaint_batch(pmh_ctx* ctx, const pmh_pair* pairs, int npairs, int P, int iterations,
                   const pmh_options* opts) {
    check_ctx(ctx); check_ptr(pairs, "pairs"); check_patch(P);
    if (npairs < 1 || npairs > 4096) throw ArgError("npairs must be in [1, 4096]");
    // the reference's final check reads plane fields written by iteration 0
    if (iterations < 1) throw ArgError("iterations must be >= 1 for depthMapInpaint");
    pmh_options o;
    pmh_default_inpaint_options(&o);
    if (opts) o = *opts;
    check_options(o);
    std::vector<PairIO> io(npairs);
    const int nc = channels(pairs[0].format);
    for (int i = 0; i < npairs; ++i) {
        const pmh_pair& q = pairs[i];
        const std::string at = " (pair " + std::to_string(i) + ")";
        if (!q.left || !q.right || !q.dmap) throw ArgError("null left, right or dmap" + at);
        check_size(q.width, q.height);
        if (!(q.depthmax > 0.0f)) throw ArgError("depthmax must be positive" + at);
        if (channels(q.format) != nc)
            throw ArgError("all pairs of a batch must have the same format" + at);
        io[i] = PairIO{q.left, q.right, q.width, q.height, nc, q.depthmax, 1.0f, q.dmap,
                       reinterpret_cast<float4*>(q.normalmap),
                       reinterpret_cast<float4*>(q.left_states),
                       reinterpret_cast<float4*>(q.right_states)};
    }
    depthmap_batch(ctx->eng, io.data(), npairs, P, iterations, o, ctx->factor_fn(), true);
}

}  // namespace

int pmh_depthmap_inpaint(pmh_ctx* ctx, const float* left, const float* right, int W, int H,
                         int P, float depthmax, int iterations, int viewsamples, float* dmap,
                         const pmh_options* opts, float* left_states, float* right_states) {
    (void)viewsamples;  // unused, as in the reference
    return guarded([&] {
        pmh_pair q{};
        q.left = left;
        q.right = right;
        q.width = W;
        q.height = H;
        q.format = PMH_FORMAT_GRAY;
        q.depthmax = depthmax;
        q.dmap = dmap;
        q.left_states = left_states;
        q.right_states = right_states;
        inpaint_batch(ctx, &q, 1, P, iterations, opts);
    });
}

int pmh_batch_depthmap_inpaint(pmh_ctx* ctx, const pmh_pair* pairs, int npairs, int P,
                               int iterations, const pmh_options* opts) {
    return guarded([&] { inpaint_batch(ctx, pairs, npairs, P, iterations, opts); });
}

int pmh_occmap(pmh_ctx* ctx, const float* left, const float* right, int W, int H,
               float maxdispdiff, float maxanglediff, int lr, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(left, "left"); check_ptr(right, "right"); check_ptr(out, "out");
        check_size(W, H);
        Plan plan;
        Launch* l = one(plan, Kern::Occmap,
                        ConsParams{reinterpret_cast<const float4*>(left),
                                   reinterpret_cast<const float4*>(right),
                                   reinterpret_cast<float4*>(out), W, H, lr ? 1 : 0},
                        tiles_2d(W, H), "occmap");
        l->f0 = maxdispdiff;
        l->f1 = maxanglediff;
        ctx->eng.run(plan, false);
    });
}

int pmh_normal_depth(pmh_ctx* ctx, const float* in, int W, int H, float depthmax, int lr,
                     float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(out, "out"); check_size(W, H);
        Plan plan;
        one(plan, Kern::NormalDepth,
            FillParams{reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out), W, H,
                       lr ? 1 : 0, depthmax},
            tiles_2d(W, H), "normal_depth");
        ctx->eng.run(plan, false);
    });
}

int pmh_pd_step_inpaint(pmh_ctx* ctx, const float* last, const float* plast, const float* f,
                        const float* w, const float* mask, int W, int H, float lambda,
                        float* out, float* pout) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(last, "last"); check_ptr(plast, "plast"); check_ptr(f, "f");
        check_ptr(w, "w"); check_ptr(mask, "mask"); check_ptr(out, "out");
        check_ptr(pout, "pout"); check_size(W, H);
        Plan plan;
        Launch* l = one(plan, Kern::PdStep,
                        PdParams{reinterpret_cast<const float4*>(last),
                                 reinterpret_cast<const float4*>(plast),
                                 reinterpret_cast<const float4*>(f), w,
                                 reinterpret_cast<float4*>(out), reinterpret_cast<float4*>(pout),
                                 W, H, reinterpret_cast<const float4*>(mask)},
                        tiles_2d(W, H), "pd_step_inpaint");
        l->f0 = lambda;
        l->i0 = 1;
        ctx->eng.run(plan, false);
    });
}

int pmh_pdrof_inpaint(pmh_ctx* ctx, float* u, const float* f, const float* w, const float* mask,
                      int W, int H, float lambda, int iters) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(u, "u"); check_ptr(f, "f"); check_ptr(w, "w");
        check_ptr(mask, "mask"); check_size(W, H);
        if (iters < 1) throw ArgError("iters must be >= 1");
        const size_t n = (size_t)W * H;
        PdItem it{reinterpret_cast<float4*>(u), reinterpret_cast<const float4*>(f), w,
                  ctx->eng.get<float4>("pdrof.tmp", n), ctx->eng.get<float4>("pdrof.p0", 2 * n),
                  ctx->eng.get<float4>("pdrof.p1", 2 * n), W, H,
                  reinterpret_cast<const float4*>(mask)};
        Plan plan;
        add_pdrof(plan, {it}, lambda, iters, false, true);
        ctx->eng.run(plan, false);
    });
}

int pmh_autotune_enable(pmh_ctx* ctx, int enable) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->eng.autotune = enable != 0;
    });
}

int pmh_autotune_sample(pmh_ctx* ctx, int tiles) {
    return guarded([&] {
        check_ctx(ctx);
        if (tiles != 0 && (tiles < 8 || tiles > (1 << 30)))
            throw ArgError("autotune sample must be 0 or in [8, 2^30] tiles");
        ctx->eng.sample_tiles = tiles;
    });
}

int pmh_tuning_load(pmh_ctx* ctx, const char* path) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(path, "path");
        ctx->eng.load_tuning(path);
    });
}

int pmh_tuning_save(pmh_ctx* ctx, const char* path) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(path, "path");
        ctx->eng.save_tuning(path);
    });
}

int pmh_graph_stats(pmh_ctx* ctx, int* graphs_cached, long long* graph_launches) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(graphs_cached, "graphs_cached");
        check_ptr(graph_launches, "graph_launches");
        *graphs_cached = (int)ctx->eng.graphs_cached();
        *graph_launches = ctx->eng.graph_launches;
    });
}

int pmh_pyr_down(pmh_ctx* ctx, const float* img, int W, int H, float* out, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(img, "img"); check_ptr(out, "out"); check_size(W, H);
        const int nc = channels(format);
        Plan plan;
        one(plan, Kern::PyrDown, PyrParams{img, out, W, H, nc},
            tiles_2d((W + 1) / 2, (H + 1) / 2), "pyr_down");
        ctx->eng.run(plan, false);
    });
}

int pmh_plane_up(pmh_ctx* ctx, const float* in, int w, int h, int W, int H, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(out, "out"); check_size(W, H);
        check_size(w, h);
        if (w != (W + 1) / 2 || h != (H + 1) / 2)
            throw ArgError("plane_up: input must be ((W+1)/2, (H+1)/2)");
        Plan plan;
        one(plan, Kern::PlaneUp,
            UpParams{reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out), W, H,
                     w, h},
            tiles_2d(W, H), "plane_up");
        ctx->eng.run(plan, false);
    });
}

int pmh_cost_backward(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2,
                      const float* G2, int W, int H, int P, const float* states,
                      const float* gout, float* dstates, float* dI1, float* dG1, float* dI2,
                      float* dG2, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(I1, "I1"); check_ptr(G1, "G1"); check_ptr(I2, "I2");
        check_ptr(G2, "G2"); check_ptr(states, "states"); check_ptr(gout, "gout");
        check_ptr(dstates, "dstates"); check_size(W, H); check_patch(P);
        const int nc = channels(format);
        const float* F = ctx->factors(P);
        Plan plan;
        Gather g = pack(ctx, plan, "other", I2, reinterpret_cast<const float4*>(G2), W, H, nc);
        ctx->eng.run(plan, false);
        launch_cost_backward(I1, reinterpret_cast<const float4*>(G1), g, F, W, H, P,
                             reinterpret_cast<const float4*>(states), gout,
                             reinterpret_cast<float4*>(dstates), dI1,
                             reinterpret_cast<float4*>(dG1), dI2,
                             reinterpret_cast<float4*>(dG2), nc, ctx->eng.stream);
    });
}

int pmh_gradxy_backward(pmh_ctx* ctx, const float* dG, int W, int H, float* dI) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(dG, "dG"); check_ptr(dI, "dI"); check_size(W, H);
        launch_gradxy_backward(reinterpret_cast<const float4*>(dG), W, H, dI, ctx->eng.stream);
    });
}

int pmh_grad_refine(pmh_ctx* ctx, float* states, const float* I1, const float* G1,
                    const float* I2, const float* G2, const float* smooth, float theta, int W,
                    int H, int P, float depthmax, int lr, int steps, float radius, int format) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(states, "states"); check_ptr(I1, "I1"); check_ptr(G1, "G1");
        check_ptr(I2, "I2"); check_ptr(G2, "G2"); check_ptr(smooth, "smooth");
        check_size(W, H); check_patch(P);
        if (steps < 0 || steps > 16) throw ArgError("steps must be in [0, 16]");
        if (!(radius > 0.0f)) throw ArgError("radius must be positive");
        const int nc = channels(format);
        const float* F = ctx->factors(P);
        Plan plan;
        Gather g = pack(ctx, plan, "other", I2, reinterpret_cast<const float4*>(G2), W, H, nc);
        Launch* l = one(plan, Kern::GradRefine,
                        RefineParams{reinterpret_cast<float4*>(states), I1,
                                     reinterpret_cast<const float4*>(G1), g,
                                     reinterpret_cast<const float4*>(smooth), W, H, lr ? 1 : 0,
                                     depthmax},
                        tiles_waves(W, H, 0), "grad_refine");
        l->F = F;
        l->P = P;
        l->f0 = theta;
        l->i0 = steps;
        l->f1 = radius;
        l->nc = nc;
        ctx->eng.run(plan, false);
    });
}

int pmh_profile_enable(pmh_ctx* ctx, int enable) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->eng.prof = enable != 0;
    });
}

int pmh_profile_reset(pmh_ctx* ctx) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->eng.resolve();
        ctx->eng.totals.clear();
        PMH_HIP(hipMemsetAsync(ctx->eng.cand_counter(), 0, 8, ctx->eng.stream));
        PMH_HIP(hipStreamSynchronize(ctx->eng.stream));
    });
}

int pmh_profile_only(pmh_ctx* ctx, const char* kernel) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->eng.prof_only = kernel ? kernel : "";
    });
}

int pmh_profile_candidates(pmh_ctx* ctx, long long* candidates) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(candidates, "candidates");
        unsigned long long v = 0;
        PMH_HIP(hipMemcpyAsync(&v, ctx->eng.cand_counter(), 8, hipMemcpyDeviceToHost,
                               ctx->eng.stream));
        PMH_HIP(hipStreamSynchronize(ctx->eng.stream));
        *candidates = (long long)v;
    });
}

int pmh_profile_read(pmh_ctx* ctx, const char* kernel, double* total_ms, long long* launches) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(kernel, "kernel"); check_ptr(total_ms, "total_ms");
        check_ptr(launches, "launches");
        ctx->eng.resolve();
        auto it = ctx->eng.totals.find(kernel);
        *total_ms = it == ctx->eng.totals.end() ? 0.0 : it->second.first;
        *launches = it == ctx->eng.totals.end() ? 0 : it->second.second;
    });
}


// ---------------------------------------------------------------- multi-view --
namespace {

void check_scene(pmh_ctx* ctx, pmh_mv_scene* sc) {
    check_ctx(ctx);
    if (!sc) throw ArgError("null pmh_mv_scene");
    if (sc->device != ctx->device)
        throw ArgError("the scene was prepared on another device than the context's");
}

void check_range(float dmin, float dmax) {
    if (!(std::isfinite(dmin) && std::isfinite(dmax) && dmin >= 0.0f && dmin < dmax))
        throw ArgError("depth range must satisfy 0 <= dmin < dmax (finite)");
}

void check_refine(int refine) {
    if (refine < 0 || refine > kMVMaxRefine) throw ArgError("refine must be in [0, 8]");
}

void mv_propagate_pass(pmh_ctx* ctx, pmh_mv_scene* sc, const float* Fw, int P,
                       const float4* old, float4* out, float dmin, float dmax, int iter,
                       const float* draws, int sweep, int refine) {
    ctx->eng.timed_launch("mv_propagate", [&] {
        launch_mv_propagate(sc->im, sc->k, sc->nsrc, Fw, P, old, out, dmin, dmax, iter, draws,
                            sweep, refine, ctx->eng.stream);
    });
}

}  // namespace

int pmh_mv_scene_create(pmh_ctx* ctx, const pmh_mv_view* ref, const pmh_mv_view* srcs,
                        int nsrc, int W, int H, pmh_mv_scene** scene) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(ref, "ref"); check_ptr(srcs, "srcs");
        check_ptr(scene, "scene"); check_ptr(ref->image, "ref->image"); check_size(W, H);
        *scene = nullptr;
        if (nsrc < 1 || nsrc > kMVMaxViews) throw ArgError("nsrc must be in [1, 4]");
        for (int v = 0; v < nsrc; ++v) check_ptr(srcs[v].image, "srcs[v].image");
        const size_t n = (size_t)W * H;
        if (mv_texel_floats(W, H) > 0xffffffffull)
            throw ArgError("image too large for 32-bit texel offsets");
        std::vector<float> Ks(9 * nsrc), Ts(12 * nsrc);
        for (int v = 0; v < nsrc; ++v) {
            std::memcpy(&Ks[9 * v], srcs[v].K, sizeof(srcs[v].K));
            std::memcpy(&Ts[12 * v], srcs[v].T, sizeof(srcs[v].T));
        }
        auto sc = std::make_unique<pmh_mv_scene>();
        mv_constants(ref->K, ref->T, nsrc, Ks.data(), Ts.data(), sc->k);
        sc->nsrc = nsrc;
        sc->device = ctx->device;
        const size_t ir_bytes = (n * sizeof(float) + 255) & ~(size_t)255;
        const size_t gr_bytes = n * sizeof(float4);
        const size_t tstride = mv_texel_floats(W, H);
        // + 16 floats: the stage copies of the last texel row read up to 6 past it
        const size_t tex_bytes = (tstride * nsrc + 16) * sizeof(float);
        PMH_HIP(hipMalloc(&sc->mem, ir_bytes + gr_bytes + tex_bytes));
        auto* base = static_cast<unsigned char*>(sc->mem);
        float* Ir = reinterpret_cast<float*>(base);
        float4* Gr = reinterpret_cast<float4*>(base + ir_bytes);
        float* tex = reinterpret_cast<float*>(base + ir_bytes + gr_bytes);
        sc->im = MVImages{Ir, Gr, tex, W, H, (unsigned)tstride};
        hipStream_t st = ctx->eng.stream;
        PMH_HIP(hipMemcpyAsync(Ir, ref->image, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        {
            Plan plan;
            one(plan, Kern::Prep, PrepParams{ref->image, nullptr, Gr, nullptr, nullptr, W, H, 1},
                tiles_2d(W, H), "prep");
            ctx->eng.run(plan, false);
        }
        float4* g = ctx->eng.get<float4>("mv.source_grad", n);
        for (int v = 0; v < nsrc; ++v) {
            Plan plan;
            one(plan, Kern::Prep,
                PrepParams{srcs[v].image, nullptr, g, nullptr, nullptr, W, H, 1}, tiles_2d(W, H),
                "prep");
            ctx->eng.run(plan, false);
            launch_mv_texels(srcs[v].image, g, W, H, tex + (size_t)v * tstride, st);
        }
        PMH_HIP(hipStreamSynchronize(st));
        *scene = sc.release();
    });
}

int pmh_mv_scene_destroy(pmh_mv_scene* scene) {
    return guarded([&] {
        if (!scene) return;
        int cur = 0;
        PMH_HIP(hipGetDevice(&cur));
        if (cur != scene->device) PMH_HIP(hipSetDevice(scene->device));
        const hipError_t e = hipFree(scene->mem);
        if (cur != scene->device) (void)hipSetDevice(cur);
        delete scene;
        PMH_HIP(e);
    });
}

int pmh_mv_init(pmh_ctx* ctx, pmh_mv_scene* sc, int P, float dmin, float dmax, float* out) {
    return guarded([&] {
        check_scene(ctx, sc); check_ptr(out, "out"); check_patch(P); check_range(dmin, dmax);
        const float* Fw = ctx->mv_factors(P);
        ctx->eng.timed_launch("mv_init", [&] {
            launch_mv_init(sc->im, sc->k, sc->nsrc, Fw, P, dmin, dmax,
                           reinterpret_cast<float4*>(out), ctx->eng.stream);
        });
    });
}

int pmh_mv_cost(pmh_ctx* ctx, pmh_mv_scene* sc, int P, const float* particles, float* cost) {
    return guarded([&] {
        check_scene(ctx, sc); check_ptr(particles, "particles"); check_ptr(cost, "cost");
        check_patch(P);
        launch_mv_cost(sc->im, sc->k, sc->nsrc, ctx->mv_factors(P), P,
                       reinterpret_cast<const float4*>(particles), cost, ctx->eng.stream);
    });
}

int pmh_mv_propagate(pmh_ctx* ctx, pmh_mv_scene* sc, const float* old, float* out, int P,
                     float dmin, float dmax, int iter, const float* draws, int sweep,
                     int refine) {
    return guarded([&] {
        check_scene(ctx, sc); check_ptr(old, "old"); check_ptr(out, "out"); check_patch(P);
        check_range(dmin, dmax); check_refine(refine);
        if (sweep < PMH_SWEEP_ALL || sweep > PMH_SWEEP_BLACK)
            throw ArgError("sweep must be PMH_SWEEP_ALL, RED or BLACK");
        if (old == out) throw ArgError("propagate needs distinct old and out fields");
        if (iter < 0) throw ArgError("iter must be >= 0");
        mv_propagate_pass(ctx, sc, ctx->mv_factors(P), P, reinterpret_cast<const float4*>(old),
                          reinterpret_cast<float4*>(out), dmin, dmax, iter, draws, sweep,
                          refine);
    });
}

int pmh_mv_homography(pmh_ctx* ctx, pmh_mv_scene* sc, const float* particles, int count,
                      float* H) {
    return guarded([&] {
        check_scene(ctx, sc); check_ptr(particles, "particles"); check_ptr(H, "H");
        if (count < 0) throw ArgError("count must be >= 0");
        launch_mv_homography(reinterpret_cast<const float4*>(particles), count, sc->k, sc->nsrc,
                             H, ctx->eng.stream);
    });
}

int pmh_mv_depth(pmh_ctx* ctx, pmh_mv_scene* sc, const float* particles, float* depth) {
    return guarded([&] {
        check_scene(ctx, sc); check_ptr(particles, "particles"); check_ptr(depth, "depth");
        launch_mv_depth(reinterpret_cast<const float4*>(particles), sc->k, sc->im.W, sc->im.H,
                        depth, ctx->eng.stream);
    });
}

int pmh_mv_depthmap(pmh_ctx* ctx, pmh_mv_scene* sc, int P, float dmin, float dmax,
                    int iterations, int sweep, int refine, float* depth, float* particles) {
    return guarded([&] {
        check_scene(ctx, sc); check_ptr(depth, "depth"); check_patch(P);
        check_range(dmin, dmax); check_refine(refine);
        if (iterations < 0) throw ArgError("iterations must be >= 0");
        if (sweep != PMH_SWEEP_ALL && sweep != PMH_SWEEP_REDBLACK)
            throw ArgError("sweep must be PMH_SWEEP_ALL or PMH_SWEEP_REDBLACK");
        const size_t n = (size_t)sc->im.W * sc->im.H;
        float4* a = ctx->eng.get<float4>("mv.field_a", n);
        float4* b = ctx->eng.get<float4>("mv.field_b", n);
        const float* Fw = ctx->mv_factors(P);
        hipStream_t st = ctx->eng.stream;
        ctx->eng.timed_launch("mv_init", [&] {
            launch_mv_init(sc->im, sc->k, sc->nsrc, Fw, P, dmin, dmax, a, st);
        });
        for (int it = 0; it < iterations; ++it) {
            if (sweep == PMH_SWEEP_REDBLACK) {
                // the red pass writes every red pixel of b, the black pass every black one
                mv_propagate_pass(ctx, sc, Fw, P, a, b, dmin, dmax, it, nullptr, PMH_SWEEP_RED,
                                  refine);
                mv_propagate_pass(ctx, sc, Fw, P, a, b, dmin, dmax, it, nullptr,
                                  PMH_SWEEP_BLACK, refine);
            } else {
                mv_propagate_pass(ctx, sc, Fw, P, a, b, dmin, dmax, it, nullptr, PMH_SWEEP_ALL,
                                  refine);
            }
            std::swap(a, b);
        }
        launch_mv_depth(a, sc->k, sc->im.W, sc->im.H, depth, st);
        if (particles)
            PMH_HIP(hipMemcpyAsync(particles, a, n * sizeof(float4), hipMemcpyDeviceToDevice, st));
        PMH_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
