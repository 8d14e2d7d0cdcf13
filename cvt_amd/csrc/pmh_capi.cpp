// pmh_capi.cpp -- C-ABI of the gfx950 PatchMatch stereo path (include/pmh_capi.h).
//
// The context owns one HIP stream, a grow-only scratch pool keyed by role and
// the device copy of the support-factor table.  pmh_depthmap restates the host
// sequence of PMHuberStereo::depthMap (cvt/vision/PMHuberStereo.cpp:64-291)
// without the debug block (:153-185, which saves PNGs and waits on getchar).
#include "pmh_capi.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pmh_internal.hpp"

using namespace pmh;

static thread_local std::string g_last_error;

struct pmh_ctx {
    // measurement hook: events around launches, resolved on read
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    bool prof = false;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> spare;
    std::map<std::string, std::pair<double, long long>> totals;

    hipEvent_t event() {
        if (!spare.empty()) {
            hipEvent_t e = spare.back();
            spare.pop_back();
            return e;
        }
        hipEvent_t e;
        PMH_HIP(hipEventCreate(&e));
        return e;
    }
    void resolve() {
        if (pending.empty()) return;
        PMH_HIP(hipStreamSynchronize(stream));
        for (auto& p : pending) {
            float ms = 0.0f;
            PMH_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            auto& t = totals[p.name];
            t.first += ms;
            t.second += 1;
            spare.push_back(p.a);
            spare.push_back(p.b);
        }
        pending.clear();
    }

    hipStream_t stream = nullptr;
    int device = 0;
    std::map<std::string, std::pair<void*, size_t>> pool;
    float* ftab = nullptr;
    int ftab_P = -1;

    void* buf(const std::string& role, size_t bytes) {
        auto it = pool.find(role);
        if (it != pool.end() && it->second.second >= bytes) return it->second.first;
        if (it != pool.end()) {
            PMH_HIP(hipStreamSynchronize(stream));
            PMH_HIP(hipFree(it->second.first));
            pool.erase(it);
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw HipError("hipMalloc(" + std::to_string(bytes) + " B) for " + role + " failed",
                           (int)e);
        }
        pool[role] = {p, bytes};
        return p;
    }
    template <class T>
    T* get(const std::string& role, size_t count) {
        return static_cast<T*>(buf(role, count * sizeof(T)));
    }
    const float* factors(int P) {
        if (P != ftab_P) {
            const int S = 2 * P + 1;
            std::vector<float> h((size_t)S * S);
            factor_table(P, h.data());
            float* d = get<float>("factor_table", h.size());
            PMH_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                                   stream));
            PMH_HIP(hipStreamSynchronize(stream));
            ftab = d;
            ftab_P = P;
        }
        return ftab;
    }
    ~pmh_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : spare) (void)hipEventDestroy(e);
        for (auto& kv : pool) (void)hipFree(kv.second.first);
    }
};

namespace {

// Times the launches enqueued during its lifetime when profiling is on.
struct Timed {
    pmh_ctx* c;
    const char* name;
    hipEvent_t a = nullptr;
    Timed(pmh_ctx* c_, const char* n) : c(c_), name(n) {
        if (c->prof) {
            a = c->event();
            PMH_HIP(hipEventRecord(a, c->stream));
        }
    }
    ~Timed() noexcept(false) {
        if (a) {
            hipEvent_t b = c->event();
            PMH_HIP(hipEventRecord(b, c->stream));
            c->pending.push_back({name, a, b});
        }
    }
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return PMH_OK;
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

// Other-view images packed for the bilinear gathers of the window cost.
Gather pack(pmh_ctx* c, const std::string& role, const float* I, const float4* G, int W, int H) {
    const size_t n = (size_t)W * H;
    if (n * kRecBytes > 0x7fffffffull) throw ArgError("image too large for 32-bit gather offsets");
    float4* rec = c->get<float4>(role + ".rec", (n * kRecBytes + 15) / 16);
    launch_pack_other(I, G, W, H, rec, c->stream);
    return Gather{rec, (unsigned)n};
}

void zero_bytes(pmh_ctx* c, void* p, size_t bytes) {
    // a fill kernel (not hipMemsetAsync) so the sequence stays graph-capturable
    launch_fill_zero(p, bytes / sizeof(uint32_t), c->stream);
}

void pdrof(pmh_ctx* c, float4* u, const float4* f, const float* w, int W, int H, float lambda,
           int iters, const std::string& tag) {
    const size_t n = (size_t)W * H;
    float4* tmp = c->get<float4>(tag + ".pd_img", n);
    float4* p0 = c->get<float4>(tag + ".pd_p0", 2 * n);
    float4* p1 = c->get<float4>(tag + ".pd_p1", 2 * n);
    float4* img[2] = {tmp, u};
    float4* p[2] = {p0, p1};
    launch_fill_zero(p1, 2 * n * 4, c->stream);  // PDROF.h:62-64
    int swap = 0;
    for (int i = 0; i < iters; ++i) {  // PDROF.h:67-81
        swap = i & 1;
        Timed t(c, "pd_step");
        launch_pd_step(img[1 - swap], p[1 - swap], f, w, W, H, lambda, img[swap], p[swap],
                       c->stream);
    }
    if (!swap) {  // PDROF.h:83-95
        Timed t(c, "pd_step");
        launch_pd_step(img[0], p[0], f, w, W, H, lambda, img[1], p[1], c->stream);
    }
}

float smoothstep01(float x) {  // cvt/math/Math.h:88-92
    float t = (x < 0.0f) ? 0.0f : (x > 1.0f) ? 1.0f : x;
    return t * t * (3.0f - 2.0f * t);
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
}

int pmh_create(pmh_ctx** out, void* stream) {
    return guarded([&] {
        check_ptr(out, "ctx");
        *out = nullptr;
        int dev = 0;
        PMH_HIP(hipGetDevice(&dev));
        auto* c = new pmh_ctx();
        c->device = dev;
        c->stream = static_cast<hipStream_t>(stream);  // NULL: the default stream
        *out = c;
    });
}

int pmh_destroy(pmh_ctx* ctx) {
    return guarded([&] { delete ctx; });
}

int pmh_synchronize(pmh_ctx* ctx) {
    return guarded([&] {
        check_ctx(ctx);
        PMH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pmh_gradxy(pmh_ctx* ctx, const float* img, int W, int H, float* grad) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(img, "img"); check_ptr(grad, "grad"); check_size(W, H);
        launch_gradxy(img, W, H, reinterpret_cast<float4*>(grad), ctx->stream);
    });
}

int pmh_weight(pmh_ctx* ctx, const float* img, int W, int H, float* weight) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(img, "img"); check_ptr(weight, "weight"); check_size(W, H);
        launch_weight(img, W, H, weight, ctx->stream);
    });
}

int pmh_rng_draws(pmh_ctx* ctx, int W, int H, int iter, int stride, int count, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(out, "out"); check_size(W, H);
        if (count < 0 || count > 64) throw ArgError("count must be in [0, 64]");
        if (stride != (int)kPropStride && stride != (int)kInitStride)
            throw ArgError("stride must be 45 (propagate) or 3 (init)");
        if (iter < 0) throw ArgError("iter must be >= 0");
        launch_rng_draws(W, H, iter, (uint32_t)stride, count, out, ctx->stream);
    });
}

int pmh_init(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2, const float* G2,
             int W, int H, int P, float depthmax, int lr, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(I1, "I1"); check_ptr(G1, "G1"); check_ptr(I2, "I2");
        check_ptr(G2, "G2"); check_ptr(out, "out"); check_size(W, H); check_patch(P);
        const float* F = ctx->factors(P);
        Gather g = pack(ctx, "other", I2, reinterpret_cast<const float4*>(G2), W, H);
        launch_init(I1, reinterpret_cast<const float4*>(G1), g, F, W, H, P, depthmax, lr,
                    reinterpret_cast<float4*>(out), ctx->stream);
    });
}

int pmh_cost(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2, const float* G2,
             int W, int H, int P, const float* states, float* cost) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(I1, "I1"); check_ptr(G1, "G1"); check_ptr(I2, "I2");
        check_ptr(G2, "G2"); check_ptr(states, "states"); check_ptr(cost, "cost");
        check_size(W, H); check_patch(P);
        const float* F = ctx->factors(P);
        Gather g = pack(ctx, "other", I2, reinterpret_cast<const float4*>(G2), W, H);
        launch_cost(I1, reinterpret_cast<const float4*>(G1), g, F, W, H, P,
                    reinterpret_cast<const float4*>(states), cost, ctx->stream);
    });
}

int pmh_propagate(pmh_ctx* ctx, const float* old, float* out, const float* I1, const float* G1,
                  const float* I2, const float* G2, const float* smooth, float theta, int W,
                  int H, int P, float depthmax, int lr, int iter, const int* viewin,
                  const float* viewsrc, const float* draws, int sweep) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(old, "old"); check_ptr(out, "out"); check_ptr(I1, "I1");
        check_ptr(G1, "G1"); check_ptr(I2, "I2"); check_ptr(G2, "G2");
        check_ptr(smooth, "smooth"); check_size(W, H); check_patch(P);
        if (viewin && !viewsrc) throw ArgError("viewin given without viewsrc");
        if (sweep < PMH_SWEEP_ALL || sweep > PMH_SWEEP_BLACK)
            throw ArgError("sweep must be PMH_SWEEP_ALL, RED or BLACK");
        if (old == out && sweep == PMH_SWEEP_ALL)
            throw ArgError("Jacobi sweep needs distinct old and out fields");
        if (iter < 0) throw ArgError("iter must be >= 0");
        PropParams p{};
        p.old = reinterpret_cast<const float4*>(old);
        p.out = reinterpret_cast<float4*>(out);
        p.I1 = I1;
        p.G1 = reinterpret_cast<const float4*>(G1);
        p.g2 = pack(ctx, "other", I2, reinterpret_cast<const float4*>(G2), W, H);
        p.smooth = reinterpret_cast<const float4*>(smooth);
        p.viewin = reinterpret_cast<const ViewRec*>(viewin);
        p.viewsrc = reinterpret_cast<const float4*>(viewsrc);
        p.draws = draws;
        p.F = ctx->factors(P);
        p.W = W; p.H = H; p.P = P;
        p.theta = theta; p.depthmax = depthmax;
        p.lr = lr ? 1 : 0; p.iter = iter; p.sweep = sweep;
        p.random_refine = 2;
        launch_propagate(p, ctx->stream);
    });
}

int pmh_viewbuf_build(pmh_ctx* ctx, const float* states, int W, int H, int* viewbuf) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(states, "states"); check_ptr(viewbuf, "viewbuf");
        check_size(W, H);
        launch_viewbuf_build(reinterpret_cast<const float4*>(states), W, H,
                             reinterpret_cast<ViewRec*>(viewbuf), ctx->stream);
    });
}

int pmh_consistency(pmh_ctx* ctx, const float* left, const float* right, int W, int H,
                    float maxdispdiff, float maxanglediff, int lr, float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(left, "left"); check_ptr(right, "right"); check_ptr(out, "out");
        check_size(W, H);
        launch_consistency(reinterpret_cast<const float4*>(left),
                           reinterpret_cast<const float4*>(right), W, H, maxdispdiff,
                           maxanglediff, lr, reinterpret_cast<float4*>(out), ctx->stream);
    });
}

int pmh_fill_state(pmh_ctx* ctx, const float* in, int W, int H, float depthmax, int lr,
                   float* out) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(out, "out"); check_size(W, H);
        if (in == out) throw ArgError("fill_state cannot run in place");
        launch_fill_state(reinterpret_cast<const float4*>(in), W, H, depthmax, lr,
                          reinterpret_cast<float4*>(out), ctx->stream);
    });
}

int pmh_pd_step(pmh_ctx* ctx, const float* last, const float* plast, const float* f,
                const float* w, int W, int H, float lambda, float* out, float* pout) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(last, "last"); check_ptr(plast, "plast"); check_ptr(f, "f");
        check_ptr(w, "w"); check_ptr(out, "out"); check_ptr(pout, "pout"); check_size(W, H);
        launch_pd_step(reinterpret_cast<const float4*>(last),
                       reinterpret_cast<const float4*>(plast),
                       reinterpret_cast<const float4*>(f), w, W, H, lambda,
                       reinterpret_cast<float4*>(out), reinterpret_cast<float4*>(pout),
                       ctx->stream);
    });
}

int pmh_pdrof(pmh_ctx* ctx, float* u, const float* f, const float* w, int W, int H, float lambda,
              int iters) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(u, "u"); check_ptr(f, "f"); check_ptr(w, "w");
        check_size(W, H);
        if (iters < 0) throw ArgError("iters must be >= 0");
        pdrof(ctx, reinterpret_cast<float4*>(u), reinterpret_cast<const float4*>(f), w, W, H,
              lambda, iters, "pdrof");
    });
}

int pmh_fill_depthmap(pmh_ctx* ctx, const float* in, int W, int H, float scale, float* dmap) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(dmap, "dmap"); check_size(W, H);
        launch_fill_depthmap(reinterpret_cast<const float4*>(in), W, H, scale, dmap,
                             ctx->stream);
    });
}

int pmh_fill_normalmap(pmh_ctx* ctx, const float* in, int W, int H, float* nmap) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(in, "in"); check_ptr(nmap, "nmap"); check_size(W, H);
        launch_fill_normalmap(reinterpret_cast<const float4*>(in), W, H,
                              reinterpret_cast<float4*>(nmap), ctx->stream);
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
        if (o.sweep != PMH_SWEEP_ALL && o.sweep != PMH_SWEEP_REDBLACK)
            throw ArgError("options.sweep must be PMH_SWEEP_ALL or PMH_SWEEP_REDBLACK");
        if (o.pd_iters < 0) throw ArgError("options.pd_iters must be >= 0");
        if (o.random_refine < 0 || o.random_refine > 2)
            throw ArgError("options.random_refine must be in [0, 2]");
        if (o.grad_steps < 0 || o.grad_steps > 16)
            throw ArgError("options.grad_steps must be in [0, 16]");
        if (dscale <= 0.0f) dscale = 1.0f / depthmax;  // PMHuberStereo.cpp:75-76
        const float thetascale = 50.0f;                 // PMHuberStereo.cpp:72
        float theta = 0.0f;
        hipStream_t s = ctx->stream;
        const size_t n = (size_t)W * H;

        float4* lg = ctx->get<float4>("dm.lgrad", n);
        float4* rg = ctx->get<float4>("dm.rgrad", n);
        float* lw = ctx->get<float>("dm.lweight", n);
        float* rw = ctx->get<float>("dm.rweight", n);
        float4* ls = ctx->get<float4>("dm.lsmooth", n);
        float4* rs = ctx->get<float4>("dm.rsmooth", n);
        float4* tmp = ctx->get<float4>("dm.tmp", n);
        float4* tmp2 = ctx->get<float4>("dm.tmp2", n);
        float4* L[2] = {ctx->get<float4>("dm.L0", n), ctx->get<float4>("dm.L1", n)};
        float4* R[2] = {ctx->get<float4>("dm.R0", n), ctx->get<float4>("dm.R1", n)};
        ViewRec* vb1 = ctx->get<ViewRec>("dm.vb1", n);
        ViewRec* vb2 = ctx->get<ViewRec>("dm.vb2", n);
        const float* F = ctx->factors(P);

        // PMHuberStereo.cpp:90-113
        Gather gl{}, gr{};
        {
            Timed t(ctx, "prep");
            launch_gradxy(left, W, H, lg, s);
            launch_gradxy(right, W, H, rg, s);
            launch_weight(left, W, H, lw, s);
            launch_weight(right, W, H, rw, s);
            zero_bytes(ctx, ls, n * sizeof(float4));
            zero_bytes(ctx, rs, n * sizeof(float4));
            gl = pack(ctx, "dm.left", left, lg, W, H);    // left as the other view
            gr = pack(ctx, "dm.right", right, rg, W, H);  // right as the other view
        }

        // :124-149
        {
            Timed t(ctx, "init");
            launch_init(left, lg, gr, F, W, H, P, depthmax, 1, L[0], s);
            launch_init(right, rg, gl, F, W, H, P, depthmax, 0, R[0], s);
        }
        zero_bytes(ctx, vb2, n * sizeof(ViewRec));
        const float4* vb2src = R[0];

        const int nsweep = o.sweep == PMH_SWEEP_REDBLACK ? 2 : 1;
        for (int iter = 0; iter < iterations; ++iter) {  // :151-266
            const int swap = iter & 1;
            for (int view = 0; view < 2; ++view) {
                PropParams p{};
                p.old = view ? R[swap] : L[swap];
                p.out = view ? R[1 - swap] : L[1 - swap];
                p.I1 = view ? right : left;
                p.G1 = view ? rg : lg;
                p.g2 = view ? gl : gr;
                p.smooth = view ? rs : ls;
                p.viewin = view ? vb1 : vb2;
                p.viewsrc = view ? L[1 - swap] : vb2src;
                p.draws = nullptr;
                p.F = F;
                p.W = W; p.H = H; p.P = P;
                p.theta = theta; p.depthmax = depthmax;
                p.lr = view ? 0 : 1; p.iter = iter;
                p.random_refine = o.random_refine;
                for (int sw = 0; sw < nsweep; ++sw) {
                    p.sweep = nsweep == 1 ? PMH_SWEEP_ALL : (sw == 0 ? PMH_SWEEP_RED : PMH_SWEEP_BLACK);
                    Timed t(ctx, "propagate");
                    launch_propagate(p, s);
                }
                if (o.grad_steps > 0) {
                    Timed t(ctx, "grad_refine");
                    launch_grad_refine(p.out, p.out, p.I1, p.G1, p.g2, F, p.smooth, W, H, P, theta,
                                       depthmax, p.lr, o.grad_steps, o.grad_step, s);
                }
                Timed t(ctx, "viewbuf");
                launch_viewbuf_build(p.out, W, H, view ? vb2 : vb1, s);
            }
            vb2src = R[1 - swap];
            const float lambda = theta * thetascale + 5.0f;
            for (int view = 0; view < 2; ++view) {  // :228-243 (left), :246-261 (right)
                {
                    Timed t(ctx, "consistency");
                    launch_consistency(view ? R[1 - swap] : L[1 - swap],
                                       view ? L[1 - swap] : R[1 - swap], W, H, o.maxdispdiff,
                                       o.maxanglediff, view ? 0 : 1, tmp, s);
                }
                {
                    Timed t(ctx, "fill_state");
                    launch_fill_state(tmp, W, H, depthmax, view ? 0 : 1, tmp2, s);
                }
                pdrof(ctx, view ? rs : ls, tmp2, view ? rw : lw, W, H, lambda, o.pd_iters, "dm");
            }
            if (iter >= 5)  // :264-265
                theta = smoothstep01(((float)iter - 5.0f) / ((float)iterations - 5.0f)) * 1.0f;
        }

        const int fin = 1 - ((iterations - 1) & 1);  // :269-270
        Timed tfin(ctx, "fill_depthmap");
        launch_consistency(L[fin], R[fin], W, H, o.maxdispdiff, o.maxanglediff, 1, tmp, s);
        launch_fill_depthmap(tmp, W, H, dscale, dmap, s);  // :278-282
        if (normalmap)
            launch_fill_normalmap(tmp, W, H, reinterpret_cast<float4*>(normalmap), s);
        if (left_states)
            PMH_HIP(hipMemcpyAsync(left_states, L[fin], n * sizeof(float4),
                                   hipMemcpyDeviceToDevice, s));
        if (right_states)
            PMH_HIP(hipMemcpyAsync(right_states, R[fin], n * sizeof(float4),
                                   hipMemcpyDeviceToDevice, s));
        PMH_HIP(hipStreamSynchronize(s));
    });
}

int pmh_cost_backward(pmh_ctx* ctx, const float* I1, const float* G1, const float* I2,
                      const float* G2, int W, int H, int P, const float* states,
                      const float* gout, float* dstates, float* dI1, float* dG1, float* dI2,
                      float* dG2) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(I1, "I1"); check_ptr(G1, "G1"); check_ptr(I2, "I2");
        check_ptr(G2, "G2"); check_ptr(states, "states"); check_ptr(gout, "gout");
        check_ptr(dstates, "dstates"); check_size(W, H); check_patch(P);
        const float* F = ctx->factors(P);
        Gather g = pack(ctx, "other", I2, reinterpret_cast<const float4*>(G2), W, H);
        launch_cost_backward(I1, reinterpret_cast<const float4*>(G1), g, F, W, H, P,
                             reinterpret_cast<const float4*>(states), gout,
                             reinterpret_cast<float4*>(dstates), dI1,
                             reinterpret_cast<float4*>(dG1), dI2,
                             reinterpret_cast<float4*>(dG2), ctx->stream);
    });
}

int pmh_gradxy_backward(pmh_ctx* ctx, const float* dG, int W, int H, float* dI) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(dG, "dG"); check_ptr(dI, "dI"); check_size(W, H);
        launch_gradxy_backward(reinterpret_cast<const float4*>(dG), W, H, dI, ctx->stream);
    });
}

int pmh_grad_refine(pmh_ctx* ctx, float* states, const float* I1, const float* G1,
                    const float* I2, const float* G2, const float* smooth, float theta, int W,
                    int H, int P, float depthmax, int lr, int steps, float radius) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(states, "states"); check_ptr(I1, "I1"); check_ptr(G1, "G1");
        check_ptr(I2, "I2"); check_ptr(G2, "G2"); check_ptr(smooth, "smooth");
        check_size(W, H); check_patch(P);
        if (steps < 0 || steps > 16) throw ArgError("steps must be in [0, 16]");
        if (!(radius > 0.0f)) throw ArgError("radius must be positive");
        const float* F = ctx->factors(P);
        Gather g = pack(ctx, "other", I2, reinterpret_cast<const float4*>(G2), W, H);
        launch_grad_refine(reinterpret_cast<const float4*>(states),
                           reinterpret_cast<float4*>(states), I1,
                           reinterpret_cast<const float4*>(G1), g, F,
                           reinterpret_cast<const float4*>(smooth), W, H, P, theta, depthmax,
                           lr ? 1 : 0, steps, radius, ctx->stream);
    });
}

int pmh_profile_enable(pmh_ctx* ctx, int enable) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->prof = enable != 0;
    });
}

int pmh_profile_reset(pmh_ctx* ctx) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->resolve();
        ctx->totals.clear();
    });
}

int pmh_profile_read(pmh_ctx* ctx, const char* kernel, double* total_ms, long long* launches) {
    return guarded([&] {
        check_ctx(ctx); check_ptr(kernel, "kernel"); check_ptr(total_ms, "total_ms");
        check_ptr(launches, "launches");
        ctx->resolve();
        auto it = ctx->totals.find(kernel);
        *total_ms = it == ctx->totals.end() ? 0.0 : it->second.first;
        *launches = it == ctx->totals.end() ? 0 : it->second.second;
    });
}

}  // extern "C"
