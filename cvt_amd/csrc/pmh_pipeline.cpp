// pmh_pipeline.cpp -- the host sequence of PMHuberStereo::depthMap
// (cvt/vision/PMHuberStereo.cpp:64-291, without the debug block :153-185 that
// saves PNGs and waits on getchar) as one launch plan over a batch of pairs.
//
// Every launch covers the same step for all pairs (and both views where the
// reference's steps are independent: initialisation, consistency, fill and
// the Huber-ROF smoothing), so small pairs share the GPU instead of leaving it
// idle, and the whole sequence is one graph.
#include "pmh_pipeline.hpp"

#include <string>

namespace pmh {

int size_class(long long v) {
    int c = 0;
    while ((1ll << c) < v) ++c;
    return c;
}

// The first iteration (random planes; the left view without view candidates)
// is a workload of its own and is tuned apart from the later ones.
std::string propagate_tune_key(const PropParams& p, int nc, int total) {
    return "propagate|p" + std::to_string(p.P) + "|c" + std::to_string(nc) + "|s" +
           std::to_string(p.sweep) + "|t" + std::to_string(size_class(total)) + "|v" +
           std::to_string(p.lr ? 0 : 1) + "|r" + std::to_string(p.random_refine) + "|n" +
           std::to_string(p.skip_nbhd ? 0 : 6) + (p.iter == 0 ? "|i0" : "");
}

namespace {

float smoothstep01(float x) {  // cvt/math/Math.h:88-92
    float t = (x < 0.0f) ? 0.0f : (x > 1.0f) ? 1.0f : x;
    return t * t * (3.0f - 2.0f * t);
}

int zero_tiles(size_t words) { return (int)((words + kZeroTileWords - 1) / kZeroTileWords); }

}  // namespace

void add_pdrof(Plan& plan, const std::vector<PdItem>& items, float lambda, int iters,
               bool tunable, bool inpaint) {
    if (items.empty()) return;
    // PDROF.h:66-95: images {tmp, u}, duals {p0, p1}; step i writes index
    // (i & 1) from index 1 - (i & 1), starting from u and a cleared p1
    // (PDROF.h:62-64); an odd count takes one more step, so the result lands
    // in u after an even number of steps.  The steps run in groups of up to
    // kPdMaxSteps per launch (each group reads one buffer pair and writes the
    // other); an even number of groups ends in u as well.
    const int steps = iters + (iters & 1);
    int groups = (steps + kPdMaxSteps - 1) / kPdMaxSteps;
    groups += groups & 1;
    Batch<ZeroParams> z;
    for (const PdItem& it : items) {
        const size_t words = (size_t)it.W * it.H * 8;
        z.push(ZeroParams{reinterpret_cast<uint32_t*>(it.p1), (unsigned)words}, zero_tiles(words));
    }
    add_batch(plan, Kern::Zero, z, "zero");
    const char* name = inpaint ? "pd_steps_inpaint" : "pd_steps";
    for (int g = 0; g < groups; ++g) {
        const int to = g & 1 ? 1 : 0;  // group g writes index to, from 1 - to
        const int n = steps / groups + (g < steps % groups ? 1 : 0);
        Batch<PdParams> b;
        for (const PdItem& it : items) {
            float4* img[2] = {it.tmp, it.u};
            float4* p[2] = {it.p0, it.p1};
            PdParams q{img[1 - to], p[1 - to], it.f, it.w, img[to], p[to], it.W, it.H, it.mask};
            b.push(q, tiles_pd(it.W, it.H));
        }
        Launch* l = add_batch(plan, Kern::PdSteps, b, name);
        l->f0 = lambda;
        l->i0 = inpaint ? 1 : 0;
        l->i1 = n;
        if (tunable) {
            l->tunable = true;
            l->tune_key = std::string(name) + "|t" + std::to_string(size_class(l->total));
        }
    }
}

namespace {

// Plane fields and work buffers of one pyramid level; the final fields of a
// level are (*L[fin]), (*R[fin]).
struct Level {
    std::vector<float4*> L0, L1, R0, R1, cL;
    int fin = 0;
};

// The depthMap sequence of one level (PMHuberStereo.cpp:90-266) appended to
// the plan.  init: NULL (random planes, :124-149) or per-pair start fields.
Level add_level(Engine& e, Plan& plan, const float* F, const std::vector<PairIO>& pr, int P,
                int iterations, const pmh_options& o, const std::string& tag,
                const std::vector<float4*>* initL, const std::vector<float4*>* initR,
                bool inpaint, bool joint) {
    const int n = (int)pr.size();
    const int nc = pr[0].nc;
    std::vector<size_t> px(n), px2(n), rec(n);
    for (int i = 0; i < n; ++i) {
        px[i] = (size_t)pr[i].W * pr[i].H;
        px2[i] = 2 * px[i];
        rec[i] = (px[i] * rec_bytes(nc) + 15) / 16;
    }
    auto S4 = [&](const char* r) { return e.slab<float4>(tag + r, px); };
    auto lg = S4("lgrad"), rg = S4("rgrad");
    auto lw = e.slab<float>(tag + "lweight", px), rw = e.slab<float>(tag + "rweight", px);
    auto lsm = S4("lsmooth"), rsm = S4("rsmooth");
    Level lv;
    lv.L0 = S4("L0");
    lv.L1 = S4("L1");
    lv.R0 = S4("R0");
    lv.R1 = S4("R1");
    auto vb1 = e.slab<ViewRec>(tag + "vb1", px), vb2 = e.slab<ViewRec>(tag + "vb2", px);
    auto recL = e.slab<float4>(tag + "recL", rec), recR = e.slab<float4>(tag + "recR", rec);
    lv.cL = S4("consL");
    auto cR = S4("consR");
    auto fL = S4("fillL"), fR = S4("fillR");
    auto tL = S4("pdtmpL"), tR = S4("pdtmpR");
    auto p0L = e.slab<float4>(tag + "p0L", px2), p0R = e.slab<float4>(tag + "p0R", px2);
    auto p1L = e.slab<float4>(tag + "p1L", px2), p1R = e.slab<float4>(tag + "p1R", px2);
    std::vector<float4*>* L[2] = {&lv.L0, &lv.L1};
    std::vector<float4*>* R[2] = {&lv.R0, &lv.R1};
    auto gath = [&](const std::vector<float4*>& r, int i) {
        return Gather{r[i], (unsigned)px[i]};
    };

    {  // PMHuberStereo.cpp:90-113: gradients, weights, smooth fields cleared
        Batch<PrepParams> b;
        for (int i = 0; i < n; ++i) {
            const PairIO& q = pr[i];
            b.push(PrepParams{q.left, nullptr, lg[i], lw[i], recL[i], q.W, q.H, nc},
                   tiles_2d(q.W, q.H));
            b.push(PrepParams{q.right, nullptr, rg[i], rw[i], recR[i], q.W, q.H, nc},
                   tiles_2d(q.W, q.H));
        }
        add_batch(plan, Kern::Prep, b, "prep");
        Batch<ZeroParams> z;
        for (int i = 0; i < n; ++i) {
            const size_t w4 = px[i] * 4, wv = px[i] * (sizeof(ViewRec) / 4);
            z.push(ZeroParams{reinterpret_cast<uint32_t*>(lsm[i]), (unsigned)w4}, zero_tiles(w4));
            z.push(ZeroParams{reinterpret_cast<uint32_t*>(rsm[i]), (unsigned)w4}, zero_tiles(w4));
            z.push(ZeroParams{reinterpret_cast<uint32_t*>(vb2[i]), (unsigned)wv}, zero_tiles(wv));
            if (joint)  // the right view's first pass reads the left view's buffer too
                z.push(ZeroParams{reinterpret_cast<uint32_t*>(vb1[i]), (unsigned)wv},
                       zero_tiles(wv));
        }
        add_batch(plan, Kern::Zero, z, "zero");
    }
    {  // :124-149
        Batch<InitParams> b;
        for (int i = 0; i < n; ++i) {
            const PairIO& q = pr[i];
            b.push(InitParams{q.left, lg[i], gath(recR, i), initL ? (*initL)[i] : nullptr,
                              lv.L0[i], q.W, q.H, 1, q.depthmax},
                   tiles_waves(q.W, q.H, 0));
            b.push(InitParams{q.right, rg[i], gath(recL, i), initR ? (*initR)[i] : nullptr,
                              lv.R0[i], q.W, q.H, 0, q.depthmax},
                   tiles_waves(q.W, q.H, 0));
        }
        Launch* l = add_batch(plan, Kern::Init, b, "init");
        l->F = F;
        l->P = P;
        l->nc = nc;
    }

    const float thetascale = 50.0f;  // PMHuberStereo.cpp:72
    float theta = 0.0f;
    const bool rb = o.sweep == PMH_SWEEP_REDBLACK;
    for (int iter = 0; iter < iterations; ++iter) {  // :151-266
        const int swap = iter & 1;
        // joint: both views in each launch (v0 = 0, v1 = 2); otherwise the
        // left view's launches, then the right view's
        for (int v0 = 0; v0 < 2; v0 += joint ? 2 : 1) {
            const int v1 = joint ? 2 : v0 + 1;
            // the left view reads the view buffer built from the right
            // view's previous result (R[swap]); the right view reads the
            // left view's new result, or with joint views its previous one
            // (L[swap], scattered at the end of the previous iteration)
            for (int sw = 0; sw < (rb ? 2 : 1); ++sw) {
                const int sweep = rb ? (sw == 0 ? PMH_SWEEP_RED : PMH_SWEEP_BLACK) : PMH_SWEEP_ALL;
                Batch<PropParams> b;
                for (int view = v0; view < v1; ++view)
                for (int i = 0; i < n; ++i) {
                    const PairIO& q = pr[i];
                    PropParams p{};
                    p.old = view ? (*R[swap])[i] : (*L[swap])[i];
                    p.out = view ? (*R[1 - swap])[i] : (*L[1 - swap])[i];
                    p.I1 = view ? q.right : q.left;
                    p.G1 = view ? rg[i] : lg[i];
                    p.g2 = view ? gath(recL, i) : gath(recR, i);
                    p.smooth = view ? rsm[i] : lsm[i];
                    p.viewin = view ? vb1[i] : vb2[i];
                    p.viewsrc = view ? (*L[joint ? swap : 1 - swap])[i] : (*R[swap])[i];
                    p.draws = nullptr;
                    p.F = F;
                    p.W = q.W;
                    p.H = q.H;
                    p.P = P;
                    p.theta = theta;
                    p.depthmax = q.depthmax;
                    p.lr = view ? 0 : 1;
                    p.iter = iter;
                    p.sweep = sweep;
                    p.random_refine = o.random_refine;
                    p.skip_nbhd = o.random_neighbours == 0;
                    p.cands = e.prof ? e.cand_counter() : nullptr;
                    b.push(p, tiles_prop(q.W, q.H, sweep));
                }
                Launch* l = add_batch(plan, Kern::Propagate, b, "propagate");
                l->nc = nc;
                l->chunk = 0;  // persistent blocks by default
                l->tunable = true;
                l->tune_key = propagate_tune_key(b.items[0], nc, l->total);
            }
            if (o.grad_steps > 0) {
                Batch<RefineParams> b;
                for (int view = v0; view < v1; ++view)
                for (int i = 0; i < n; ++i) {
                    const PairIO& q = pr[i];
                    b.push(RefineParams{view ? (*R[1 - swap])[i] : (*L[1 - swap])[i],
                                        view ? q.right : q.left, view ? rg[i] : lg[i],
                                        view ? gath(recL, i) : gath(recR, i),
                                        view ? rsm[i] : lsm[i], q.W, q.H, view ? 0 : 1,
                                        q.depthmax},
                           tiles_waves(q.W, q.H, 0));
                }
                Launch* l = add_batch(plan, Kern::GradRefine, b, "grad_refine");
                l->F = F;
                l->P = P;
                l->nc = nc;
                l->f0 = theta;
                l->i0 = o.grad_steps;
                l->f1 = o.grad_step;
            }
            Batch<ViewParams> v;  // the scatter of pmhstereo.cl:391-401
            for (int view = v0; view < v1; ++view)
            for (int i = 0; i < n; ++i)
                v.push(ViewParams{view ? (*R[1 - swap])[i] : (*L[1 - swap])[i],
                                  view ? vb2[i] : vb1[i], pr[i].W, pr[i].H},
                       tiles_rows(pr[i].H));
            add_batch(plan, Kern::Viewbuf, v, "viewbuf");
        }
        // The smoothing below feeds only the next iteration's propagation (the
        // smooth prior and the smooth-plane candidate): after the last
        // iteration nothing reads it -- the final check reads the plane fields,
        // and a coarse level hands up its planes -- so it is not run.  The
        // outputs are those of the reference's sequence, which computes it.
        if (iter == iterations - 1) break;
        if (inpaint) {
            // depthMapInpaint :451-488: occlusion map, plane -> (n, d), masked smoother
            Batch<ConsParams> c;
            Batch<FillParams> f;
            std::vector<PdItem> pd;
            for (int i = 0; i < n; ++i) {
                const PairIO& q = pr[i];
                float4* Ln = (*L[1 - swap])[i];
                float4* Rn = (*R[1 - swap])[i];
                c.push(ConsParams{Ln, Rn, lv.cL[i], q.W, q.H, 1}, tiles_2d(q.W, q.H));
                c.push(ConsParams{Rn, Ln, cR[i], q.W, q.H, 0}, tiles_2d(q.W, q.H));
                f.push(FillParams{Ln, fL[i], q.W, q.H, 1, q.depthmax}, tiles_2d(q.W, q.H));
                f.push(FillParams{Rn, fR[i], q.W, q.H, 0, q.depthmax}, tiles_2d(q.W, q.H));
                pd.push_back(PdItem{lsm[i], fL[i], lw[i], tL[i], p0L[i], p1L[i], q.W, q.H,
                                    lv.cL[i]});
                pd.push_back(PdItem{rsm[i], fR[i], rw[i], tR[i], p0R[i], p1R[i], q.W, q.H,
                                    cR[i]});
            }
            Launch* l = add_batch(plan, Kern::Occmap, c, "occmap");
            l->f0 = o.maxdispdiff;
            l->f1 = o.maxanglediff;
            add_batch(plan, Kern::NormalDepth, f, "normal_depth");
            add_pdrof(plan, pd, theta * 20.0f + 5.0f, o.pd_iters, true, true);
            if (iter >= 4)  // :487-488
                theta = smoothstep01(((float)iter - 4.0f) / ((float)iterations - 4.0f)) * 1.0f;
            continue;
        }
        // :228-261: consistency, fill and Huber-ROF smoothing of both views
        Batch<ConsParams> c;
        Batch<FillParams> f;
        std::vector<PdItem> pd;
        for (int i = 0; i < n; ++i) {
            const PairIO& q = pr[i];
            float4* Ln = (*L[1 - swap])[i];
            float4* Rn = (*R[1 - swap])[i];
            c.push(ConsParams{Ln, Rn, lv.cL[i], q.W, q.H, 1}, tiles_2d(q.W, q.H));
            c.push(ConsParams{Rn, Ln, cR[i], q.W, q.H, 0}, tiles_2d(q.W, q.H));
            f.push(FillParams{lv.cL[i], fL[i], q.W, q.H, 1, q.depthmax}, tiles_rows(q.H));
            f.push(FillParams{cR[i], fR[i], q.W, q.H, 0, q.depthmax}, tiles_rows(q.H));
            pd.push_back(PdItem{lsm[i], fL[i], lw[i], tL[i], p0L[i], p1L[i], q.W, q.H, nullptr});
            pd.push_back(PdItem{rsm[i], fR[i], rw[i], tR[i], p0R[i], p1R[i], q.W, q.H, nullptr});
        }
        Launch* l = add_batch(plan, Kern::Consistency, c, "consistency");
        l->f0 = o.maxdispdiff;
        l->f1 = o.maxanglediff;
        add_batch(plan, Kern::FillState, f, "fill_state");
        add_pdrof(plan, pd, theta * thetascale + 5.0f, o.pd_iters, true);
        if (iter >= 5)  // :264-265
            theta = smoothstep01(((float)iter - 5.0f) / ((float)iterations - 5.0f)) * 1.0f;
    }
    lv.fin = 1 - ((iterations - 1) & 1);  // :269-270
    return lv;
}

}  // namespace

void depthmap_batch(Engine& e, const PairIO* pairs, int n, int P, int iterations,
                    const pmh_options& o, const FactorFn& factors, bool inpaint) {
    Plan plan;
    std::vector<PairIO> lv0(pairs, pairs + n);
    const int levels = o.coarse_levels;
    const int nc = pairs[0].nc;
    // image pyramids of every pair, level 1 .. levels (2x2 box reduction)
    std::vector<std::vector<PairIO>> pyr(levels + 1);
    pyr[0] = lv0;
    for (int l = 1; l <= levels; ++l) {
        const auto& up = pyr[l - 1];
        std::vector<size_t> px(n);
        for (int i = 0; i < n; ++i)
            px[i] = (size_t)((up[i].W + 1) / 2) * ((up[i].H + 1) / 2) * nc;
        const std::string tag = "pyr" + std::to_string(l) + ".";
        auto li = e.slab<float>(tag + "left", px), ri = e.slab<float>(tag + "right", px);
        Batch<PyrParams> b;
        pyr[l].resize(n);
        for (int i = 0; i < n; ++i) {
            const int w = (up[i].W + 1) / 2, h = (up[i].H + 1) / 2;
            b.push(PyrParams{up[i].left, li[i], up[i].W, up[i].H, nc}, tiles_2d(w, h));
            b.push(PyrParams{up[i].right, ri[i], up[i].W, up[i].H, nc}, tiles_2d(w, h));
            pyr[l][i] = PairIO{li[i], ri[i], w, h, nc, 0.5f * up[i].depthmax, 0.0f,
                               nullptr, nullptr, nullptr, nullptr};
        }
        add_batch(plan, Kern::PyrDown, b, "pyr_down");
    }
    // coarse to fine: each level starts from the planes of the level below
    std::vector<float4*> upL, upR;
    Level last;
    for (int l = levels; l >= 0; --l) {
        const int Pl = l > 0 && o.coarse_patch > 0 ? o.coarse_patch : P;
        const int it = l > 0 ? o.coarse_iters : iterations;
        pmh_options ol = o;
        if (l > 0 && o.coarse_pd_iters > 0) ol.pd_iters = o.coarse_pd_iters;
        if (l > 0 && o.coarse_random_refine >= 0) ol.random_refine = o.coarse_random_refine;
        if (l > 0 && o.coarse_random_neighbours >= 0)
            ol.random_neighbours = o.coarse_random_neighbours;
        const std::string tag = "b" + std::to_string(l) + ".";
        last = add_level(e, plan, factors(Pl), pyr[l], Pl, it, ol, tag, l < levels ? &upL : nullptr,
                         l < levels ? &upR : nullptr, inpaint, l > 0 && o.coarse_joint_views);
        if (l == 0) break;
        const auto& fine = pyr[l - 1];
        std::vector<size_t> px(n);
        for (int i = 0; i < n; ++i) px[i] = (size_t)fine[i].W * fine[i].H;
        upL = e.slab<float4>("up" + std::to_string(l - 1) + ".L", px);
        upR = e.slab<float4>("up" + std::to_string(l - 1) + ".R", px);
        const auto& Lf = last.fin ? last.L1 : last.L0;
        const auto& Rf = last.fin ? last.R1 : last.R0;
        Batch<UpParams> b;
        for (int i = 0; i < n; ++i) {
            const PairIO& c = pyr[l][i];
            b.push(UpParams{Lf[i], upL[i], fine[i].W, fine[i].H, c.W, c.H},
                   tiles_2d(fine[i].W, fine[i].H));
            b.push(UpParams{Rf[i], upR[i], fine[i].W, fine[i].H, c.W, c.H},
                   tiles_2d(fine[i].W, fine[i].H));
        }
        add_batch(plan, Kern::PlaneUp, b, "plane_up");
    }

    // depthMapInpaint reads the fields of index 1 whatever the iteration count
    // (PMHuberStereo.cpp:491-497), depthMap the last written ones (:269-270)
    const int fin = inpaint ? 1 : last.fin;
    const auto& Lfin = fin ? last.L1 : last.L0;
    const auto& Rfin = fin ? last.R1 : last.R0;
    const float cdisp = inpaint ? 1.0f : o.maxdispdiff;   // :494-495
    const float cdeg = inpaint ? 5.0f : o.maxanglediff;
    {  // :269-290: final consistency, disparity and normal maps
        Batch<ConsParams> c;
        Batch<FillDepthParams> d, nm;
        for (int i = 0; i < n; ++i) {
            const PairIO& q = pairs[i];
            c.push(ConsParams{Lfin[i], Rfin[i], last.cL[i], q.W, q.H, 1}, tiles_2d(q.W, q.H));
            d.push(FillDepthParams{last.cL[i], q.dmap, nullptr, q.W, q.H, q.dscale},
                   tiles_rows(q.H));  // :278-282
            if (q.nmap)
                nm.push(FillDepthParams{last.cL[i], nullptr, q.nmap, q.W, q.H, q.dscale},
                        tiles_rows(q.H));
        }
        Launch* l = add_batch(plan, Kern::Consistency, c, "consistency");
        l->f0 = cdisp;
        l->f1 = cdeg;
        add_batch(plan, Kern::FillDepth, d, "fill_depthmap");
        l = add_batch(plan, Kern::FillDepth, nm, "fill_normalmap");
        if (l) l->i0 = 1;
    }

    e.run(plan, o.use_graph != 0);
    for (int i = 0; i < n; ++i) {
        const size_t bytes = (size_t)pairs[i].W * pairs[i].H * sizeof(float4);
        if (pairs[i].lstates)
            PMH_HIP(hipMemcpyAsync(pairs[i].lstates, Lfin[i], bytes, hipMemcpyDeviceToDevice,
                                   e.stream));
        if (pairs[i].rstates)
            PMH_HIP(hipMemcpyAsync(pairs[i].rstates, Rfin[i], bytes, hipMemcpyDeviceToDevice,
                                   e.stream));
    }
    PMH_HIP(hipStreamSynchronize(e.stream));
}

}  // namespace pmh
