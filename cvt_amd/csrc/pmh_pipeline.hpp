// pmh_pipeline.hpp -- launch plans of PMHuberStereo::depthMap over a batch of
// pairs and of PDROF::apply.
#pragma once

#include <functional>
#include <string>

#include "pmh_capi.h"
#include "pmh_engine.hpp"

namespace pmh {

struct PairIO {
    const float* left;   // nc floats per pixel
    const float* right;
    int W, H;
    int nc;              // 1 gray, 4 RGBA (the same for every pair of a batch)
    float depthmax;
    float dscale;      // > 0
    float* dmap;
    float4* nmap;      // optional
    float4* lstates;   // optional: final left plane field
    float4* rstates;   // optional: final right plane field
};

// One PDROF::apply (PDROF.h:50-96) per item, all items in the same launches.
struct PdItem {
    float4* u;         // warm start in, result out
    const float4* f;
    const float* w;
    float4* tmp;       // W*H scratch
    float4* p0;        // 2*W*H scratch (dual)
    float4* p1;        // 2*W*H scratch (dual)
    int W, H;
    const float4* mask;  // PDROFInpaint::apply (PDROFInpaint.h:50-98): lambda scale in .x
};
// inpaint: the items are PDROFInpaint::apply calls (every item has a mask)
void add_pdrof(Plan& plan, const std::vector<PdItem>& items, float lambda, int iters,
               bool tunable, bool inpaint = false);

// Device support-factor table of a half window P.
using FactorFn = std::function<const float*(int)>;

// PMHuberStereo::depthMap (PMHuberStereo.cpp:64-291) for every pair, the pairs
// sharing each launch; with o.coarse_levels > 0 the planes start from a
// coarse-to-fine pass over 2x2-reduced images.  Synchronizes the stream
// before returning.
// inpaint: PMHuberStereo::depthMapInpaint (PMHuberStereo.cpp:293-507) instead.
void depthmap_batch(Engine& e, const PairIO* pairs, int n, int P, int iterations,
                    const pmh_options& o, const FactorFn& factors, bool inpaint = false);

// ceil(log2(v)): the size class of a work list in tuning keys
int size_class(long long v);

// Tuning key of a propagate launch: half window, colour format, sweep, size
// class of the work list, view (0 left, 1 right: the right view's random
// planes are steeper, so its rows overflow a narrow stage more often) and the
// options that change a wave's candidate list (randomized refinements,
// random-neighbourhood candidates).
std::string propagate_tune_key(const PropParams& p, int nc, int total);

}  // namespace pmh
