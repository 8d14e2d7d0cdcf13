// pmh_internal.hpp -- host-side declarations shared by the C-ABI and the
// kernel launchers.  Not installed; the public interface is include/pmh_capi.h.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "pmh_device.hpp"

namespace pmh {

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define PMH_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::pmh::HipError(std::string(#call) + ": " + hipGetErrorString(e_),   \
                                  (int)e_);                                            \
    } while (0)

// Other-view images packed for bilinear gathers in x (see pack_other): three
// planes in one allocation, so that lanes reading neighbouring texels touch
// as few cache lines as possible -- G[x] (float4), G[x+1] - G[x] (float4) and
// (I[x], I[x+1] - I[x]) (float2); the last column has zero deltas.
struct Gather {
    const float4* rec;  // planes G | dG | ID, 40 * W * H bytes
    unsigned n;         // texels per plane (W * H)
};
constexpr unsigned kRecBytes = 40;

struct PropParams {
    const float4* old;
    float4* out;
    const float* I1;
    const float4* G1;
    Gather g2;
    const float4* smooth;
    const ViewRec* viewin;
    const float4* viewsrc;
    const float* draws;  // optional, 21 per pixel
    const float* F;      // (2P+1)^2 support factors
    int W, H, P;
    float theta, depthmax;
    int lr, iter, sweep;  // sweep: 0 all, 1 red, 2 black
    int random_refine;    // randomized refinements per pixel (reference: 2)
};

// SkipTable for RNG_init strides 45 and 3, computed once on the host.
const SkipTable& skip_table(uint32_t stride);

void launch_pack_other(const float* I, const float4* G, int W, int H, float4* rec, hipStream_t s);
void launch_gradxy(const float* img, int W, int H, float4* grad, hipStream_t s);
void launch_weight(const float* img, int W, int H, float* w, hipStream_t s);
void launch_init(const float* I1, const float4* G1, Gather g2, const float* F, int W, int H,
                 int P, float depthmax, int lr, float4* out, hipStream_t s);
void launch_cost(const float* I1, const float4* G1, Gather g2, const float* F, int W, int H,
                 int P, const float4* states, float* cost, hipStream_t s);
void launch_propagate(const PropParams& p, hipStream_t s);
void launch_viewbuf_build(const float4* states, int W, int H, ViewRec* out, hipStream_t s);
void launch_consistency(const float4* left, const float4* right, int W, int H, float maxdisp,
                        float maxdeg, int lr, float4* out, hipStream_t s);
void launch_fill_state(const float4* in, int W, int H, float depthmax, int lr, float4* out,
                       hipStream_t s);
void launch_pd_step(const float4* last, const float4* plast, const float4* f, const float* w,
                    int W, int H, float lambda, float4* out, float4* pout, hipStream_t s);
// zero `words` 32-bit words with a kernel (graph-capturable, unlike hipMemsetAsync)
void launch_fill_zero(void* p, size_t words, hipStream_t s);
void launch_fill_depthmap(const float4* in, int W, int H, float scale, float* dmap,
                          hipStream_t s);
void launch_fill_normalmap(const float4* in, int W, int H, float4* nmap, hipStream_t s);
void launch_rng_draws(int W, int H, int iter, uint32_t stride, int count, float* out,
                      hipStream_t s);

void launch_cost_backward(const float* I1, const float4* G1, Gather g2, const float* F, int W,
                          int H, int P, const float4* states, const float* gout, float4* dst,
                          float* dI1, float4* dG1, float* dI2, float4* dG2, hipStream_t s);
void launch_gradxy_backward(const float4* dG, int W, int H, float* dI, hipStream_t s);
void launch_grad_refine(const float4* in, float4* out, const float* I1, const float4* G1,
                        Gather g2, const float* F, const float4* smooth, int W, int H, int P,
                        float theta, float depthmax, int lr, int steps, float radius,
                        hipStream_t s);

// Host factor table of pmhstereo.cl:199 (same expression order as the kernel).
void factor_table(int P, float* table);

constexpr int kMaxRowWidth = 8192;  // row-scan kernels keep one row in LDS

}  // namespace pmh
