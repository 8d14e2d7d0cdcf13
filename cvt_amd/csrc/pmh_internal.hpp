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
struct Unsupported : std::runtime_error {  // valid request outside what this build does
    using std::runtime_error::runtime_error;
};

#define PMH_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::pmh::HipError(std::string(#call) + ": " + hipGetErrorString(e_),   \
                                  (int)e_);                                            \
    } while (0)

// Input images hold nc floats per pixel: 1 (gray, CL_INTENSITY) or 4 (RGBA).
constexpr int kGray = 1, kRGBA = 4;

// Other-view images packed for bilinear gathers in x (prep_kernel): three
// planes in one allocation, so that lanes reading neighbouring texels touch
// as few cache lines as possible -- G[x] (float4), G[x+1] - G[x] (float4) and
// (I[x], I[x+1] - I[x]) (float2 gray, 2 x float4 RGBA); the last column has
// zero deltas.
struct Gather {
    const float4* rec;  // planes G | dG | ID, rec_bytes(nc) * W * H bytes
    unsigned n;         // texels per plane (W * H)
    unsigned pad_ = 0;
};
inline unsigned rec_bytes(int nc) { return nc == kRGBA ? 64u : 40u; }

// Work list of a batched launch (see for_tiles in pmh_window.hpp).
struct Tiles {
    const int* prefix;  // items + 1 entries, device memory
    int items;
    int total;
    int chunk;          // tiles per block; 0: persistent blocks pulling tiles from `queue`
    int stride = 1;     // tile t of the launch is tile t * stride of the work list (a sample)
    int* queue = nullptr;  // kTileQueueWords counters of the persistent mode, zero between launches
};
constexpr int kTileQueueWords = 9;  // 8 per-XCD band counters + blocks finished
inline int tiles_grid(int total, int chunk) {
    const int nb = (total + chunk - 1) / chunk;
    return ((nb + 7) >> 3) << 3;
}
// Grid of a persistent launch of `kernel` with `threads` per block: as many
// blocks as the device holds at once (a multiple of 8), never more than tiles.
int persistent_grid(const void* kernel, int threads, int total);
template <class K>
inline int grid_of(K kernel, int threads, const Tiles& T) {
    if (T.chunk > 0) return tiles_grid(T.total, T.chunk);
    return persistent_grid(reinterpret_cast<const void*>(kernel), threads, T.total);
}

// Per-item parameters of the batched kernels (device arrays, one per item).
// A plan stores identical parameter arrays once and the graph cache keys on
// their offsets, both by comparing bytes: these structs have no implicit
// padding (explicit pad_ members, zero), so equal parameters are equal bytes
// (tests/test_capi.py compiles this header with -Wpadded).
struct PrepParams {     // gradxy + weight + gather planes of one image
    const float* img;   // nc floats per pixel
    const float4* gin;  // optional: gradient image to pack (default: computed from img)
    float4* grad;       // optional
    float* weight;      // optional
    float4* rec;        // optional (gather planes of this image as the other view)
    int W, H, nc;
    int pad_ = 0;
};
struct ZeroParams {
    uint32_t* p;
    unsigned words;
    unsigned pad_ = 0;
};
struct InitParams {
    const float* I1;    // nc floats per pixel (the launch's nc)
    const float4* G1;
    Gather g2;
    const float4* src;  // optional: start from these planes instead of random ones
    float4* out;
    int W, H, lr;
    float depthmax;
};
struct PyrParams {      // 2x2 box reduction of an input image (per channel)
    const float* in;
    float* out;
    int W, H;           // input size; output is ((W+1)/2, (H+1)/2)
    int nc;
    int pad_ = 0;
};
struct UpParams {       // plane field of a half-resolution level -> this level
    const float4* in;
    float4* out;
    int W, H;           // output size
    int w, h;           // input size
};
struct ViewParams {
    const float4* states;
    ViewRec* out;
    int W, H;
};
struct ConsParams {
    const float4* left;
    const float4* right;
    float4* out;
    int W, H, lr;
    int pad_ = 0;
};
struct FillParams {
    const float4* in;
    float4* out;
    int W, H, lr;
    float depthmax;
};
struct PdParams {
    const float4* last;
    const float4* plast;
    const float4* f;
    const float* w;
    float4* out;
    float4* pout;
    int W, H;
    const float4* mask;  // inpainting step only: lambda scale in .x
};
struct FillDepthParams {
    const float4* in;
    float* dmap;     // optional
    float4* nmap;    // optional
    int W, H;
    float scale;
    int pad_ = 0;
};
struct RefineParams {
    float4* states;  // refined in place
    const float* I1;
    const float4* G1;
    Gather g2;
    const float4* smooth;
    int W, H, lr;
    float depthmax;
};

// tiles per item of each kernel family
inline int tiles_2d(int W, int H) { return ((W + 63) / 64) * ((H + 3) / 4); }  // 64x4 px
// fused Huber-ROF steps: output tile of kPdTileW x kPdTileH px, up to kPdMaxSteps steps
constexpr int kPdTileW = 32, kPdTileH = 16, kPdMaxSteps = 4;
inline int tiles_pd(int W, int H) {
    return ((W + kPdTileW - 1) / kPdTileW) * ((H + kPdTileH - 1) / kPdTileH);
}
inline int tiles_rows(int H) { return H; }                                    // one row
inline int tiles_waves(int W, int H, int sweep) {                             // 64-lane segments
    const int lanes = sweep == 0 ? W : (W + 1) / 2;
    return ((lanes + 63) / 64) * H;
}
// Propagate passes: a row segment of the block-stage kernels holds 63 pixels
// of a red or black pass (62 of a Jacobi sweep) on its 64 lanes; the first
// lanes re-hold the previous segment's last pixels, so that every owned
// pixel's left neighbour column sits on a lane of the same wave.  Every
// propagate launch shape decodes its tiles with this count (the 64-pixel
// shapes leave a tile past their last segment empty).
__host__ __device__ constexpr int prop_seg_pixels(int sweep) { return sweep == 0 ? 62 : 63; }
__host__ __device__ inline int prop_tiles_x(int W, int sweep) {
    const int lanes = sweep == 0 ? W : (W + 1) / 2;
    return (lanes + prop_seg_pixels(sweep) - 1) / prop_seg_pixels(sweep);
}
inline int tiles_prop(int W, int H, int sweep) { return prop_tiles_x(W, sweep) * H; }
constexpr unsigned kZeroTileWords = 4096;

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
    int skip_nbhd;        // 1: the six random-neighbourhood candidates (k = 9 .. 14) are skipped
    unsigned long long* cands;  // optional: += window-cost candidates of the algorithm, per pixel
                                // 16 (10 without the neighbourhood) + its view candidates
                                // + random_refine (a measurement counter)
};

// SkipTable for RNG_init strides 45, 3 and 33, computed once on the host.
const SkipTable& skip_table(uint32_t stride);

void launch_prep(const PrepParams* pp, Tiles T, hipStream_t s);
void launch_zero(const ZeroParams* pp, Tiles T, hipStream_t s);
void launch_pyr_down(const PyrParams* pp, Tiles T, hipStream_t s);
void launch_plane_up(const UpParams* pp, Tiles T, hipStream_t s);
void launch_init(const InitParams* pp, Tiles T, const float* F, int P, int nc, hipStream_t s);
// stage: LDS stage width of the window cost (texels of the other view per
// window row) of the one-wave kernel, one of kStageChoices; kStageSplit: the
// format's default stage width with a row segment's candidates spread over
// four waves of a block, one window sweep per candidate; kStageSplitMulti (and
// 0, the default): the same with each wave sweeping its candidates together;
// kStageSplit3: that kernel at a 576 / 352-texel stage and three waves per
// SIMD.  Stage widths: 20 B per staged texel gray, 32 B RGBA; the defaults are
// about 18 / 16 KB per wave.
// kStageBlock: the four waves of a block share one stage of kBlockStage
// 40-byte records (G, dG, I, dI) and sweep the window rows in lockstep, each
// on its 4-5 candidates; kStageBlock3: the same at kBlockStage3 records and
// three waves per SIMD.  Gray inputs (RGBA launches take the default shape).
constexpr int kStageSplit = 1;
constexpr int kStageSplitMulti = 2;
constexpr int kStageSplit3 = 3;
constexpr int kStageBlock = 4;
constexpr int kStageBlock3 = 5;
constexpr int kBlockStage = 1792, kBlockStage3 = 1152;
// The shape of stage code 0: the block-stage kernel for gray inputs, the
// shared-sweep four-wave kernel for RGBA.
constexpr int propagate_stage_default(int nc) { return nc == 1 ? kStageBlock : kStageSplitMulti; }
constexpr int kStageChoices[8] = {512, 704, 896, kStageSplit, kStageSplitMulti, kStageSplit3,
                                  kStageBlock, kStageBlock3};
constexpr int stage_default(int nc) { return nc == kGray ? 896 : 512; }
// The narrower stage of the one-plane kernels (init, cost, gradient
// refinement) and of stage code 3: 11.5 KB per wave gray, 11.3 KB RGBA, so the
// LDS holds three or more one-wave blocks per SIMD instead of two (gradient
// refinement 14% faster; the rows of one plane over 64 contiguous pixels
// rarely span more).
constexpr int stage_narrow(int nc) { return nc == kGray ? 576 : 352; }
void launch_propagate(const PropParams* pp, Tiles T, int stage, int nc, hipStream_t s);
void launch_viewbuf(const ViewParams* pp, Tiles T, hipStream_t s);
void launch_consistency(const ConsParams* pp, Tiles T, float maxdisp, float maxdeg,
                        hipStream_t s);
void launch_fill_state(const FillParams* pp, Tiles T, hipStream_t s);
void launch_pd_step(const PdParams* pp, Tiles T, float lambda, bool inpaint, hipStream_t s);
// `steps` (1..kPdMaxSteps) consecutive pd steps in one launch (tiles_pd work
// list); the tile's results equal `steps` launches of launch_pd_step bit for
// bit.  Reads last/plast, writes out/pout (distinct buffers).
void launch_pd_steps(const PdParams* pp, Tiles T, float lambda, bool inpaint, int steps,
                     hipStream_t s);
void launch_occmap(const ConsParams* pp, Tiles T, float maxdisp, float maxdeg, hipStream_t s);
void launch_normal_depth(const FillParams* pp, Tiles T, hipStream_t s);
void launch_fill_depth(const FillDepthParams* pp, Tiles T, bool normals, hipStream_t s);
void launch_grad_refine(const RefineParams* pp, Tiles T, const float* F, int P, float theta,
                        int steps, float radius, int nc, hipStream_t s);

// single-pair kernels outside the depthMap pipeline
void launch_cost(const float* I1, const float4* G1, Gather g2, const float* F, int W, int H,
                 int P, const float4* states, float* cost, int nc, hipStream_t s);
void launch_rng_draws(int W, int H, int iter, uint32_t stride, int count, float* out,
                      hipStream_t s);
void launch_cost_backward(const float* I1, const float4* G1, Gather g2, const float* F, int W,
                          int H, int P, const float4* states, const float* gout, float4* dst,
                          float* dI1, float4* dG1, float* dI2, float4* dG2, int nc,
                          hipStream_t s);
void launch_gradxy_backward(const float4* dG, int W, int H, float* dI, hipStream_t s);

// Host factor table of pmhstereo.cl:199 (same expression order as the kernel).
// Support-factor tables of a half window P, kFactorTables * (2P+1)^2 floats:
// the factors F(dx, dy) of the reference, then the weight exponents
// -3 F log2(e) (gray: the RGB sum is 3 |c0 - v1|) and -F log2(e) (RGBA), so
// that a tap's weight exp(-sad F) is one multiply and one exp2.
constexpr int kFactorTables = 3;
void factor_table(int P, float* table);

constexpr int kMaxRowWidth = 8192;  // row-scan kernels keep one row in LDS

// pmh_pdrof_noalloc: PDROF::apply with its temporary image and duals (80
// bytes per pixel) in a file-scope device array of kNoAllocPixels pixels
// (80 MiB per GPU); the address is looked up per device by
// noalloc_scratch_init (called by pmh_create with the device current).
constexpr int kNoAllocPixels = 1 << 20;
constexpr size_t kNoAllocScratchFloat4 = (size_t)kNoAllocPixels * 5;  // tmp + 2 x 2 dual
constexpr int kMaxDevices = 64;
void noalloc_scratch_init(int device);
void launch_pdrof_noalloc(int device, float4* u, const float4* f, const float* w, int W, int H,
                          float lambda, int iters, hipStream_t s);

}  // namespace pmh
