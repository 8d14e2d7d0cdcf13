// pmh_engine.hpp -- launch plans over batches of items, their execution
// (eager or as one captured HIP graph) and the launch-shape autotuner.
//
// A plan is a list of kernel launches plus one host "arena" holding every
// launch's per-item parameter array and tile prefix table.  Executing a plan
// uploads the arena to device memory with one copy and then enqueues the
// launches, whose only pointer arguments are into that arena (and the
// support-factor table).  A captured graph therefore stays valid for any call
// with the same launch list: new buffers only change the arena contents.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "pmh_internal.hpp"

namespace pmh {

enum class Kern : int {
    Prep, Zero, Init, Propagate, Viewbuf, Consistency, FillState, PdStep, FillDepth, GradRefine,
    PyrDown, PlaneUp, Occmap, NormalDepth, PdSteps
};

struct Launch {
    Kern kind;
    size_t params;        // arena offset of the per-item parameter array
    size_t prefix;        // arena offset of the tile prefix table
    int items, total;     // items, tiles over all items
    float f0 = 0, f1 = 0; // maxdisp / maxdeg, lambda, theta / radius
    int i0 = 0;           // normals flag, refine steps, inpainting smoother flag
    int i1 = 0;           // fused Huber-ROF steps per launch
    int P = 0;
    int nc = 1;           // floats per input pixel (1 gray, 4 RGBA)
    const float* F = nullptr;
    bool tunable = false; // pick stage / chunk with the autotuner
    std::string tune_key; // shape class of a tunable launch
    int stage = 0;        // propagate: LDS stage width (0: the format's default)
    int chunk = 1;        // tiles per block (pixels per wave = 64 * chunk for wave kernels);
                          // 0: persistent blocks taking tiles from per-XCD queues
    int stride = 1;       // tile stride (the autotuner's sample of a launch)
    const char* name = "";
};

class Plan {
public:
    // Appends a parameter array (16-byte aligned); identical arrays are
    // stored once.
    template <class P>
    size_t params(const std::vector<P>& v) {
        return blob(v.data(), v.size() * sizeof(P));
    }
    // Prefix table of per-item tile counts; returns its offset, sets total.
    size_t prefix(const std::vector<int>& counts, int& total);
    Launch& add(Kern k, size_t params, size_t prefix, int items, int total, const char* name);

    std::vector<unsigned char> arena;
    std::vector<Launch> launches;

private:
    size_t blob(const void* p, size_t bytes);
    std::unordered_map<std::string, size_t> seen_;
};

// Tiled parameter list of one launch: items with their tile counts.
template <class P>
struct Batch {
    std::vector<P> items;
    std::vector<int> tiles;
    void push(const P& p, int t) {
        if (t <= 0) return;  // nothing to do for this item
        items.push_back(p);
        tiles.push_back(t);
    }
    bool empty() const { return items.empty(); }
};

template <class P>
Launch* add_batch(Plan& plan, Kern k, const Batch<P>& b, const char* name) {
    if (b.empty()) return nullptr;
    int total = 0;
    const size_t pre = plan.prefix(b.tiles, total);
    const size_t par = plan.params(b.items);
    return &plan.add(k, par, pre, (int)b.items.size(), total, name);
}

// Device memory, pinned staging, graphs and tuned launch shapes of a context.
class Engine {
public:
    explicit Engine(hipStream_t s) : stream(s) {}
    ~Engine();

    hipStream_t stream;

    // grow-only scratch pool keyed by role
    void* buf(const std::string& role, size_t bytes);
    template <class T>
    T* get(const std::string& role, size_t count) {
        return static_cast<T*>(buf(role, count * sizeof(T)));
    }
    // one allocation per role, split into 256-byte aligned slices
    template <class T>
    std::vector<T*> slab(const std::string& role, const std::vector<size_t>& counts) {
        std::vector<size_t> off(counts.size());
        size_t total = 0;
        for (size_t i = 0; i < counts.size(); ++i) {
            off[i] = total;
            total += (counts[i] * sizeof(T) + 255) & ~(size_t)255;
        }
        auto* base = static_cast<unsigned char*>(buf(role, total));
        std::vector<T*> out(counts.size());
        for (size_t i = 0; i < counts.size(); ++i) out[i] = reinterpret_cast<T*>(base + off[i]);
        return out;
    }

    // Runs a plan on the stream.  graph: capture the launch list once and
    // replay it on later calls with the same list (the first call of a list
    // with untuned launches runs eagerly and tunes them).
    void run(Plan& plan, bool graph);

    // measurement hook: events around eager launches (all, or those named
    // prof_only), resolved on read; with prof, propagate launches also count
    // their candidates into cand_counter()
    bool prof = false;
    std::string prof_only;
    unsigned long long* cand_counter();
    void resolve();
    // Brackets a launch made outside a plan with events under the same rules.
    template <class F>
    void timed_launch(const std::string& name, F&& launch) {
        const bool timed = prof && (prof_only.empty() || prof_only == name);
        hipEvent_t a = nullptr;
        if (timed) {
            a = event();
            PMH_HIP(hipEventRecord(a, stream));
        }
        launch();
        if (timed) {
            hipEvent_t b = event();
            PMH_HIP(hipEventRecord(b, stream));
            pending_.push_back({name, a, b});
        }
    }
    std::map<std::string, std::pair<double, long long>> totals;

    // tuned launch shapes, "kind|shape" -> (stage, chunk)
    std::map<std::string, std::pair<int, int>> tuned;
    int tune_reps = 2;
    bool autotune = true;
    int sample_tiles = 0;  // largest timed sample of a launch (0: per-kind defaults)
    void load_tuning(const char* path);
    void save_tuning(const char* path) const;

    size_t graphs_cached() const { return graphs_.size(); }
    long long graph_launches = 0;

private:
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> spare_;
    hipEvent_t event();

    std::map<std::string, std::pair<void*, size_t>> pool_;
    int* queue_ = nullptr;  // tile-queue counters of persistent launches (Tiles::queue)

    struct Staging {
        void* host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
    };
    Staging ring_[4];
    int next_ = 0;
    const unsigned char* upload(const Plan& plan);

    struct GraphEntry {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::string, GraphEntry> graphs_;
    hipStream_t capture_ = nullptr;

    void enqueue(const Launch& l, const unsigned char* dev, hipStream_t s, bool timed);
    void tune(Launch& l, const unsigned char* dev);
    bool apply_tuning(Launch& l) const;
    std::string graph_key(const Plan& plan, const unsigned char* dev) const;
};

}  // namespace pmh
