// pmh_engine.cpp -- plan execution, graph capture and the autotuner.
#include "pmh_engine.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace pmh {

// ------------------------------------------------------------------- Plan ----
size_t Plan::blob(const void* p, size_t bytes) {
    std::string key(static_cast<const char*>(p), bytes);
    auto it = seen_.find(key);
    if (it != seen_.end()) return it->second;
    const size_t off = (arena.size() + 15) & ~(size_t)15;
    arena.resize(off + bytes);
    if (bytes) std::memcpy(arena.data() + off, p, bytes);
    seen_.emplace(std::move(key), off);
    return off;
}

size_t Plan::prefix(const std::vector<int>& counts, int& total) {
    std::vector<int> pre(counts.size() + 1, 0);
    long long t = 0;
    for (size_t i = 0; i < counts.size(); ++i) {
        pre[i] = (int)t;
        t += counts[i];
    }
    if (t > 0x3fffffff) throw ArgError("batch too large: more than 2^30 tiles in one launch");
    pre[counts.size()] = (int)t;
    total = (int)t;
    return blob(pre.data(), pre.size() * sizeof(int));
}

Launch& Plan::add(Kern k, size_t params, size_t prefix, int items, int total, const char* name) {
    Launch l;
    l.kind = k;
    l.params = params;
    l.prefix = prefix;
    l.items = items;
    l.total = total;
    l.name = name;
    launches.push_back(l);
    return launches.back();
}

// ----------------------------------------------------------------- Engine ----
Engine::~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& kv : graphs_) {
        (void)hipGraphExecDestroy(kv.second.exec);
        (void)hipGraphDestroy(kv.second.g);
    }
    if (capture_) (void)hipStreamDestroy(capture_);
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : spare_) (void)hipEventDestroy(e);
    for (auto& s : ring_) {
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.host) (void)hipHostFree(s.host);
    }
    for (auto& kv : pool_) (void)hipFree(kv.second.first);
}

void* Engine::buf(const std::string& role, size_t bytes) {
    auto it = pool_.find(role);
    if (it != pool_.end() && it->second.second >= bytes) return it->second.first;
    if (it != pool_.end()) {
        PMH_HIP(hipStreamSynchronize(stream));
        PMH_HIP(hipFree(it->second.first));
        pool_.erase(it);
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw HipError("hipMalloc(" + std::to_string(bytes) + " B) for " + role + " failed",
                       (int)e);
    }
    pool_[role] = {p, bytes};
    return p;
}

unsigned long long* Engine::cand_counter() {
    const bool fresh = pool_.find("profile.candidates") == pool_.end();
    auto* p = static_cast<unsigned long long*>(buf("profile.candidates", 8));
    if (fresh) PMH_HIP(hipMemsetAsync(p, 0, 8, stream));
    return p;
}

hipEvent_t Engine::event() {
    if (!spare_.empty()) {
        hipEvent_t e = spare_.back();
        spare_.pop_back();
        return e;
    }
    hipEvent_t e;
    PMH_HIP(hipEventCreate(&e));
    return e;
}

void Engine::resolve() {
    if (pending_.empty()) return;
    PMH_HIP(hipStreamSynchronize(stream));
    for (auto& p : pending_) {
        float ms = 0.0f;
        PMH_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto& t = totals[p.name];
        t.first += ms;
        t.second += 1;
        spare_.push_back(p.a);
        spare_.push_back(p.b);
    }
    pending_.clear();
}

// The arena goes through a ring of pinned staging buffers: a slot is reused
// only after the copy that last read it has completed, so the host never
// waits on the kernels of the call before.
const unsigned char* Engine::upload(const Plan& plan) {
    const size_t bytes = std::max<size_t>(plan.arena.size(), 16);
    auto* dev = static_cast<unsigned char*>(buf("arena", bytes));
    Staging& s = ring_[next_];
    next_ = (next_ + 1) & 3;
    if (s.done) PMH_HIP(hipEventSynchronize(s.done));
    else PMH_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    if (s.cap < bytes) {
        if (s.host) PMH_HIP(hipHostFree(s.host));
        s.host = nullptr;
        s.cap = 0;
        PMH_HIP(hipHostMalloc(&s.host, bytes, hipHostMallocDefault));
        s.cap = bytes;
    }
    std::memcpy(s.host, plan.arena.data(), plan.arena.size());
    PMH_HIP(hipMemcpyAsync(dev, s.host, bytes, hipMemcpyHostToDevice, stream));
    PMH_HIP(hipEventRecord(s.done, stream));
    return dev;
}

void Engine::enqueue(const Launch& l, const unsigned char* dev, hipStream_t s, bool timed) {
    timed = timed && (prof_only.empty() || prof_only == l.name);
    hipEvent_t a = nullptr;
    if (timed) {
        a = event();
        PMH_HIP(hipEventRecord(a, s));
    }
    const Tiles T{reinterpret_cast<const int*>(dev + l.prefix), l.items, l.total, l.chunk,
                  l.stride, queue_};
    const void* pp = dev + l.params;
    switch (l.kind) {
    case Kern::Prep: launch_prep(static_cast<const PrepParams*>(pp), T, s); break;
    case Kern::Zero: launch_zero(static_cast<const ZeroParams*>(pp), T, s); break;
    case Kern::Init: launch_init(static_cast<const InitParams*>(pp), T, l.F, l.P, l.nc, s); break;
    case Kern::Propagate:
        launch_propagate(static_cast<const PropParams*>(pp), T, l.stage, l.nc, s);
        break;
    case Kern::Viewbuf: launch_viewbuf(static_cast<const ViewParams*>(pp), T, s); break;
    case Kern::Consistency:
        launch_consistency(static_cast<const ConsParams*>(pp), T, l.f0, l.f1, s);
        break;
    case Kern::FillState: launch_fill_state(static_cast<const FillParams*>(pp), T, s); break;
    case Kern::PdStep:
        launch_pd_step(static_cast<const PdParams*>(pp), T, l.f0, l.i0 != 0, s);
        break;
    case Kern::PdSteps:
        launch_pd_steps(static_cast<const PdParams*>(pp), T, l.f0, l.i0 != 0, l.i1, s);
        break;
    case Kern::Occmap:
        launch_occmap(static_cast<const ConsParams*>(pp), T, l.f0, l.f1, s);
        break;
    case Kern::NormalDepth: launch_normal_depth(static_cast<const FillParams*>(pp), T, s); break;
    case Kern::FillDepth:
        launch_fill_depth(static_cast<const FillDepthParams*>(pp), T, l.i0 != 0, s);
        break;
    case Kern::PyrDown: launch_pyr_down(static_cast<const PyrParams*>(pp), T, s); break;
    case Kern::PlaneUp: launch_plane_up(static_cast<const UpParams*>(pp), T, s); break;
    case Kern::GradRefine:
        launch_grad_refine(static_cast<const RefineParams*>(pp), T, l.F, l.P, l.f0, l.i0, l.f1,
                           l.nc, s);
        break;
    }
    if (timed) {
        hipEvent_t b = event();
        PMH_HIP(hipEventRecord(b, s));
        pending_.push_back({l.name, a, b});
    }
}

bool Engine::apply_tuning(Launch& l) const {
    if (!l.tunable) return true;
    auto it = tuned.find(l.tune_key);
    if (it != tuned.end()) {
        l.stage = it->second.first;
        l.chunk = it->second.second;
        return true;
    }
    return !autotune;  // defaults stand
}

// A candidate must beat the best so far by this factor to replace it.
constexpr double kTuneMargin = 0.97;

// Times every candidate shape of one launch on its own inputs and keeps the
// fastest.  The candidates compute identical results and the tunable kernels
// read and write disjoint buffers, so repeating a launch is harmless.  A
// bounded sample of the work list, every stride-th tile, keeps tuning cheap on
// very large batches and still sees every part of it (pairs of every size).
void Engine::tune(Launch& l, const unsigned char* dev) {
    // the default shape first: another candidate replaces it only when
    // clearly faster (timing noise must not flip the choice)
    std::vector<std::pair<int, int>> cand{{l.stage, l.chunk}};
    if (l.kind == Kern::Propagate) {  // more tiles per block never won a trial
        for (int st : kStageChoices)  // the block-stage kernels are gray only
            if (st != propagate_stage_default(l.nc) &&
                (l.nc == kGray || (st != kStageBlock && st != kStageBlock3)))
                cand.push_back({st, 1});
        cand.push_back({l.stage, l.chunk == 0 ? 1 : 0});  // the default shape, the other mode
    } else {
        for (int ch : {2, 4, 8, 0}) cand.push_back({l.stage, ch});  // 0: persistent blocks
    }
    Launch t = l;
    const int sample =
        sample_tiles > 0 ? sample_tiles : (l.kind == Kern::Propagate ? 16384 : 65536);
    t.total = std::min(l.total, sample);
    t.stride = std::max(1, l.total / t.total);  // (total - 1) * stride < l.total
    // A launch much bigger than the sample runs many rounds of blocks, the
    // sample only a few, and shapes with fewer waves per tile lose more of a
    // few rounds to the last one: there the score is the marginal time,
    // time(sample) - time(half the sample), which cancels the launch's tail
    // and fixed costs.
    const bool marginal = l.total >= 2 * (long long)sample;
    Launch th = t;
    if (marginal) {  // t.total = sample >= 8, so th.total = sample / 2 > 0
        th.total = t.total / 2;
        th.stride = std::max(1, l.total / th.total);
    }
    // every sampled tile lies in the launch's work list
    if (t.total < 1 || (long long)(t.total - 1) * t.stride >= l.total ||
        (marginal && (th.total < 1 || (long long)(th.total - 1) * th.stride >= l.total)))
        throw ArgError("autotuner sample outside the work list");
    hipEvent_t a = event(), b = event();
    auto time_ms = [&](const Launch& x) {
        PMH_HIP(hipEventRecord(a, stream));
        enqueue(x, dev, stream, false);
        PMH_HIP(hipEventRecord(b, stream));
        PMH_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        PMH_HIP(hipEventElapsedTime(&ms, a, b));
        return (double)ms;
    };
    // Two passes over all candidates, each keeping its faster pass: a clock
    // or cache state that drifts while the first candidates run cannot
    // favour the later ones.
    std::vector<double> ms_min(cand.size(), 1e30);
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t ci = 0; ci < cand.size(); ++ci) {
            t.stage = th.stage = cand[ci].first;
            t.chunk = th.chunk = cand[ci].second;
            enqueue(t, dev, stream, false);  // warm-up (instruction cache, L2)
            double ms_sum = 0.0;
            for (int r = 0; r < tune_reps; ++r) {
                ms_sum += time_ms(t);
                if (marginal) ms_sum -= time_ms(th);
            }
            ms_min[ci] = std::min(ms_min[ci], ms_sum);
        }
    }
    double best = ms_min[0];
    std::pair<int, int> pick = cand[0];
    for (size_t ci = 1; ci < cand.size(); ++ci)
        if (ms_min[ci] < kTuneMargin * best) {
            best = ms_min[ci];
            pick = cand[ci];
        }
    spare_.push_back(a);
    spare_.push_back(b);
    tuned[l.tune_key] = pick;
    l.stage = pick.first;
    l.chunk = pick.second;
}

std::string Engine::graph_key(const Plan& plan, const unsigned char* dev) const {
    std::string k;
    auto put = [&](const void* p, size_t n) { k.append(static_cast<const char*>(p), n); };
    put(&dev, sizeof(dev));
    for (const Launch& l : plan.launches) {
        put(&l.kind, sizeof(l.kind));
        put(&l.params, sizeof(l.params));
        put(&l.prefix, sizeof(l.prefix));
        put(&l.items, sizeof(l.items));
        put(&l.total, sizeof(l.total));
        put(&l.f0, sizeof(l.f0));
        put(&l.f1, sizeof(l.f1));
        put(&l.i0, sizeof(l.i0));
        put(&l.i1, sizeof(l.i1));
        put(&l.P, sizeof(l.P));
        put(&l.nc, sizeof(l.nc));
        put(&l.F, sizeof(l.F));
        put(&l.stage, sizeof(l.stage));
        put(&l.chunk, sizeof(l.chunk));
        put(&l.stride, sizeof(l.stride));
    }
    return k;
}

void Engine::run(Plan& plan, bool graph) {
    if (!queue_) {  // counters of persistent launches, cleared once (never inside a capture)
        queue_ = get<int>("tiles.queue", 64);
        PMH_HIP(hipMemsetAsync(queue_, 0, 64 * sizeof(int), stream));
    }
    const unsigned char* dev = upload(plan);
    bool ready = true;
    for (Launch& l : plan.launches) ready = apply_tuning(l) && ready;
    if (!graph || !ready) {
        for (Launch& l : plan.launches) {
            if (!apply_tuning(l)) tune(l, dev);
            enqueue(l, dev, stream, prof);
        }
        return;
    }
    const std::string key = graph_key(plan, dev);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        if (graphs_.size() >= 16) {  // bounded cache: drop every graph
            PMH_HIP(hipStreamSynchronize(stream));
            for (auto& kv : graphs_) {
                (void)hipGraphExecDestroy(kv.second.exec);
                (void)hipGraphDestroy(kv.second.g);
            }
            graphs_.clear();
        }
        if (!capture_) PMH_HIP(hipStreamCreateWithFlags(&capture_, hipStreamNonBlocking));
        GraphEntry ge;
        PMH_HIP(hipStreamBeginCapture(capture_, hipStreamCaptureModeThreadLocal));
        try {
            for (const Launch& l : plan.launches) enqueue(l, dev, capture_, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(capture_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        PMH_HIP(hipStreamEndCapture(capture_, &ge.g));
        hipError_t e = hipGraphInstantiate(&ge.exec, ge.g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(ge.g);
            PMH_HIP(e);
        }
        it = graphs_.emplace(key, ge).first;
    }
    hipEvent_t a = nullptr;
    if (prof) {
        a = event();
        PMH_HIP(hipEventRecord(a, stream));
    }
    PMH_HIP(hipGraphLaunch(it->second.exec, stream));
    ++graph_launches;
    if (prof) {
        hipEvent_t b = event();
        PMH_HIP(hipEventRecord(b, stream));
        pending_.push_back({"graph", a, b});
    }
}

// Tuning cache file: one "key stage chunk" line per shape class.
void Engine::load_tuning(const char* path) {
    std::ifstream f(path);
    if (!f) throw ArgError(std::string("cannot read tuning file ") + path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string key;
        int st = 0, ch = 0;
        if (!(ls >> key >> st >> ch)) continue;
        if ((st != 0 && std::find(std::begin(kStageChoices), std::end(kStageChoices), st) ==
                            std::end(kStageChoices)) ||
            ch < 0 || ch > 64)  // chunk 0: persistent blocks
            throw ArgError("bad tuning entry: " + line);
        tuned[key] = {st, ch};
    }
}

void Engine::save_tuning(const char* path) const {
    std::ofstream f(path);
    if (!f) throw ArgError(std::string("cannot write tuning file ") + path);
    for (auto& kv : tuned) f << kv.first << ' ' << kv.second.first << ' ' << kv.second.second << '\n';
}

}  // namespace pmh
