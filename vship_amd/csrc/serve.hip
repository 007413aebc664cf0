// Continuous-batching SSIMULACRA2 server.
//
// Clips (sequences of device-resident frame pairs) are submitted at any time.
// A worker thread runs GPU steps back to back. Before each step it admits
// waiting clips in arrival order while their per-pair arena footprint fits
// the device-memory budget and a slot is free, then packs the next frames of
// every admitted clip round-robin into one uneven batch. Two engines on two
// streams alternate, so the next step is planned and enqueued while the
// previous one runs. A clip leaves as soon as its last score is back, freeing
// its budget and slot for the next waiting clip. Each pair's score is the
// same whatever batch it lands in (fixed tile decomposition, per-tile
// centring, fixed-order FP64 sums).
//
// No counterpart in the reference (its filter scores one frame per stream,
// src/ssimu2/vapoursynth.hpp:16-71).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/VshipSSIMU2Batch.h"
#include "api_internal.hpp"
#include "engine.hpp"

namespace vship {

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Clip {
    int id = 0;
    std::vector<PairInput> frames;
    std::vector<double> scores;
    int next = 0;      // next frame to schedule
    int done = 0;      // frames scored
    int64_t footprint = 0;  // arena bytes one in-flight pair of this clip needs
    double submit_ms = 0, admit_ms = 0, finish_ms = 0;
    int error = 0;
    bool finished = false;
};

struct Slot {  // one frame of one clip inside a step
    std::shared_ptr<Clip> clip;
    int frame;
};

}  // namespace

class Server {
public:
    explicit Server(const Vship_SSIMU2ServerConfig& c, int device) : cfg_(c), device_(device) {
        if (cfg_.max_pairs_per_step < 1 || cfg_.max_width < 1 || cfg_.max_height < 1 || cfg_.memory_budget == 0)
            VSHIP_THROW(E_BadPointer);
        if (cfg_.max_inflight_clips < 1) cfg_.max_inflight_clips = cfg_.max_pairs_per_step;
        EngineOptions o;
        o.device = device;
        o.max_pairs = cfg_.max_pairs_per_step;
        o.fp16 = cfg_.fp16_planes != 0;
        o.map_th = cfg_.map_tile_height == 64 || cfg_.map_tile_height == 112 || cfg_.map_tile_height == kMapMfma
                       ? cfg_.map_tile_height
                       : default_map_th(cfg_.max_width, cfg_.max_height);
        // the budget covers the arenas of both in-flight steps
        o.arena_bytes = (int64_t)(cfg_.memory_budget / 2);
        o.max_map_tiles = (int64_t)cfg_.max_pairs_per_step * map_tiles(cfg_.max_width, cfg_.max_height, 32);
        o.persistent_grid = cfg_.persistent_grid;
        o.use_graph = cfg_.use_graph != 0;
        step_budget_ = o.arena_bytes;
        fp16_ = o.fp16;
        for (int i = 0; i < 2; i++) eng_[i] = std::make_unique<Engine>(o);
        worker_ = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
    }

    int submit(const Vship_SSIMU2Pair* frames, int n, double arrival_ms) {
        if (!frames || n < 1) VSHIP_THROW(E_BadPointer);
        auto c = std::make_shared<Clip>();
        c->frames.resize(n);
        for (int i = 0; i < n; i++) {
            const Vship_SSIMU2Pair& p = frames[i];
            PairInput& in = c->frames[i];
            for (int k = 0; k < 3; k++) {
                in.src[0][k] = p.ref[k];
                in.src[1][k] = p.dist[k];
                in.stride[0][k] = p.ref_stride[k];
                in.stride[1][k] = p.dist_stride[k];
            }
            in.w = p.width;
            in.h = p.height;
            in.fmt = p.format;
            in.bits = p.bits;
            in.matrix = p.matrix;
            in.chroma_loc = p.chroma_location;
            if (p.width < 1 || p.height < 1 || p.width > cfg_.max_width || p.height > cfg_.max_height)
                VSHIP_THROW(E_BadPointer);
            c->footprint = std::max<int64_t>(c->footprint, pair_bytes(p.width, p.height, fp16_));
        }
        if (c->footprint > step_budget_) VSHIP_THROW(E_OutOfVRAM);
        c->scores.assign(n, 0.0);
        c->submit_ms = arrival_ms > 0 ? arrival_ms : now_ms();
        std::lock_guard<std::mutex> g(mu_);
        c->id = next_id_++;
        clips_[c->id] = c;
        waiting_.push_back(c);
        cv_.notify_all();
        return c->id;
    }

    // 1 when finished (scores copied), 0 when not yet; throws for unknown ids
    int poll(int id, double* scores, double* latency_ms, bool block) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = clips_.find(id);
        if (it == clips_.end()) VSHIP_THROW(E_BadHandler);
        std::shared_ptr<Clip> c = it->second;
        if (block) done_cv_.wait(lk, [&] { return c->finished; });
        if (!c->finished) return 0;
        if (c->error) {
            const int e = c->error;
            clips_.erase(it);
            throw VshipErr((ErrCode)e, __FILE__, __LINE__);
        }
        if (scores) std::memcpy(scores, c->scores.data(), sizeof(double) * c->scores.size());
        if (latency_ms) *latency_ms = c->finish_ms - c->submit_ms;
        clips_.erase(it);
        return 1;
    }

    void stats(Vship_SSIMU2ServerStats* s) {
        std::lock_guard<std::mutex> g(mu_);
        *s = stats_;
        s->waiting_clips = (int)waiting_.size();
        s->active_clips = (int)active_.size();
    }

private:
    // Pack the next step: admit, then take frames round-robin (caller holds mu_).
    void build_step(std::vector<Slot>& slots) {
        slots.clear();
        int64_t used = 0;
        // admission: arrival order, footprint within the step budget, free slot
        while (!waiting_.empty() && (int)active_.size() < cfg_.max_inflight_clips) {
            std::shared_ptr<Clip> c = waiting_.front();
            int64_t need = c->footprint;
            for (auto& a : active_) need += a->footprint;
            if (need > step_budget_ && !active_.empty()) break;
            waiting_.pop_front();
            c->admit_ms = now_ms();
            active_.push_back(c);
        }
        // round-robin packing over the admitted clips with frames left to schedule
        bool progress = true;
        while (progress && (int)slots.size() < cfg_.max_pairs_per_step) {
            progress = false;
            for (auto& c : active_) {
                if ((int)slots.size() >= cfg_.max_pairs_per_step) break;
                if (c->next >= (int)c->frames.size()) continue;
                if (used + c->footprint > step_budget_) continue;
                slots.push_back({c, c->next++});
                used += c->footprint;
                progress = true;
            }
        }
    }

    void finish_step(std::vector<Slot>& slots, const std::vector<double>& scores, int err) {
        const double t = now_ms();
        std::lock_guard<std::mutex> g(mu_);
        for (size_t i = 0; i < slots.size(); i++) {
            Clip& c = *slots[i].clip;
            if (err) c.error = err;
            else c.scores[slots[i].frame] = scores[i];
            c.done++;
        }
        for (auto it = active_.begin(); it != active_.end();) {
            Clip& c = **it;
            if (c.done == (int)c.frames.size()) {
                c.finish_ms = t;
                c.finished = true;
                stats_.clips_done++;
                it = active_.erase(it);
            } else {
                ++it;
            }
        }
        stats_.pairs_done += (int64_t)slots.size();
        stats_.steps++;
        done_cv_.notify_all();
    }

    void loop() {
        if (hipSetDevice(device_) != hipSuccess) return;
        std::vector<Slot> slots[2];
        std::vector<PairInput> in;
        std::vector<double> sc;
        bool pending[2] = {false, false};
        int cur = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                if (!pending[0] && !pending[1])
                    cv_.wait(lk, [&] { return stop_ || !waiting_.empty() || has_work(); });
                if (stop_ && !pending[0] && !pending[1] && waiting_.empty() && !has_work()) break;
                build_step(slots[cur]);
            }
            if (!slots[cur].empty()) {
                in.resize(slots[cur].size());
                for (size_t i = 0; i < slots[cur].size(); i++) in[i] = slots[cur][i].clip->frames[slots[cur][i].frame];
                const double t0 = now_ms();
                try {
                    eng_[cur]->submit(in.data(), (int)in.size());
                    pending[cur] = true;
                } catch (const VshipErr& e) {
                    finish_step(slots[cur], sc, e.code);
                }
                std::lock_guard<std::mutex> g(mu_);
                stats_.host_ms += now_ms() - t0;
            }
            // retire the other engine's step while this one runs
            const int other = cur ^ 1;
            if (pending[other]) {
                sc.assign(slots[other].size(), 0.0);
                int err = 0;
                try {
                    eng_[other]->wait(sc.data(), nullptr);
                } catch (const VshipErr& e) {
                    err = e.code;
                }
                pending[other] = false;
                finish_step(slots[other], sc, err);
            }
            cur = other;
        }
        // drain
        for (int k = 0; k < 2; k++)
            if (pending[k]) {
                sc.assign(slots[k].size(), 0.0);
                int err = 0;
                try {
                    eng_[k]->wait(sc.data(), nullptr);
                } catch (const VshipErr& e) {
                    err = e.code;
                }
                finish_step(slots[k], sc, err);
            }
    }

    bool has_work() const {
        for (auto& c : active_)
            if (c->next < (int)c->frames.size()) return true;
        return false;
    }

    Vship_SSIMU2ServerConfig cfg_;
    int device_;
    int64_t step_budget_ = 0;
    bool fp16_ = false;
    std::unique_ptr<Engine> eng_[2];
    std::thread worker_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    bool stop_ = false;
    int next_id_ = 1;
    std::map<int, std::shared_ptr<Clip>> clips_;
    std::deque<std::shared_ptr<Clip>> waiting_;
    std::vector<std::shared_ptr<Clip>> active_;
    Vship_SSIMU2ServerStats stats_{};
};

namespace {

std::mutex g_srv_mu;
std::map<int, std::shared_ptr<Server>> g_servers;
int g_srv_next = 1;

std::shared_ptr<Server> get_server(int id) {
    std::lock_guard<std::mutex> g(g_srv_mu);
    auto it = g_servers.find(id);
    return it == g_servers.end() ? nullptr : it->second;
}

template <typename F>
Vship_Exception guard(F&& fn) {
    try {
        fn();
    } catch (const VshipErr& e) {
        return (Vship_Exception)e.code;
    } catch (const std::bad_alloc&) {
        return Vship_OutOfRAM;
    } catch (...) {
        return Vship_BadErrorType;
    }
    return Vship_NoError;
}

}  // namespace

}  // namespace vship

using namespace vship;

extern "C" {

Vship_Exception Vship_SSIMU2ServerInit(Vship_SSIMU2ServerHandler* handler, const Vship_SSIMU2ServerConfig* cfg) {
    if (!handler || !cfg) return Vship_BadPointer;
    return guard([&] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
        auto s = std::make_shared<Server>(*cfg, dev);
        std::lock_guard<std::mutex> g(g_srv_mu);
        handler->id = g_srv_next++;
        g_servers[handler->id] = s;
    });
}

Vship_Exception Vship_SSIMU2ServerFree(Vship_SSIMU2ServerHandler handler) {
    std::shared_ptr<Server> s;
    {
        std::lock_guard<std::mutex> g(g_srv_mu);
        auto it = g_servers.find(handler.id);
        if (it == g_servers.end()) return Vship_BadHandler;
        s = it->second;
        g_servers.erase(it);
    }
    s.reset();  // joins the worker after the in-flight steps drain
    return Vship_NoError;
}

Vship_Exception Vship_SSIMU2ServerSubmitClip(Vship_SSIMU2ServerHandler handler, const Vship_SSIMU2Pair* frames,
                                             int nframes, double arrival_ms, int* clip_id) {
    std::shared_ptr<Server> s = get_server(handler.id);
    if (!s) return Vship_BadHandler;
    if (!clip_id) return Vship_BadPointer;
    return guard([&] { *clip_id = s->submit(frames, nframes, arrival_ms); });
}

Vship_Exception Vship_SSIMU2ServerPollClip(Vship_SSIMU2ServerHandler handler, int clip_id, int* done,
                                           double* scores, double* latency_ms) {
    std::shared_ptr<Server> s = get_server(handler.id);
    if (!s) return Vship_BadHandler;
    if (!done) return Vship_BadPointer;
    return guard([&] { *done = s->poll(clip_id, scores, latency_ms, false); });
}

Vship_Exception Vship_SSIMU2ServerWaitClip(Vship_SSIMU2ServerHandler handler, int clip_id, double* scores,
                                           double* latency_ms) {
    std::shared_ptr<Server> s = get_server(handler.id);
    if (!s) return Vship_BadHandler;
    return guard([&] { s->poll(clip_id, scores, latency_ms, true); });
}

Vship_Exception Vship_SSIMU2ServerGetStats(Vship_SSIMU2ServerHandler handler, Vship_SSIMU2ServerStats* stats) {
    std::shared_ptr<Server> s = get_server(handler.id);
    if (!s) return Vship_BadHandler;
    if (!stats) return Vship_BadPointer;
    s->stats(stats);
    return Vship_NoError;
}

double Vship_SSIMU2ServerClockMs(void) { return now_ms(); }

}  // extern "C"
