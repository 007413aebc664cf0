// Autotuner of the batch launch shape: map tile height, persistent grid size
// and graph capture, measured on a caller-supplied sample batch. Results can
// be kept in a text table (one line per frame-size class) that
// Vship_SSIMU2BatchInit consults when the configuration leaves the shape at 0.
//
// No counterpart in the reference (its launch shapes are fixed,
// src/ssimu2/score.hpp:211-217).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/VshipSSIMU2Batch.h"
#include "api_internal.hpp"
#include "engine.hpp"

namespace vship {

namespace {

std::mutex g_tune_mu;
std::vector<TuneEntry> g_table;

double wall_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int size_class(int w, int h) {
    const int64_t p = (int64_t)w * h;
    if (p <= 640 * 360) return 0;
    if (p <= 1280 * 720) return 1;
    if (p <= 1920 * 1080) return 2;
    if (p <= 2560 * 1440) return 3;
    return 4;
}

bool tuned_shape(int w, int h, bool fp16, int* th, int* grid, int* graph) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    const int c = size_class(w, h);
    for (const TuneEntry& e : g_table)
        if (e.size_class == c && e.fp16 == (int)fp16) {
            *th = e.th;
            *grid = e.grid;
            *graph = e.graph;
            return true;
        }
    return false;
}

void record_shape(const TuneEntry& e) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (TuneEntry& x : g_table)
        if (x.size_class == e.size_class && x.fp16 == e.fp16) {
            x = e;
            return;
        }
    g_table.push_back(e);
}

}  // namespace vship

using namespace vship;

extern "C" {

Vship_Exception Vship_SSIMU2BatchTune(const Vship_SSIMU2Pair* pairs, int npairs, int fp16_planes, int repeats,
                                      Vship_SSIMU2TuneResult* result) {
    if (!pairs || npairs < 1 || !result) return Vship_BadPointer;
    try {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            VSHIP_THROW(E_BadDeviceCode);
        std::vector<PairInput> in(npairs);
        int mw = 1, mh = 1;
        int64_t arena = 0;
        for (int i = 0; i < npairs; i++) {
            const Vship_SSIMU2Pair& p = pairs[i];
            PairInput& x = in[i];
            for (int k = 0; k < 3; k++) {
                x.src[0][k] = p.ref[k];
                x.src[1][k] = p.dist[k];
                x.stride[0][k] = p.ref_stride[k];
                x.stride[1][k] = p.dist_stride[k];
            }
            x.w = p.width;
            x.h = p.height;
            x.fmt = p.format;
            x.bits = p.bits;
            x.matrix = p.matrix;
            x.chroma_loc = p.chroma_location;
            mw = std::max(mw, p.width);
            mh = std::max(mh, p.height);
            arena += pair_bytes(p.width, p.height, fp16_planes != 0);
        }
        const int ths[4] = {kMapMfma, 48, 64, 112};
        const int grids[4] = {0, 2 * cus, 4 * cus, 8 * cus};
        std::memset(result, 0, sizeof(*result));
        result->best_ms = 1e30f;
        int k = 0;
        for (int th : ths)
            for (int gi = 0; gi < 4; gi++)
                for (int graph = 0; graph < 2; graph++) {
                    if (graph && grids[gi] == 0) continue;  // a captured graph needs a fixed grid
                    if (k >= VSHIP_TUNE_MAX_CANDIDATES) break;
                    EngineOptions o;
                    o.device = dev;
                    o.max_pairs = npairs;
                    o.fp16 = fp16_planes != 0;
                    o.map_th = th;
                    o.arena_bytes = arena;
                    o.max_map_tiles = (int64_t)npairs * map_tiles(mw, mh, 32);  // bound for every shape
                    o.persistent_grid = grids[gi];
                    o.use_graph = graph != 0;
                    Engine e(o);
                    e.submit(in.data(), npairs);  // warm-up (and graph capture)
                    e.wait(nullptr, nullptr);
                    float best = 1e30f;
                    for (int r = 0; r < std::max(1, repeats); r++) {
                        const double t0 = wall_ms();
                        e.submit(in.data(), npairs);
                        e.wait(nullptr, nullptr);
                        best = std::min(best, (float)(wall_ms() - t0));
                    }
                    Vship_SSIMU2TuneCandidate& c = result->candidates[k++];
                    c.map_tile_height = th;
                    c.persistent_grid = grids[gi];
                    c.use_graph = graph;
                    c.ms = best;
                    if (best < result->best_ms) {
                        result->best_ms = best;
                        result->map_tile_height = th;
                        result->persistent_grid = grids[gi];
                        result->use_graph = graph;
                    }
                }
        result->ncandidates = k;
        TuneEntry te;
        te.size_class = size_class(mw, mh);
        te.fp16 = fp16_planes != 0;
        te.th = result->map_tile_height;
        te.grid = result->persistent_grid;
        te.graph = result->use_graph;
        te.ms_per_pair = result->best_ms / npairs;
        record_shape(te);
    } catch (const VshipErr& e) {
        return (Vship_Exception)e.code;
    } catch (...) {
        return Vship_BadErrorType;
    }
    return Vship_NoError;
}

Vship_Exception Vship_SSIMU2TuneSave(const char* path) {
    if (!path) return Vship_BadPointer;
    FILE* f = std::fopen(path, "w");
    if (!f) return Vship_BadPointer;
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::fprintf(f, "# size_class fp16 map_tile_height persistent_grid use_graph ms_per_pair\n");
    for (const TuneEntry& e : g_table)
        std::fprintf(f, "%d %d %d %d %d %.6f\n", e.size_class, e.fp16, e.th, e.grid, e.graph, e.ms_per_pair);
    std::fclose(f);
    return Vship_NoError;
}

Vship_Exception Vship_SSIMU2TuneLoad(const char* path) {
    if (!path) return Vship_BadPointer;
    FILE* f = std::fopen(path, "r");
    if (!f) return Vship_BadPointer;
    char line[256];
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        TuneEntry e;
        if (std::sscanf(line, "%d %d %d %d %d %lf", &e.size_class, &e.fp16, &e.th, &e.grid, &e.graph,
                        &e.ms_per_pair) == 6) {
            if (e.th != kMapMfma && e.th != 48 && e.th != 64 && e.th != 112) continue;
            record_shape(e);
        }
    }
    std::fclose(f);
    return Vship_NoError;
}

}  // extern "C"
