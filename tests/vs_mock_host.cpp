// A minimal VapourSynth host for testing the vship plugin without
// VapourSynth: it loads libvship.so, calls VapourSynthPluginInit2 with a
// plugin API that records what is registered, and drives the registered
// functions through a VSAPI table whose used entries are implemented here
// (maps, nodes of in-memory 32-bit float RGB frames, frame contexts). Entries
// the plugin is not expected to use abort with their slot index.
//
// Built by vship_amd/build.py into tests/_build/libvs_mock_host.so and
// driven from tests/test_vs_filter.py (ctypes).
#include <dlfcn.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../vship_amd/csrc/vs4.hpp"

namespace vs4 {

struct Value {
    int64_t i = 0;
    double f = 0;
    std::string data;
    Node* node = nullptr;
    int kind = 0;  // 1 int, 2 float, 3 data, 5 node
};
struct Map {
    std::map<std::string, std::vector<Value>> kv;
    std::string error;
};
struct Frame {
    int w = 0, h = 0;
    ptrdiff_t stride = 0;
    const uint8_t* planes[3] = {nullptr, nullptr, nullptr};
    std::vector<uint8_t> own;
    Map props;
    bool output = false;
    bool shared = false;  // planes belong to the frame this one was copied from
};
struct Node {
    VideoInfo vi{};
    std::vector<Frame*> frames;  // source clip
    FilterGetFrame get = nullptr;
    FilterFree free = nullptr;
    void* inst = nullptr;
    std::atomic<int> freed{0};
    std::atomic<int> gets{0};  // references handed out by mapGetNode
};
struct FrameContext {
    std::vector<std::pair<int, Node*>> requested;
    std::string error;
};
struct Core {
    int threads = 4;
};
struct Plugin {
    std::string id, ns, name;
    int version = 0, api = 0;
    struct Fn {
        std::string name, args, ret;
        PublicFunction fn;
    };
    std::vector<Fn> fns;
};

}  // namespace vs4

using namespace vs4;

namespace {

std::atomic<int> g_output_frames{0};

// --- the VSAPI entries the filter uses -----------------------------------
void m_createVideoFilter(Map* out, const char*, const VideoInfo* vi, FilterGetFrame getFrame, FilterFree free,
                         int, const FilterDependency*, int, void* inst, Core*) {
    Node* n = new Node;
    n->vi = *vi;
    n->get = getFrame;
    n->free = free;
    n->inst = inst;
    Value v;
    v.kind = 5;
    v.node = n;
    out->kv["clip"] = {v};
}
void m_freeNode(Node* n) {
    if (n) n->freed++;
}
const VideoInfo* m_getVideoInfo(Node* n) { return &n->vi; }
void m_freeFrame(const Frame* f) {
    if (f && f->output) {
        delete f;
        g_output_frames--;
    }
}
// As in VapourSynth, a copied frame shares the source's planes until one of
// them is written (m_getWritePtr copies them then).
Frame* m_copyFrame(const Frame* f, Core*) {
    Frame* c = new Frame;
    c->w = f->w;
    c->h = f->h;
    c->stride = f->stride;
    for (int p = 0; p < 3; p++) c->planes[p] = f->planes[p];
    c->shared = true;
    c->props = f->props;
    c->output = true;
    g_output_frames++;
    return c;
}
Frame* m_newVideoFrame(const VideoFormat* fmt, int w, int h, const Frame* prop_src, Core*) {
    Frame* c = new Frame;
    c->w = w;
    c->h = h;
    c->stride = (ptrdiff_t)w * fmt->bytesPerSample;
    c->own.assign((size_t)c->stride * h * fmt->numPlanes, 0);
    for (int p = 0; p < fmt->numPlanes && p < 3; p++) c->planes[p] = c->own.data() + (size_t)p * c->stride * h;
    if (prop_src) c->props = prop_src->props;
    c->output = true;
    g_output_frames++;
    return c;
}
uint8_t* m_getWritePtr(Frame* f, int p) {
    if (f->shared) {
        f->own.resize((size_t)f->stride * f->h * 3);
        for (int q = 0; q < 3; q++) {
            std::memcpy(f->own.data() + (size_t)q * f->stride * f->h, f->planes[q], (size_t)f->stride * f->h);
            f->planes[q] = f->own.data() + (size_t)q * f->stride * f->h;
        }
        f->shared = false;
    }
    return const_cast<uint8_t*>(f->planes[p]);
}
int m_queryVideoFormat(VideoFormat* f, int family, int sample, int bits, int ssw, int ssh, Core*) {
    f->colorFamily = family;
    f->sampleType = sample;
    f->bitsPerSample = bits;
    f->bytesPerSample = (bits + 7) / 8;
    f->subSamplingW = ssw;
    f->subSamplingH = ssh;
    f->numPlanes = family == cfGray ? 1 : 3;
    return 1;
}
Map* m_getFramePropertiesRW(Frame* f) { return &f->props; }
ptrdiff_t m_getStride(const Frame* f, int) { return f->stride; }
const uint8_t* m_getReadPtr(const Frame* f, int p) { return f->planes[p]; }
int m_getFrameWidth(const Frame* f, int) { return f->w; }
int m_getFrameHeight(const Frame* f, int) { return f->h; }
const Frame* m_getFrameFilter(int n, Node* node, FrameContext*) {
    return (n >= 0 && n < (int)node->frames.size()) ? node->frames[n] : nullptr;
}
void m_requestFrameFilter(int n, Node* node, FrameContext* ctx) { ctx->requested.push_back({n, node}); }
void m_setFilterError(const char* msg, FrameContext* ctx) { ctx->error = msg; }
Map* m_createMap() { return new Map; }
void m_freeMap(Map* m) { delete m; }
void m_mapSetError(Map* m, const char* msg) { m->error = msg; }
const char* m_mapGetError(const Map* m) { return m->error.empty() ? nullptr : m->error.c_str(); }
int64_t m_mapGetInt(const Map* m, const char* key, int index, int* err) {
    auto it = m->kv.find(key);
    if (it == m->kv.end() || index >= (int)it->second.size() || it->second[index].kind != 1) {
        if (err) *err = peUnset;
        return 0;
    }
    if (err) *err = peSuccess;
    return it->second[index].i;
}
int m_mapSetInt(Map* m, const char* key, int64_t i, int) {
    Value v;
    v.kind = 1;
    v.i = i;
    m->kv[key] = {v};
    return 0;
}
double m_mapGetFloat(const Map* m, const char* key, int index, int* err) {
    auto it = m->kv.find(key);
    if (it == m->kv.end() || index >= (int)it->second.size() || it->second[index].kind != 2) {
        if (err) *err = peUnset;
        return 0;
    }
    if (err) *err = peSuccess;
    return it->second[index].f;
}
int m_mapSetFloat(Map* m, const char* key, double d, int) {
    Value v;
    v.kind = 2;
    v.f = d;
    m->kv[key] = {v};
    return 0;
}
int m_mapSetData(Map* m, const char* key, const char* data, int size, int, int) {
    Value v;
    v.kind = 3;
    v.data.assign(data, size < 0 ? std::strlen(data) : (size_t)size);
    m->kv[key] = {v};
    return 0;
}
Node* m_mapGetNode(const Map* m, const char* key, int index, int* err) {
    auto it = m->kv.find(key);
    if (it == m->kv.end() || index >= (int)it->second.size() || it->second[index].kind != 5) {
        if (err) *err = peUnset;
        return nullptr;
    }
    if (err) *err = peSuccess;
    it->second[index].node->gets++;
    return it->second[index].node;
}
int m_mapConsumeNode(Map* m, const char* key, Node* n, int) {
    Value v;
    v.kind = 5;
    v.node = n;
    m->kv[key] = {v};
    return 0;
}
Plugin* m_getPluginByID(const char*, Core*) { return nullptr; }  // no resize plugin in this host
Map* m_invoke(Plugin*, const char*, const Map*) { return nullptr; }
void m_getCoreInfo(Core* core, CoreInfo* info) {
    std::memset(info, 0, sizeof(*info));
    info->versionString = "mock";
    info->core = 70;
    info->api = kApiVersion;
    info->numThreads = core->threads;
}

template <int K>
void unexpected() {
    std::fprintf(stderr, "vs mock host: unexpected VSAPI entry %d\n", K);
    std::abort();
}
template <int... K>
void fill_stubs(Api& a, std::integer_sequence<int, K...>) {
    ((a.slot[K] = reinterpret_cast<void (*)()>(&unexpected<K>)), ...);
}

Api make_api() {
    Api a;
    fill_stubs(a, std::make_integer_sequence<int, kSlotsUsed>{});
#define SET(name) a.slot[name] = reinterpret_cast<void (*)()>(&m_##name)
    SET(createVideoFilter);
    SET(freeNode);
    SET(getVideoInfo);
    SET(freeFrame);
    SET(copyFrame);
    SET(getFramePropertiesRW);
    SET(getStride);
    SET(getReadPtr);
    SET(getFrameWidth);
    SET(getFrameHeight);
    SET(getFrameFilter);
    SET(requestFrameFilter);
    SET(setFilterError);
    SET(createMap);
    SET(freeMap);
    SET(mapSetError);
    SET(mapGetError);
    SET(mapGetInt);
    SET(mapSetInt);
    SET(mapSetFloat);
    SET(mapSetData);
    SET(mapGetNode);
    SET(mapConsumeNode);
    SET(getPluginByID);
    SET(invoke);
    SET(getCoreInfo);
    SET(newVideoFrame);
    SET(getWritePtr);
    SET(queryVideoFormat);
    SET(mapGetFloat);
#undef SET
    return a;
}

// --- plugin API given to VapourSynthPluginInit2 ---------------------------
int p_getAPIVersion() { return kApiVersion; }
int p_configPlugin(const char* id, const char* ns, const char* name, int version, int api, int, Plugin* p) {
    p->id = id;
    p->ns = ns;
    p->name = name;
    p->version = version;
    p->api = api;
    return 1;
}
int p_registerFunction(const char* name, const char* args, const char* ret, PublicFunction fn, void*, Plugin* p) {
    p->fns.push_back({name, args, ret, fn});
    return 1;
}

bool load_plugin(const char* libpath, Plugin* p, std::string* err) {
    void* h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        *err = dlerror();
        return false;
    }
    typedef void (*Init)(Plugin*, const PluginApi*);
    Init init = reinterpret_cast<Init>(dlsym(h, "VapourSynthPluginInit2"));
    if (!init) {
        *err = "VapourSynthPluginInit2 not exported";
        return false;
    }
    static const PluginApi papi = {&p_getAPIVersion, &p_configPlugin, &p_registerFunction};
    init(p, &papi);
    return true;
}

PublicFunction find_fn(const Plugin& p, const char* name) {
    for (const auto& f : p.fns)
        if (f.name == name) return f.fn;
    return nullptr;
}

void copy_out(const std::string& s, char* buf, int len) {
    if (!buf || len <= 0) return;
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
}

Node* source_node(int w, int h, int nframes, const uint8_t* const* planes, int stride, VideoFormat fmt) {
    Node* n = new Node;
    n->vi.format = fmt;
    n->vi.fpsNum = 24;
    n->vi.fpsDen = 1;
    n->vi.width = w;
    n->vi.height = h;
    n->vi.numFrames = nframes;
    for (int i = 0; i < nframes; i++) {
        Frame* f = new Frame;
        f->w = w;
        f->h = h;
        f->stride = stride;
        for (int p = 0; p < 3; p++) f->planes[p] = planes ? planes[i * 3 + p] : nullptr;
        n->frames.push_back(f);
    }
    return n;
}

VideoFormat rgbs() { return VideoFormat{cfRGB, stFloat, 32, 4, 0, 0, 3}; }
VideoFormat yuv8() { return VideoFormat{cfYUV, stInteger, 8, 1, 1, 1, 3}; }

}  // namespace

extern "C" {

// Registration as text: "config|id|ns|name|version|api" then one
// "fn|name|args|ret" line per registered function.
int vsmock_registration(const char* libpath, char* out, int len) {
    Plugin p;
    std::string err;
    if (!load_plugin(libpath, &p, &err)) {
        copy_out(err, out, len);
        return 1;
    }
    std::string s = "config|" + p.id + "|" + p.ns + "|" + p.name + "|" + std::to_string(p.version) + "|" +
                    std::to_string(p.api) + "\n";
    for (const auto& f : p.fns) s += "fn|" + f.name + "|" + f.args + "|" + f.ret + "\n";
    copy_out(s, out, len);
    return 0;
}

// Calls SSIMULACRA2 on two clips and scores every frame from `nthreads`
// threads. frames*: nframes*3 plane pointers (32-bit float, `stride` bytes
// per row); rgbs1/rgbs2: 0 makes that clip 8-bit YUV 4:2:0 instead (to reach
// the conversion path); w2/h2: size of the distorted clip. numStream/gpu_id
// < 0: argument absent. Returns 0 and scores[nframes] on success; 1 with the
// create error in err; 2 with the first frame error in err.
int vsmock_ssimu2_timed(const char* libpath, int w, int h, int w2, int h2, int nframes, const uint8_t* const* ref,
                        const uint8_t* const* dist, int stride, int rgbs1, int rgbs2, int numStream, int gpu_id,
                        int core_threads, int nthreads, double* scores, int* output_matches, char* err, int errlen,
                        double* loop_ms);

int vsmock_ssimu2(const char* libpath, int w, int h, int w2, int h2, int nframes, const uint8_t* const* ref,
                  const uint8_t* const* dist, int stride, int rgbs1, int rgbs2, int numStream, int gpu_id,
                  int core_threads, int nthreads, double* scores, int* output_matches, char* err, int errlen) {
    return vsmock_ssimu2_timed(libpath, w, h, w2, h2, nframes, ref, dist, stride, rgbs1, rgbs2, numStream, gpu_id,
                               core_threads, nthreads, scores, output_matches, err, errlen, nullptr);
}

// vsmock_ssimu2, and *loop_ms (if not null) receives the wall time of the
// frame requests alone: from the first request to the last frame returned,
// without the filter's creation and free.
int vsmock_ssimu2_timed(const char* libpath, int w, int h, int w2, int h2, int nframes, const uint8_t* const* ref,
                        const uint8_t* const* dist, int stride, int rgbs1, int rgbs2, int numStream, int gpu_id,
                        int core_threads, int nthreads, double* scores, int* output_matches, char* err, int errlen,
                        double* loop_ms) {
    Plugin p;
    std::string e;
    if (!load_plugin(libpath, &p, &e)) {
        copy_out(e, err, errlen);
        return 1;
    }
    PublicFunction create = find_fn(p, "SSIMULACRA2");
    if (!create) {
        copy_out("SSIMULACRA2 not registered", err, errlen);
        return 1;
    }
    const Api api = make_api();
    Core core;
    core.threads = core_threads;
    Node* n1 = source_node(w, h, nframes, ref, stride, rgbs1 ? rgbs() : yuv8());
    Node* n2 = source_node(w2, h2, nframes, dist, stride, rgbs2 ? rgbs() : yuv8());
    Map in, out;
    m_mapConsumeNode(&in, "reference", n1, 0);
    m_mapConsumeNode(&in, "distorted", n2, 0);
    if (numStream >= 0) m_mapSetInt(&in, "numStream", numStream, 0);
    if (gpu_id >= 0) m_mapSetInt(&in, "gpu_id", gpu_id, 0);
    create(&in, &out, nullptr, &core, &api);
    if (!out.error.empty()) {
        copy_out(out.error, err, errlen);
        const bool released = n1->freed == n1->gets && n2->freed == n2->gets;
        return released ? 1 : 3;  // 3: an error path leaked or double-freed an input node
    }
    Node* filt = out.kv["clip"][0].node;
    std::mutex mu;
    std::string first_err;
    int matches = 0;
    auto work = [&](int t) {
        for (int i = t; i < nframes; i += nthreads) {
            FrameContext ctx;
            void* fd = nullptr;
            const Frame* r = filt->get(i, arInitial, filt->inst, &fd, &ctx, &core, &api);
            bool ok = r == nullptr && ctx.requested.size() == 2 && ctx.requested[0].first == i &&
                      ctx.requested[1].first == i;
            r = filt->get(i, arAllFramesReady, filt->inst, &fd, &ctx, &core, &api);
            std::lock_guard<std::mutex> g(mu);
            if (!r) {
                if (first_err.empty()) first_err = ctx.error.empty() ? "no frame and no error" : ctx.error;
                continue;
            }
            auto it = r->props.kv.find("_SSIMULACRA2");
            if (it == r->props.kv.end() || it->second[0].kind != 2) {
                if (first_err.empty()) first_err = "no _SSIMULACRA2 property";
            } else {
                scores[i] = it->second[0].f;
            }
            for (int pl = 0; pl < 3 && ok; pl++)
                ok = r->planes[pl] == dist[i * 3 + pl] ||
                     std::memcmp(r->planes[pl], dist[i * 3 + pl], (size_t)stride * h) == 0;
            matches += ok ? 1 : 0;
            m_freeFrame(r);
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    if (loop_ms) *loop_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    filt->free(filt->inst, &core, &api);
    if (output_matches) *output_matches = matches;
    if (!first_err.empty()) {
        copy_out(first_err, err, errlen);
        return 2;
    }
    if (n1->freed != n1->gets || n2->freed != n2->gets || g_output_frames != 0) {
        copy_out("node or frame references not released exactly once", err, errlen);
        return 4;
    }
    return 0;
}

// Calls BUTTERAUGLI on two RGBS clips of nframes frames and scores every
// frame from `nthreads` threads. qnorm/numStream/gpu_id < 0 and intensity < 0:
// argument absent. norms[3 * i ...] receive the three frame properties; with
// distmap, maps[i * w * h ...] the GrayS output frame, otherwise the output
// frame must equal the distorted one (counted in *output_matches). Returns as
// vsmock_ssimu2; 5: the output clip's format is not the expected one.
int vsmock_butter(const char* libpath, int w, int h, int w2, int h2, int nframes, const uint8_t* const* ref,
                  const uint8_t* const* dist, int stride, int qnorm, double intensity, int distmap, int numStream,
                  int gpu_id, int core_threads, int nthreads, double* norms, float* maps, int* output_matches,
                  char* err, int errlen) {
    Plugin p;
    std::string e;
    if (!load_plugin(libpath, &p, &e)) {
        copy_out(e, err, errlen);
        return 1;
    }
    PublicFunction create = find_fn(p, "BUTTERAUGLI");
    if (!create) {
        copy_out("BUTTERAUGLI not registered", err, errlen);
        return 1;
    }
    const Api api = make_api();
    Core core;
    core.threads = core_threads;
    Node* n1 = source_node(w, h, nframes, ref, stride, rgbs());
    Node* n2 = source_node(w2, h2, nframes, dist, stride, rgbs());
    Map in, out;
    m_mapConsumeNode(&in, "reference", n1, 0);
    m_mapConsumeNode(&in, "distorted", n2, 0);
    if (qnorm >= 0) m_mapSetInt(&in, "qnorm", qnorm, 0);
    if (intensity >= 0) m_mapSetFloat(&in, "intensity_multiplier", intensity, 0);
    if (distmap >= 0) m_mapSetInt(&in, "distmap", distmap, 0);
    if (numStream >= 0) m_mapSetInt(&in, "numStream", numStream, 0);
    if (gpu_id >= 0) m_mapSetInt(&in, "gpu_id", gpu_id, 0);
    create(&in, &out, nullptr, &core, &api);
    if (!out.error.empty()) {
        copy_out(out.error, err, errlen);
        const bool released = n1->freed == n1->gets && n2->freed == n2->gets;
        return released ? 1 : 3;
    }
    Node* filt = out.kv["clip"][0].node;
    const bool want_gray = distmap > 0;
    if ((filt->vi.format.colorFamily == cfGray) != want_gray || filt->vi.format.bitsPerSample != 32) {
        filt->free(filt->inst, &core, &api);
        copy_out("unexpected output format", err, errlen);
        return 5;
    }
    std::mutex mu;
    std::string first_err;
    int matches = 0;
    static const char* keys[3] = {"_BUTTERAUGLI_QNorm", "_BUTTERAUGLI_3Norm", "_BUTTERAUGLI_INFNorm"};
    auto work = [&](int t) {
        for (int i = t; i < nframes; i += nthreads) {
            FrameContext ctx;
            void* fd = nullptr;
            const Frame* r = filt->get(i, arInitial, filt->inst, &fd, &ctx, &core, &api);
            bool ok = r == nullptr && ctx.requested.size() == 2;
            r = filt->get(i, arAllFramesReady, filt->inst, &fd, &ctx, &core, &api);
            std::lock_guard<std::mutex> g(mu);
            if (!r) {
                if (first_err.empty()) first_err = ctx.error.empty() ? "no frame and no error" : ctx.error;
                continue;
            }
            for (int k = 0; k < 3; k++) {
                auto it = r->props.kv.find(keys[k]);
                if (it == r->props.kv.end() || it->second[0].kind != 2) {
                    if (first_err.empty()) first_err = std::string("no property ") + keys[k];
                } else {
                    norms[3 * i + k] = it->second[0].f;
                }
            }
            if (want_gray) {
                for (int y = 0; y < h; y++)
                    std::memcpy(maps + (size_t)i * w * h + (size_t)y * w, r->planes[0] + (size_t)y * r->stride,
                                (size_t)w * 4);
            } else {
                for (int pl = 0; pl < 3 && ok; pl++)
                    ok = std::memcmp(r->planes[pl], dist[i * 3 + pl], (size_t)stride * h) == 0;
                matches += ok ? 1 : 0;
            }
            m_freeFrame(r);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    filt->free(filt->inst, &core, &api);
    if (output_matches) *output_matches = matches;
    if (!first_err.empty()) {
        copy_out(first_err, err, errlen);
        return 2;
    }
    if (n1->freed != n1->gets || n2->freed != n2->gets || g_output_frames != 0) {
        copy_out("node or frame references not released exactly once", err, errlen);
        return 4;
    }
    return 0;
}

// GpuInfo with (gpu_id >= 0) or without the argument; the data or the error.
int vsmock_gpuinfo(const char* libpath, int gpu_id, char* out, int len) {
    Plugin p;
    std::string e;
    if (!load_plugin(libpath, &p, &e)) {
        copy_out(e, out, len);
        return 1;
    }
    PublicFunction fn = find_fn(p, "GpuInfo");
    if (!fn) {
        copy_out("GpuInfo not registered", out, len);
        return 1;
    }
    const Api api = make_api();
    Core core;
    Map in, res;
    if (gpu_id >= 0) m_mapSetInt(&in, "gpu_id", gpu_id, 0);
    fn(&in, &res, nullptr, &core, &api);
    if (!res.error.empty()) {
        copy_out(res.error, out, len);
        return 1;
    }
    copy_out(res.kv["gpu_human_data"][0].data, out, len);
    return 0;
}

}  // extern "C"
