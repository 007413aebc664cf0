// VapourSynth plugin of libvship.so: vship.SSIMULACRA2, vship.BUTTERAUGLI and
// vship.GpuInfo, the filter interface of the reference (src/VshipLib.cpp:8-64,
// src/ssimu2/vapoursynth.hpp:8-166, src/butter/vapoursynth.hpp:6-193).
//
//   SSIMULACRA2(reference:vnode, distorted:vnode, numStream:int:opt, gpu_id:int:opt) -> clip
//     output frame n = the RGBS form of distorted frame n with the float
//     property _SSIMULACRA2 (the score of the pair). Inputs that are not
//     RGBS are converted with the core's resize plugin (Bicubic, BT.709
//     transfer and primaries, limited-range YUV in, full-range RGB out; the
//     input matrix is BT.709 above 650 rows and BT.601 below), as
//     src/util/torgbs.hpp:8-36 does, so the scores and output frames are
//     the reference's.
//   BUTTERAUGLI(reference:vnode, distorted:vnode, qnorm:int:opt,
//               intensity_multiplier:float:opt, distmap:int:opt, numStream:int:opt,
//               gpu_id:int:opt) -> clip
//     output frame n = the RGBS form of distorted frame n, or with distmap=1
//     the GrayS distortion map, carrying the float properties
//     _BUTTERAUGLI_QNorm, _BUTTERAUGLI_3Norm and _BUTTERAUGLI_INFNorm.
//     Defaults: qnorm 2, intensity_multiplier 203, numStream 4 (clamped to
//     [1, threads]), as there.
//   GpuInfo(gpu_id:int:opt) -> gpu_human_data
//
// numStream per-frame engines (default: the core's thread count, clamped
// to [1, threads]) serve the frame requests VapourSynth makes in parallel.
// Errors are reported in the reference's text format
// ("VshipException\n<Enum>: <message>\n - At line <L> of <file>\n").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "engine.hpp"
#include "vs4.hpp"

#define VSHIP_PLUGIN_EXPORT extern "C" __attribute__((visibility("default")))

namespace vship {
namespace {

using vs4::Api;

std::string vs_error_text(const VshipErr& e) {
    std::ostringstream ss;
    ss << "VshipException\n" << error_message(e.code) << "\n - At line " << e.line << " of " << e.file << "\n";
    return ss.str();
}

std::string vs_error_text(ErrCode c, const char* file, int line) { return vs_error_text(VshipErr(c, file, line)); }

// A pool of per-frame engines; a frame request takes one, scores, returns it.
template <typename H>
struct EnginePool {
    std::vector<std::unique_ptr<H>> handles;
    std::vector<int> free_ids;
    std::mutex mu;
    std::condition_variable cv;

    int take() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !free_ids.empty(); });
        const int id = free_ids.back();
        free_ids.pop_back();
        return id;
    }
    void give(int id) {
        {
            std::lock_guard<std::mutex> lk(mu);
            free_ids.push_back(id);
        }
        cv.notify_one();
    }
};

struct Ssimu2Filter {
    vs4::Node* reference = nullptr;
    vs4::Node* distorted = nullptr;
    EnginePool<FrameHandle> pool;
};

struct ButterFilter {
    vs4::Node* reference = nullptr;
    vs4::Node* distorted = nullptr;
    bool distmap = false;
    vs4::VideoFormat gray{};
    EnginePool<ButterHandle> pool;
};

// The RGBS form of a clip (src/util/torgbs.hpp:8-36): the clip itself when it
// is already 32-bit float RGB, else the resize plugin's Bicubic conversion.
// Takes ownership of `source`; returns nullptr and sets *err on failure.
vs4::Node* to_rgbs(vs4::Node* source, vs4::Core* core, const Api* api, std::string* err) {
    const vs4::VideoInfo* vi = api->call<vs4::getVideoInfo>(source);
    if (vi->format.colorFamily == vs4::cfRGB && vi->format.sampleType == vs4::stFloat &&
        vi->format.bitsPerSample == 32)
        return source;
    vs4::Plugin* resize = api->call<vs4::getPluginByID>("com.vapoursynth.resize", core);
    if (!resize) {
        api->call<vs4::freeNode>(source);
        *err = vs_error_text(E_NonRGBSInput, __FILE__, __LINE__);
        return nullptr;
    }
    vs4::Map* args = api->call<vs4::createMap>();
    api->call<vs4::mapConsumeNode>(args, "clip", source, (int)vs4::maReplace);
    const struct {
        const char* key;
        int64_t value;
    } opts[] = {
        {"matrix_in", vi->height > 650 ? 1 : 6},  // BT.709, or BT.601 (SMPTE 170M) for SD
        {"transfer_in", 1}, {"primaries_in", 1}, {"range_in", 0},
        {"matrix", 0}, {"transfer", 1}, {"primaries", 1}, {"range", 1},
        {"format", (int64_t)vs4::pfRGBS},
    };
    for (const auto& o : opts) api->call<vs4::mapSetInt>(args, o.key, o.value, (int)vs4::maReplace);
    vs4::Map* ret = api->call<vs4::invoke>(resize, "Bicubic", (const vs4::Map*)args);
    vs4::Node* out = nullptr;
    if (ret) {
        const char* e = api->call<vs4::mapGetError>((const vs4::Map*)ret);
        if (e)
            *err = e;
        else
            out = api->call<vs4::mapGetNode>((const vs4::Map*)ret, "clip", 0, (int*)nullptr);
        api->call<vs4::freeMap>(ret);
    } else {
        *err = vs_error_text(E_NonRGBSInput, __FILE__, __LINE__);
    }
    api->call<vs4::freeMap>(args);
    return out;
}

bool same_video_info(const vs4::VideoInfo* a, const vs4::VideoInfo* b) {
    return a->width == b->width && a->height == b->height && a->numFrames == b->numFrames &&
           a->fpsNum == b->fpsNum && a->fpsDen == b->fpsDen && a->format.colorFamily == b->format.colorFamily &&
           a->format.sampleType == b->format.sampleType && a->format.bitsPerSample == b->format.bitsPerSample &&
           a->format.subSamplingW == b->format.subSamplingW && a->format.subSamplingH == b->format.subSamplingH;
}

const vs4::Frame* ssimu2_get_frame(int n, int reason, void* inst, void** /*frameData*/, vs4::FrameContext* ctx,
                                   vs4::Core* core, const Api* api) {
    Ssimu2Filter* d = static_cast<Ssimu2Filter*>(inst);
    if (reason == vs4::arInitial) {
        api->call<vs4::requestFrameFilter>(n, d->reference, ctx);
        api->call<vs4::requestFrameFilter>(n, d->distorted, ctx);
        return nullptr;
    }
    if (reason != vs4::arAllFramesReady) return nullptr;
    const vs4::Frame* f1 = api->call<vs4::getFrameFilter>(n, d->reference, ctx);
    const vs4::Frame* f2 = api->call<vs4::getFrameFilter>(n, d->distorted, ctx);
    const uint8_t* p1[3];
    const uint8_t* p2[3];
    for (int i = 0; i < 3; i++) {
        p1[i] = api->call<vs4::getReadPtr>(f1, i);
        p2[i] = api->call<vs4::getReadPtr>(f2, i);
    }
    const int64_t s1 = (int64_t)api->call<vs4::getStride>(f1, 0);
    const int64_t s2 = (int64_t)api->call<vs4::getStride>(f2, 0);
    double score = 0.0;
    std::string err;
    const int id = d->pool.take();
    try {
        score = d->pool.handles[id]->compute(IN_RGB_F32, p1, p2, s1, s2);
    } catch (const VshipErr& e) {
        err = vs_error_text(e);
    } catch (...) {
        err = vs_error_text(E_BadErrorType, __FILE__, __LINE__);
    }
    d->pool.give(id);
    if (!err.empty()) {
        api->call<vs4::setFilterError>(err.c_str(), ctx);
        api->call<vs4::freeFrame>(f1);
        api->call<vs4::freeFrame>(f2);
        return nullptr;
    }
    vs4::Frame* dst = api->call<vs4::copyFrame>(f2, core);
    api->call<vs4::mapSetFloat>(api->call<vs4::getFramePropertiesRW>(dst), "_SSIMULACRA2", score,
                                (int)vs4::maReplace);
    api->call<vs4::freeFrame>(f1);
    api->call<vs4::freeFrame>(f2);
    return dst;
}

void ssimu2_free(void* inst, vs4::Core*, const Api* api) {
    Ssimu2Filter* d = static_cast<Ssimu2Filter*>(inst);
    api->call<vs4::freeNode>(d->reference);
    api->call<vs4::freeNode>(d->distorted);
    delete d;
}

void ssimu2_create(const vs4::Map* in, vs4::Map* out, void*, vs4::Core* core, const Api* api) {
    std::string err;
    vs4::Node* ref = to_rgbs(api->call<vs4::mapGetNode>(in, "reference", 0, (int*)nullptr), core, api, &err);
    vs4::Node* dis = ref ? to_rgbs(api->call<vs4::mapGetNode>(in, "distorted", 0, (int*)nullptr), core, api, &err)
                         : nullptr;
    if (!ref || !dis) {
        if (ref) api->call<vs4::freeNode>(ref);
        api->call<vs4::mapSetError>(out, err.c_str());
        return;
    }
    auto fail = [&](const std::string& msg) {
        api->call<vs4::freeNode>(ref);
        api->call<vs4::freeNode>(dis);
        api->call<vs4::mapSetError>(out, msg.c_str());
    };
    const vs4::VideoInfo* vi1 = api->call<vs4::getVideoInfo>(ref);
    const vs4::VideoInfo* vi2 = api->call<vs4::getVideoInfo>(dis);
    if (!same_video_info(vi1, vi2)) return fail(vs_error_text(E_DifferingInputType, __FILE__, __LINE__));
    if (vi1->format.bitsPerSample != 32 || vi1->format.colorFamily != vs4::cfRGB ||
        vi1->format.sampleType != vs4::stFloat)
        return fail(vs_error_text(E_NonRGBSInput, __FILE__, __LINE__));
    int e = 0;
    int64_t gpu = api->call<vs4::mapGetInt>(in, "gpu_id", 0, &e);
    if (e != vs4::peSuccess) gpu = 0;
    vs4::CoreInfo info{};
    api->call<vs4::getCoreInfo>(core, &info);
    int64_t streams = api->call<vs4::mapGetInt>(in, "numStream", 0, &e);
    if (e != vs4::peSuccess) streams = info.numThreads;
    streams = std::max<int64_t>(1, std::min<int64_t>(streams, std::max(1, info.numThreads)));
    auto d = std::make_unique<Ssimu2Filter>();
    try {
        gpu_full_check((int)gpu);  // also makes gpu the current device
        for (int i = 0; i < (int)streams; i++) {
            d->pool.handles.push_back(std::make_unique<FrameHandle>(vi1->width, vi1->height, (int)gpu));
            d->pool.free_ids.push_back(i);
        }
    } catch (const VshipErr& x) {
        return fail(vs_error_text(x));
    } catch (...) {
        return fail(vs_error_text(E_OutOfRAM, __FILE__, __LINE__));
    }
    d->reference = ref;
    d->distorted = dis;
    const vs4::FilterDependency deps[2] = {{ref, vs4::rpStrictSpatial}, {dis, vs4::rpStrictSpatial}};
    api->call<vs4::createVideoFilter>(out, "vship", vi1, &ssimu2_get_frame, &ssimu2_free, (int)vs4::fmParallel,
                                      deps, 2, (void*)d.release(), core);
}

const vs4::Frame* butter_get_frame(int n, int reason, void* inst, void** /*frameData*/, vs4::FrameContext* ctx,
                                   vs4::Core* core, const Api* api) {
    ButterFilter* d = static_cast<ButterFilter*>(inst);
    if (reason == vs4::arInitial) {
        api->call<vs4::requestFrameFilter>(n, d->reference, ctx);
        api->call<vs4::requestFrameFilter>(n, d->distorted, ctx);
        return nullptr;
    }
    if (reason != vs4::arAllFramesReady) return nullptr;
    const vs4::Frame* f1 = api->call<vs4::getFrameFilter>(n, d->reference, ctx);
    const vs4::Frame* f2 = api->call<vs4::getFrameFilter>(n, d->distorted, ctx);
    const int w = api->call<vs4::getFrameWidth>(f1, 0), h = api->call<vs4::getFrameHeight>(f1, 0);
    const uint8_t* p1[3];
    const uint8_t* p2[3];
    for (int i = 0; i < 3; i++) {
        p1[i] = api->call<vs4::getReadPtr>(f1, i);
        p2[i] = api->call<vs4::getReadPtr>(f2, i);
    }
    const int64_t s1 = (int64_t)api->call<vs4::getStride>(f1, 0);
    const int64_t s2 = (int64_t)api->call<vs4::getStride>(f2, 0);
    vs4::Frame* dst = d->distmap ? api->call<vs4::newVideoFrame>((const vs4::VideoFormat*)&d->gray, w, h,
                                                                 (const vs4::Frame*)nullptr, core)
                                 : api->call<vs4::copyFrame>(f2, core);
    float* mp = d->distmap ? reinterpret_cast<float*>(api->call<vs4::getWritePtr>(dst, 0)) : nullptr;
    const int64_t ms = d->distmap ? (int64_t)api->call<vs4::getStride>((const vs4::Frame*)dst, 0) : 0;
    double norms[3] = {0.0, 0.0, 0.0};
    std::string err;
    const int id = d->pool.take();
    try {
        d->pool.handles[id]->compute(IN_RGB_F32, p1, p2, s1, s2, mp, ms, norms);
    } catch (const VshipErr& e) {
        err = vs_error_text(e);
    } catch (...) {
        err = vs_error_text(E_BadErrorType, __FILE__, __LINE__);
    }
    d->pool.give(id);
    api->call<vs4::freeFrame>(f1);
    api->call<vs4::freeFrame>(f2);
    if (!err.empty()) {
        api->call<vs4::freeFrame>((const vs4::Frame*)dst);
        api->call<vs4::setFilterError>(err.c_str(), ctx);
        return nullptr;
    }
    vs4::Map* props = api->call<vs4::getFramePropertiesRW>(dst);
    api->call<vs4::mapSetFloat>(props, "_BUTTERAUGLI_QNorm", norms[0], (int)vs4::maReplace);
    api->call<vs4::mapSetFloat>(props, "_BUTTERAUGLI_3Norm", norms[1], (int)vs4::maReplace);
    api->call<vs4::mapSetFloat>(props, "_BUTTERAUGLI_INFNorm", norms[2], (int)vs4::maReplace);
    return dst;
}

void butter_free(void* inst, vs4::Core*, const Api* api) {
    ButterFilter* d = static_cast<ButterFilter*>(inst);
    api->call<vs4::freeNode>(d->reference);
    api->call<vs4::freeNode>(d->distorted);
    delete d;
}

void butter_create(const vs4::Map* in, vs4::Map* out, void*, vs4::Core* core, const Api* api) {
    std::string err;
    vs4::Node* ref = to_rgbs(api->call<vs4::mapGetNode>(in, "reference", 0, (int*)nullptr), core, api, &err);
    vs4::Node* dis = ref ? to_rgbs(api->call<vs4::mapGetNode>(in, "distorted", 0, (int*)nullptr), core, api, &err)
                         : nullptr;
    if (!ref || !dis) {
        if (ref) api->call<vs4::freeNode>(ref);
        api->call<vs4::mapSetError>(out, err.c_str());
        return;
    }
    auto fail = [&](const std::string& msg) {
        api->call<vs4::freeNode>(ref);
        api->call<vs4::freeNode>(dis);
        api->call<vs4::mapSetError>(out, msg.c_str());
    };
    const vs4::VideoInfo* vi1 = api->call<vs4::getVideoInfo>(ref);
    const vs4::VideoInfo* vi2 = api->call<vs4::getVideoInfo>(dis);
    if (!same_video_info(vi1, vi2)) return fail(vs_error_text(E_DifferingInputType, __FILE__, __LINE__));
    if (vi1->format.bitsPerSample != 32 || vi1->format.colorFamily != vs4::cfRGB ||
        vi1->format.sampleType != vs4::stFloat)
        return fail(vs_error_text(E_NonRGBSInput, __FILE__, __LINE__));
    int e = 0;
    double intensity = api->call<vs4::mapGetFloat>(in, "intensity_multiplier", 0, &e);
    if (e != vs4::peSuccess) intensity = 203.0;
    int64_t gpu = api->call<vs4::mapGetInt>(in, "gpu_id", 0, &e);
    if (e != vs4::peSuccess) gpu = 0;
    int64_t distmap = api->call<vs4::mapGetInt>(in, "distmap", 0, &e);
    if (e != vs4::peSuccess) distmap = 0;
    int64_t qnorm = api->call<vs4::mapGetInt>(in, "qnorm", 0, &e);
    if (e != vs4::peSuccess) qnorm = 2;
    vs4::CoreInfo info{};
    api->call<vs4::getCoreInfo>(core, &info);
    int64_t streams = api->call<vs4::mapGetInt>(in, "numStream", 0, &e);
    if (e != vs4::peSuccess) streams = 4;
    streams = std::max<int64_t>(1, std::min<int64_t>(streams, std::max(1, info.numThreads)));
    auto d = std::make_unique<ButterFilter>();
    d->distmap = distmap != 0;
    vs4::VideoInfo viout = *vi1;
    if (d->distmap) {
        api->call<vs4::queryVideoFormat>(&d->gray, (int)vs4::cfGray, (int)vs4::stFloat, 32, 0, 0, core);
        viout.format = d->gray;
    }
    try {
        gpu_full_check((int)gpu);  // also makes gpu the current device
        for (int i = 0; i < (int)streams; i++) {
            d->pool.handles.push_back(
                std::make_unique<ButterHandle>(vi1->width, vi1->height, (int)qnorm, (float)intensity, (int)gpu));
            d->pool.free_ids.push_back(i);
        }
    } catch (const VshipErr& x) {
        return fail(vs_error_text(x));
    } catch (...) {
        return fail(vs_error_text(E_OutOfRAM, __FILE__, __LINE__));
    }
    d->reference = ref;
    d->distorted = dis;
    const vs4::FilterDependency deps[2] = {{ref, vs4::rpStrictSpatial}, {dis, vs4::rpStrictSpatial}};
    api->call<vs4::createVideoFilter>(out, "vship", (const vs4::VideoInfo*)&viout, &butter_get_frame, &butter_free,
                                      (int)vs4::fmParallel, deps, 2, (void*)d.release(), core);
}

// Device description (src/VshipLib.cpp:8-56): every device's name when no
// gpu_id is given, else the properties of that device.
void gpu_info(const vs4::Map* in, vs4::Map* out, void*, vs4::Core*, const Api* api) {
    std::ostringstream ss;
    int count = 0;
    try {
        count = device_count_or_throw();
    } catch (const VshipErr& x) {
        api->call<vs4::mapSetError>(out, vs_error_text(x).c_str());
        return;
    }
    int e = 0;
    const int64_t gpu = api->call<vs4::mapGetInt>(in, "gpu_id", 0, &e);
    const bool given = e == vs4::peSuccess;
    if (given && (gpu < 0 || gpu >= count)) {
        api->call<vs4::mapSetError>(out, vs_error_text(E_BadDeviceArgument, __FILE__, __LINE__).c_str());
        return;
    }
    hipDeviceProp_t p;
    if (!given) {
        for (int i = 0; i < count; i++) {
            if (hipGetDeviceProperties(&p, i) != hipSuccess) continue;
            ss << "GPU " << i << ": " << p.name << "\n";
        }
    } else {
        (void)hipSetDevice((int)gpu);
        if (hipGetDeviceProperties(&p, (int)gpu) == hipSuccess) {
            ss << "Name: " << p.name << "\n";
            ss << "MultiProcessorCount: " << p.multiProcessorCount << "\n";
            ss << "MaxSharedMemoryPerBlock: " << p.sharedMemPerBlock << " bytes\n";
            ss << "WarpSize: " << p.warpSize << "\n";
            ss << "VRAMCapacity: " << (float)p.totalGlobalMem / 1e9f << " GB\n";
            ss << "MemoryBusWidth: " << p.memoryBusWidth << " bits\n";
            ss << "Integrated: " << p.integrated << "\n";
            ss << "PassKernelCheck : " << (int)gpu_kernel_check() << "\n";
        }
    }
    const std::string s = ss.str();
    api->call<vs4::mapSetData>(out, "gpu_human_data", s.data(), (int)s.size(), (int)vs4::dtUtf8,
                               (int)vs4::maReplace);
}

// The reference's CVVDP metric (src/VshipLib.cpp:62) is not built here; it
// stays registered so a script calling it gets this error, not a
// missing-function one.
void metric_not_built(const vs4::Map*, vs4::Map* out, void*, vs4::Core*, const Api* api) {
    const std::string msg =
        "VshipException\nBadErrorType: this vship build (MI355X) provides SSIMULACRA2 and BUTTERAUGLI only\n";
    api->call<vs4::mapSetError>(out, msg.c_str());
}

}  // namespace
}  // namespace vship

// Plugin registration (src/VshipLib.cpp:58-64): identifier, namespace and
// the argument strings are the interface scripts call, kept as they are.
VSHIP_PLUGIN_EXPORT void VapourSynthPluginInit2(vs4::Plugin* plugin, const vs4::PluginApi* vspapi) {
    vspapi->configPlugin("com.lumen.vship", "vship", "VapourSynth GPU Perceptual Metrics", vs4::make_version(3, 3),
                         vs4::kApiVersion, 0, plugin);
    vspapi->registerFunction("SSIMULACRA2", "reference:vnode;distorted:vnode;numStream:int:opt;gpu_id:int:opt;",
                             "clip:vnode;", &vship::ssimu2_create, nullptr, plugin);
    vspapi->registerFunction("BUTTERAUGLI",
                             "reference:vnode;distorted:vnode;qnorm:int:opt;intensity_multiplier:float:opt;"
                             "distmap:int:opt;numStream:int:opt;gpu_id:int:opt;",
                             "clip:vnode;", &vship::butter_create, nullptr, plugin);
    vspapi->registerFunction("CVVDP", "reference:vnode;distorted:vnode;display:data:opt;numStream:int:opt;gpu_id:int:opt;",
                             "clip:vnode;", &vship::metric_not_built, nullptr, plugin);
    vspapi->registerFunction("GpuInfo", "gpu_id:int:opt;", "gpu_human_data:data;", &vship::gpu_info, nullptr, plugin);
}
