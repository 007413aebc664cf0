// The part of the VapourSynth API v4 ABI that the vship filter uses,
// declared here so the library builds without VapourSynth's headers.
//
// VSAPI is a table of function pointers whose order is the ABI: `Slot` gives
// the index of each entry used here, `Api::call<Slot>` casts the entry to its
// signature. The struct layouts and enum values below are the ABI's as
// well. A CPU test compiles this header next to VapourSynth's own
// VapourSynth4.h and checks every slot index (offsetof / pointer size), every
// struct offset and every constant.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace vs4 {

struct Frame;
struct Node;
struct Core;
struct Plugin;
struct Map;
struct FrameContext;
struct Api;

constexpr int make_version(int major, int minor) { return (major << 16) | minor; }
constexpr int kApiVersion = make_version(4, 0);

enum ColorFamily { cfUndefined = 0, cfGray = 1, cfRGB = 2, cfYUV = 3 };
enum SampleType { stInteger = 0, stFloat = 1 };
constexpr uint32_t video_id(int family, int sample, int bits, int ssw, int ssh) {
    return ((uint32_t)family << 28) | ((uint32_t)sample << 24) | ((uint32_t)bits << 16) | ((uint32_t)ssw << 8) |
           (uint32_t)ssh;
}
constexpr uint32_t pfRGBS = video_id(cfRGB, stFloat, 32, 0, 0);
enum FilterMode { fmParallel = 0, fmParallelRequests = 1, fmUnordered = 2, fmFrameState = 3 };
enum MapPropertyError { peSuccess = 0, peUnset = 1, peType = 2, peError = 3, peIndex = 4 };
enum MapAppendMode { maReplace = 0, maAppend = 1 };
enum ActivationReason { arError = -1, arInitial = 0, arAllFramesReady = 1 };
enum DataTypeHint { dtUnknown = -1, dtBinary = 0, dtUtf8 = 1 };
enum RequestPattern { rpGeneral = 0, rpNoFrameReuse = 1, rpStrictSpatial = 2 };

struct VideoFormat {
    int colorFamily;
    int sampleType;
    int bitsPerSample;
    int bytesPerSample;
    int subSamplingW;
    int subSamplingH;
    int numPlanes;
};

struct VideoInfo {
    VideoFormat format;
    int64_t fpsNum;
    int64_t fpsDen;
    int width;
    int height;
    int numFrames;
};

struct CoreInfo {
    const char* versionString;
    int core;
    int api;
    int numThreads;
    int64_t maxFramebufferSize;
    int64_t usedFramebufferSize;
};

struct FilterDependency {
    Node* source;
    int requestPattern;
};

typedef void (*PublicFunction)(const Map* in, Map* out, void* userData, Core* core, const Api* vsapi);
typedef const Frame* (*FilterGetFrame)(int n, int activationReason, void* instanceData, void** frameData,
                                       FrameContext* frameCtx, Core* core, const Api* vsapi);
typedef void (*FilterFree)(void* instanceData, Core* core, const Api* vsapi);

// VSPLUGINAPI: the three entries a plugin's init function receives.
struct PluginApi {
    int (*getAPIVersion)();
    int (*configPlugin)(const char* identifier, const char* pluginNamespace, const char* name, int pluginVersion,
                        int apiVersion, int flags, Plugin* plugin);
    int (*registerFunction)(const char* name, const char* args, const char* returnType, PublicFunction argsFunc,
                            void* functionData, Plugin* plugin);
};

// Index of each VSAPI entry used by the filter.
enum Slot : int {
    createVideoFilter = 0,
    freeNode = 7,
    getVideoInfo = 10,
    newVideoFrame = 12,
    freeFrame = 16,
    copyFrame = 18,
    getFramePropertiesRW = 20,
    getStride = 21,
    getReadPtr = 22,
    getWritePtr = 23,
    getFrameWidth = 27,
    getFrameHeight = 28,
    queryVideoFormat = 32,
    getFrameFilter = 38,
    requestFrameFilter = 39,
    setFilterError = 42,
    createMap = 47,
    freeMap = 48,
    mapSetError = 51,
    mapGetError = 52,
    mapGetInt = 59,
    mapSetInt = 62,
    mapGetFloat = 64,
    mapSetFloat = 67,
    mapSetData = 72,
    mapGetNode = 73,
    mapConsumeNode = 75,
    getPluginByID = 83,
    invoke = 96,
    getCoreInfo = 101,
    kSlotsUsed = 102,
};

// Signature of each used entry.
template <Slot S>
struct Sig;
#define VS4_SIG(slot, ...)      \
    template <>                 \
    struct Sig<slot> {          \
        typedef __VA_ARGS__;    \
    }
VS4_SIG(createVideoFilter, void (*type)(Map* out, const char* name, const VideoInfo* vi, FilterGetFrame getFrame,
                                        FilterFree free, int filterMode, const FilterDependency* deps, int numDeps,
                                        void* instanceData, Core* core));
VS4_SIG(freeNode, void (*type)(Node* node));
VS4_SIG(getVideoInfo, const VideoInfo* (*type)(Node* node));
VS4_SIG(newVideoFrame,
        Frame* (*type)(const VideoFormat* format, int width, int height, const Frame* propSrc, Core* core));
VS4_SIG(freeFrame, void (*type)(const Frame* f));
VS4_SIG(copyFrame, Frame* (*type)(const Frame* f, Core* core));
VS4_SIG(getFramePropertiesRW, Map* (*type)(Frame* f));
VS4_SIG(getStride, ptrdiff_t (*type)(const Frame* f, int plane));
VS4_SIG(getReadPtr, const uint8_t* (*type)(const Frame* f, int plane));
VS4_SIG(getWritePtr, uint8_t* (*type)(Frame* f, int plane));
VS4_SIG(getFrameWidth, int (*type)(const Frame* f, int plane));
VS4_SIG(getFrameHeight, int (*type)(const Frame* f, int plane));
VS4_SIG(queryVideoFormat, int (*type)(VideoFormat* format, int colorFamily, int sampleType, int bitsPerSample,
                                       int subSamplingW, int subSamplingH, Core* core));
VS4_SIG(getFrameFilter, const Frame* (*type)(int n, Node* node, FrameContext* frameCtx));
VS4_SIG(requestFrameFilter, void (*type)(int n, Node* node, FrameContext* frameCtx));
VS4_SIG(setFilterError, void (*type)(const char* errorMessage, FrameContext* frameCtx));
VS4_SIG(createMap, Map* (*type)());
VS4_SIG(freeMap, void (*type)(Map* map));
VS4_SIG(mapSetError, void (*type)(Map* map, const char* errorMessage));
VS4_SIG(mapGetError, const char* (*type)(const Map* map));
VS4_SIG(mapGetInt, int64_t (*type)(const Map* map, const char* key, int index, int* error));
VS4_SIG(mapSetInt, int (*type)(Map* map, const char* key, int64_t i, int append));
VS4_SIG(mapGetFloat, double (*type)(const Map* map, const char* key, int index, int* error));
VS4_SIG(mapSetFloat, int (*type)(Map* map, const char* key, double d, int append));
VS4_SIG(mapSetData, int (*type)(Map* map, const char* key, const char* data, int size, int type, int append));
VS4_SIG(mapGetNode, Node* (*type)(const Map* map, const char* key, int index, int* error));
VS4_SIG(mapConsumeNode, int (*type)(Map* map, const char* key, Node* node, int append));
VS4_SIG(getPluginByID, Plugin* (*type)(const char* identifier, Core* core));
VS4_SIG(invoke, Map* (*type)(Plugin* plugin, const char* name, const Map* args));
VS4_SIG(getCoreInfo, void (*type)(Core* core, CoreInfo* info));
#undef VS4_SIG

struct Api {
    void (*slot[kSlotsUsed])();
    template <Slot S, typename... A>
    auto call(A... a) const {
        return reinterpret_cast<typename Sig<S>::type>(slot[S])(a...);
    }
};

}  // namespace vs4
