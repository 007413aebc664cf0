// Compiled next to VapourSynth's own VapourSynth4.h (API 4.0): every VSAPI
// entry, struct field and constant that vship_amd/csrc/vs4.hpp declares must
// sit at the same index / offset / value. Prints one line per mismatch and
// exits with their count.
#include <cstddef>
#include <cstdio>

#include "VapourSynth4.h"
#include "../vship_amd/csrc/vs4.hpp"

static int fails = 0;
static void check(const char* what, long long theirs, long long ours) {
    if (theirs != ours) {
        std::printf("MISMATCH %s: VapourSynth4.h %lld, vs4.hpp %lld\n", what, theirs, ours);
        fails++;
    }
}
#define SLOT(name) check("VSAPI::" #name, (long long)(offsetof(VSAPI, name) / sizeof(void*)), (long long)vs4::name)
#define FIELD(T, U, f) check(#T "::" #f, (long long)offsetof(T, f), (long long)offsetof(U, f))
#define SIZE(T, U) check("sizeof " #T, (long long)sizeof(T), (long long)sizeof(U))
#define CONST(a, b) check(#a, (long long)(a), (long long)(b))

int main() {
    SLOT(createVideoFilter); SLOT(freeNode); SLOT(getVideoInfo); SLOT(freeFrame); SLOT(copyFrame);
    SLOT(getFramePropertiesRW); SLOT(getStride); SLOT(getReadPtr); SLOT(getFrameWidth); SLOT(getFrameHeight);
    SLOT(getFrameFilter); SLOT(requestFrameFilter); SLOT(setFilterError); SLOT(createMap); SLOT(freeMap);
    SLOT(mapSetError); SLOT(mapGetError); SLOT(mapGetInt); SLOT(mapSetInt); SLOT(mapSetFloat); SLOT(mapSetData);
    SLOT(mapGetNode); SLOT(mapConsumeNode); SLOT(getPluginByID); SLOT(invoke); SLOT(getCoreInfo);
    SLOT(newVideoFrame); SLOT(getWritePtr); SLOT(queryVideoFormat); SLOT(mapGetFloat);
    FIELD(VSPLUGINAPI, vs4::PluginApi, getAPIVersion); FIELD(VSPLUGINAPI, vs4::PluginApi, configPlugin);
    FIELD(VSPLUGINAPI, vs4::PluginApi, registerFunction);
    FIELD(VSVideoFormat, vs4::VideoFormat, colorFamily); FIELD(VSVideoFormat, vs4::VideoFormat, sampleType);
    FIELD(VSVideoFormat, vs4::VideoFormat, bitsPerSample); FIELD(VSVideoFormat, vs4::VideoFormat, bytesPerSample);
    FIELD(VSVideoFormat, vs4::VideoFormat, subSamplingW); FIELD(VSVideoFormat, vs4::VideoFormat, subSamplingH);
    FIELD(VSVideoFormat, vs4::VideoFormat, numPlanes); SIZE(VSVideoFormat, vs4::VideoFormat);
    FIELD(VSVideoInfo, vs4::VideoInfo, format); FIELD(VSVideoInfo, vs4::VideoInfo, fpsNum);
    FIELD(VSVideoInfo, vs4::VideoInfo, fpsDen); FIELD(VSVideoInfo, vs4::VideoInfo, width);
    FIELD(VSVideoInfo, vs4::VideoInfo, height); FIELD(VSVideoInfo, vs4::VideoInfo, numFrames);
    SIZE(VSVideoInfo, vs4::VideoInfo);
    FIELD(VSCoreInfo, vs4::CoreInfo, versionString); FIELD(VSCoreInfo, vs4::CoreInfo, core);
    FIELD(VSCoreInfo, vs4::CoreInfo, api); FIELD(VSCoreInfo, vs4::CoreInfo, numThreads);
    FIELD(VSCoreInfo, vs4::CoreInfo, maxFramebufferSize); FIELD(VSCoreInfo, vs4::CoreInfo, usedFramebufferSize);
    SIZE(VSCoreInfo, vs4::CoreInfo);
    FIELD(VSFilterDependency, vs4::FilterDependency, source);
    FIELD(VSFilterDependency, vs4::FilterDependency, requestPattern);
    SIZE(VSFilterDependency, vs4::FilterDependency);
    CONST(VAPOURSYNTH_API_VERSION, vs4::kApiVersion); CONST(VS_MAKE_VERSION(3, 3), vs4::make_version(3, 3));
    CONST(cfGray, vs4::cfGray); CONST(cfRGB, vs4::cfRGB); CONST(cfYUV, vs4::cfYUV); CONST(stFloat, vs4::stFloat); CONST(stInteger, vs4::stInteger);
    CONST(pfRGBS, vs4::pfRGBS); CONST(fmParallel, vs4::fmParallel); CONST(peSuccess, vs4::peSuccess);
    CONST(peUnset, vs4::peUnset); CONST(maReplace, vs4::maReplace); CONST(arInitial, vs4::arInitial);
    CONST(arAllFramesReady, vs4::arAllFramesReady); CONST(dtUtf8, vs4::dtUtf8);
    CONST(rpStrictSpatial, vs4::rpStrictSpatial);
    std::printf("%d mismatches\n", fails);
    return fails;
}
