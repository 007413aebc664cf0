// C ABI of libvship.so: the reference's VshipAPI.h entry points
// (src/VshipLib.cpp:71-250 of the reference) and the additive batched
// entry points of VshipSSIMU2Batch.h. No C++ exception crosses this boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/VshipAPI.h"
#include "../../include/VshipButteraugliBatch.h"
#include "../../include/VshipSSIMU2Batch.h"
#include "api_internal.hpp"
#include "engine.hpp"
#include "ssimu2_kernels.hpp"

using namespace vship;

namespace vship {

// Message texts of the reference (src/util/VshipExceptions.hpp:41-86); callers
// parse them, so they are kept word for word.
const char* error_message(int code) {
    switch (code) {
    case E_NoError: return "NoError: Vship succeeded";
    case E_OutOfVRAM:
        return "OutOfVRAM: Vship was not able to perform GPU memory allocation. (Advice) Reduce or Set numStream "
               "argument";
    case E_OutOfRAM:
        return "OutOfRAM: Vship was not able to allocate CPU memory. This is a rare error that should be reported. "
               "Check your RAM usage";
    case E_DifferingInputType:
        return "DifferingInputType: Vship received 2 videos with different properties. (Advice) verify that they "
               "have the same width, height and length";
    case E_NonRGBSInput:
        return "NonRGBSInput: Vship did not manage to get RGBS format of your inputs. This should not happen. "
               "(Advice) try converting yourself to RGBS";
    case E_DeviceCountError:
        return "DeviceCountError: Vship was unable to verify the number of GPU on your system. (Advice) Did you "
               "select the correct binary for your device AMD/NVIDIA. (Advice) if linux AMD, are you in video and "
               "render groups?";
    case E_NoDeviceDetected:
        return "NoDeviceDetected: Vship found no device on your system. (Advice) Did you select the correct binary "
               "for your device AMD/NVIDIA. (Advice) if linux AMD, are you in video and render groups?";
    case E_BadDeviceArgument:
        return "BadDeviceArgument: Vship received a bad gpu_id argument either you specified a number >= to your "
               "gpu count, either it was negative";
    case E_BadDeviceCode:
        return "BadDeviceCode: Vship was unable to run a simple GPU Kernel. This usually indicate that the code was "
               "compiled for the wrong architecture. (Advice) Try to compile vship yourself, eventually replace "
               "--offload-arch=native to your arch";
    case E_BadHandler:
        return "BadHandler: The handler used was not allocated. This indicates an issue in Vship API Usage. Did you "
               "call the Init function or altered the handler id?";
    case E_BadPointer:
        return "BadPointer: The pointer passed as an argument was not valid and returned an error. (Advice) If this "
               "happened while trying to free, was it allocated using the same method and not modified?";
    default: return "BadErrorType: There was an unknown error";
    }
}

int device_count_or_throw() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
    if (n == 0) VSHIP_THROW(E_NoDeviceDetected);
    return n;
}

// Runs a one-thread kernel that writes a known value (src/util/gpuhelper.hpp:40-65).
bool gpu_kernel_check() {
    int* d = nullptr;
    int h = 0;
    if (hipMalloc(&d, sizeof(int)) != hipSuccess) return false;
    bool ok = hipMemset(d, 0, sizeof(int)) == hipSuccess && launch_check(d, nullptr) == hipSuccess &&
              hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(d);
    return ok && h == 4320984;
}

void gpu_full_check(int gpu_id) {
    const int n = device_count_or_throw();
    if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(E_BadDeviceArgument);
    if (hipSetDevice(gpu_id) != hipSuccess) VSHIP_THROW(E_BadDeviceArgument);
    if (!gpu_kernel_check()) VSHIP_THROW(E_BadDeviceCode);
}

// ---------------------------------------------------------------------------
// per-frame handles (Vship_SSIMU2Handler)
// ---------------------------------------------------------------------------
FrameHandle::FrameHandle(int w, int h, int device) : width(w), height(h), device(device) {
    EngineOptions o;
    o.device = device;
    o.max_pairs = 1;
    o.map_th = default_map_th(w, h);
    o.arena_bytes = pair_bytes(w, h, false);
    o.max_map_tiles = map_tiles(w, h, 32);
    // one 1080p float plane is 8.3 MB: 4 MiB chunks keep two of them in
    // flight per plane while the next one is copied in
    o.host_chunk_bytes = 4ll << 20;
    o.host_slots = 4;
    o.host_copy_threads = 1;  // frame-level parallelism comes from the callers' threads
    engine = std::make_unique<Engine>(o);
}

FrameHandle::~FrameHandle() = default;

// A size-independent default keeps a pair's result independent of the batch
// it is in (the tile decomposition, and so the per-tile centring, is fixed):
// the matrix-core map kernel.
int default_map_th(int, int) { return kMapMfma; }

double FrameHandle::compute(int fmt, const uint8_t* const p1[3], const uint8_t* const p2[3], int64_t s1, int64_t s2) {
    if (!p1 || !p2) VSHIP_THROW(E_BadPointer);
    for (int i = 0; i < 3; i++)
        if (!p1[i] || !p2[i]) VSHIP_THROW(E_BadPointer);
    const int bps = fmt == IN_RGB_F32 ? 4 : (fmt == IN_RGB_U8 ? 1 : 2);
    const int64_t row = (int64_t)width * bps;
    if (s1 < row || s2 < row) VSHIP_THROW(E_BadPointer);
    PairInput in{};
    for (int pl = 0; pl < 3; pl++) {
        in.src[0][pl] = p1[pl];
        in.src[1][pl] = p2[pl];
        in.stride[0][pl] = s1;
        in.stride[1][pl] = s2;
    }
    in.w = width;
    in.h = height;
    in.fmt = fmt;
    double score = 0.0;
    engine->submit_host(&in, 1);
    engine->wait(&score, nullptr);
    return score;
}

// ---------------------------------------------------------------------------
// Butteraugli handles (Vship_ButteraugliHandler)
// ---------------------------------------------------------------------------
ButterHandle::ButterHandle(int w, int h, int qnorm, float intensity, int dev) : width(w), height(h), device(dev) {
    engine = std::make_unique<bt::Engine>(w, h, 1, qnorm, intensity, dev);
    stager = std::make_unique<HostStager>(dev, (size_t)4 << 20, 4, 1);
    HIP_OR_THROW(hipMalloc(reinterpret_cast<void**>(&map), sizeof(float) * (size_t)w * h), E_OutOfVRAM);
}

ButterHandle::~ButterHandle() {
    if (staging) (void)hipFree(staging);
    if (map) (void)hipFree(map);
}

void ButterHandle::compute(int fmt, const uint8_t* const p1[3], const uint8_t* const p2[3], int64_t s1, int64_t s2,
                           float* dst, int64_t dst_stride, double norms[3]) {
    if (!p1 || !p2) VSHIP_THROW(E_BadPointer);
    for (int i = 0; i < 3; i++)
        if (!p1[i] || !p2[i]) VSHIP_THROW(E_BadPointer);
    const int64_t row = (int64_t)width * (fmt == IN_RGB_F32 ? 4 : 2);
    if (s1 < row || s2 < row) VSHIP_THROW(E_BadPointer);
    if (dst && dst_stride < (int64_t)width * 4) VSHIP_THROW(E_BadPointer);
    // each plane is one contiguous block of (height - 1) strides plus a row
    const size_t blk1 = (size_t)s1 * (height - 1) + row, blk2 = (size_t)s2 * (height - 1) + row;
    const size_t a1 = (blk1 + 255) / 256 * 256, a2 = (blk2 + 255) / 256 * 256;
    const size_t need = 3 * (a1 + a2);
    HIP_OR_THROW(hipSetDevice(device), E_BadDeviceArgument);
    if (need > staging_bytes) {
        if (staging) (void)hipFree(staging);
        staging = nullptr;
        staging_bytes = 0;
        HIP_OR_THROW(hipMalloc(reinterpret_cast<void**>(&staging), need), E_OutOfVRAM);
        staging_bytes = need;
    }
    hipStream_t st = engine->stream();
    bt::PairDesc d{};
    for (int im = 0; im < 2; im++) {
        const uint8_t* const* src = im == 0 ? p1 : p2;
        const int64_t sp = im == 0 ? s1 : s2;
        const size_t blk = im == 0 ? blk1 : blk2;
        for (int pl = 0; pl < 3; pl++) {
            uint8_t* dstp = staging + (im == 0 ? pl * a1 : 3 * a1 + pl * a2);
            stager->upload(dstp, src[pl], blk, st);
            d.src[im][pl] = dstp;
            d.stride[im][pl] = sp;
        }
    }
    d.map = dst ? map : nullptr;
    d.map_stride = (int64_t)width * 4;
    engine->run(fmt, &d, 1, norms);
    if (dst) stager->download_rows(dst, dst_stride, map, (int64_t)width * 4, (size_t)width * 4, height, st);
}

template <typename T>
struct Table {
    std::mutex lock;
    std::vector<std::shared_ptr<T>> items;
    std::vector<char> used;
    int allocate() {
        std::lock_guard<std::mutex> g(lock);
        for (size_t i = 0; i < used.size(); i++)
            if (!used[i]) {
                used[i] = 1;
                return (int)i;
            }
        items.emplace_back();
        used.push_back(1);
        return (int)items.size() - 1;
    }
    std::shared_ptr<T> get(int id) {
        std::lock_guard<std::mutex> g(lock);
        if (id < 0 || id >= (int)items.size() || !used[id]) return nullptr;
        return items[id];
    }
    bool valid(int id) {
        std::lock_guard<std::mutex> g(lock);
        return id >= 0 && id < (int)items.size() && used[id];
    }
    void set(int id, std::shared_ptr<T> v) {
        std::lock_guard<std::mutex> g(lock);
        items[id] = std::move(v);
    }
    bool release(int id) {
        std::shared_ptr<T> v;
        {
            std::lock_guard<std::mutex> g(lock);
            if (id < 0 || id >= (int)items.size() || !used[id]) return false;
            v = std::move(items[id]);
            used[id] = 0;
        }
        return true;  // v is destroyed outside the lock
    }
};

Table<FrameHandle>& frame_table() {
    static Table<FrameHandle> t;
    return t;
}

Table<ButterHandle>& butter_table() {
    static Table<ButterHandle> t;
    return t;
}

struct ButterBatchHandle {
    std::unique_ptr<bt::Engine> engine;
    std::atomic<bool> busy{false};
};

Table<ButterBatchHandle>& butter_batch_table() {
    static Table<ButterBatchHandle> t;
    return t;
}

struct BatchHandle {
    std::unique_ptr<Engine> engine;
    std::atomic<bool> busy{false};
};

Table<BatchHandle>& batch_table() {
    static Table<BatchHandle> t;
    return t;
}

}  // namespace vship

namespace {

struct BusyGuard {
    std::atomic<bool>& f;
    bool ok;
    explicit BusyGuard(std::atomic<bool>& flag) : f(flag), ok(!flag.exchange(true)) {}
    ~BusyGuard() {
        if (ok) f.store(false);
    }
};

template <typename F>
Vship_Exception guarded(F&& fn) {
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

Vship_Exception compute_frame(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3],
                              const uint8_t* srcp2[3], int64_t s1, int64_t s2, int fmt) {
    std::shared_ptr<FrameHandle> h = frame_table().get(handler.id);
    if (!h || !h->engine) return Vship_BadHandler;
    if (!score) return Vship_BadPointer;
    BusyGuard g(h->busy);
    if (!g.ok) return Vship_BadHandler;
    double r = 0.0;
    const Vship_Exception e = guarded([&] { r = h->compute(fmt, srcp1, srcp2, s1, s2); });
    if (e == Vship_NoError) *score = r;
    return e;
}

// Vship_ComputeButteraugli*: the norms are stored into the score's int
// fields by plain conversion, truncating, as the reference does
// (src/VshipLib.cpp:308-310, where a float tuple is assigned to them).
Vship_Exception compute_butter(Vship_ButteraugliHandler handler, Vship_ButteraugliScore* score, const uint8_t* dstp,
                               int64_t dststride, const uint8_t* srcp1[3], const uint8_t* srcp2[3], int64_t s1,
                               int64_t s2, int fmt) {
    std::shared_ptr<ButterHandle> h = butter_table().get(handler.id);
    if (!h || !h->engine) return Vship_BadHandler;
    if (!score) return Vship_BadPointer;
    BusyGuard g(h->busy);
    if (!g.ok) return Vship_BadHandler;
    double n[3] = {0.0, 0.0, 0.0};
    const Vship_Exception e = guarded([&] {
        h->compute(fmt, srcp1, srcp2, s1, s2, reinterpret_cast<float*>(const_cast<uint8_t*>(dstp)), dststride, n);
    });
    if (e == Vship_NoError) {
        score->norm2 = (int)(float)n[0];
        score->norm3 = (int)(float)n[1];
        score->norminf = (int)(float)n[2];
    }
    return e;
}

PairInput to_pair_input(const Vship_SSIMU2Pair& p) {
    PairInput in{};
    for (int i = 0; i < 3; i++) {
        in.src[0][i] = p.ref[i];
        in.src[1][i] = p.dist[i];
        in.stride[0][i] = p.ref_stride[i];
        in.stride[1][i] = p.dist_stride[i];
    }
    in.w = p.width;
    in.h = p.height;
    in.fmt = p.format;
    in.bits = p.bits;
    in.matrix = p.matrix;
    in.chroma_loc = p.chroma_location;
    return in;
}

}  // namespace

extern "C" {

Vship_Version Vship_GetVersion(void) {
    Vship_Version v;
    v.major = 3;
    v.minor = 1;
    v.minorMinor = 0;
    v.backend = Vship_HIP;
    return v;
}

Vship_Exception Vship_GetDeviceCount(int* number) {
    if (!number) return Vship_BadPointer;
    int n = 0;
    const Vship_Exception e = guarded([&] { n = device_count_or_throw(); });
    if (e == Vship_NoError) *number = n;
    return e;
}

Vship_Exception Vship_SetDevice(int gpu_id) {
    return guarded([&] {
        const int n = device_count_or_throw();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(E_BadDeviceArgument);
        if (hipSetDevice(gpu_id) != hipSuccess) VSHIP_THROW(E_BadDeviceArgument);
    });
}

Vship_Exception Vship_GetDeviceInfo(Vship_DeviceInfo* info, int gpu_id) {
    if (!info) return Vship_BadPointer;
    return guarded([&] {
        const int n = device_count_or_throw();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(E_BadDeviceArgument);
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, gpu_id) != hipSuccess) VSHIP_THROW(E_BadDeviceArgument);
        std::memset(info->name, 0, sizeof(info->name));
        std::memcpy(info->name, p.name, std::min(sizeof(info->name) - 1, sizeof(p.name)));
        info->VRAMSize = p.totalGlobalMem;
        info->integrated = p.integrated;
        info->MultiProcessorCount = p.multiProcessorCount;
        info->WarpSize = p.warpSize;
    });
}

Vship_Exception Vship_GPUFullCheck(int gpu_id) {
    return guarded([&] { gpu_full_check(gpu_id); });
}

int Vship_GetErrorMessage(Vship_Exception exception, char* out_message, int len) {
    const char* msg = error_message((int)exception);
    const int size = (int)std::strlen(msg);
    if (len <= 0 || !out_message) return size + 1;
    const int n = std::min(len - 1, size);
    std::memcpy(out_message, msg, n);
    out_message[n] = '\0';
    return size + 1;
}

Vship_Exception Vship_PinnedMalloc(void** ptr, uint64_t size) {
    if (!ptr) return Vship_BadPointer;
    if (hipHostMalloc(ptr, size, hipHostMallocDefault) != hipSuccess) return Vship_OutOfRAM;
    return Vship_NoError;
}

Vship_Exception Vship_PinnedFree(void* ptr) {
    if (hipHostFree(ptr) != hipSuccess) return Vship_BadPointer;
    return Vship_NoError;
}

Vship_Exception Vship_SSIMU2Init(Vship_SSIMU2Handler* handler, int width, int height) {
    if (!handler) return Vship_BadPointer;
    // the slot is taken even when initialisation fails, as in the reference
    // (src/VshipLib.cpp:162-173); Vship_SSIMU2Free releases it
    handler->id = frame_table().allocate();
    if (width <= 0 || height <= 0) return Vship_OutOfVRAM;
    return guarded([&] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
        frame_table().set(handler->id, std::make_shared<FrameHandle>(width, height, dev));
    });
}

Vship_Exception Vship_SSIMU2Free(Vship_SSIMU2Handler handler) {
    return frame_table().release(handler.id) ? Vship_NoError : Vship_BadHandler;
}

Vship_Exception Vship_ComputeSSIMU2Float(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3],
                                         const uint8_t* srcp2[3], int64_t stride) {
    return compute_frame(handler, score, srcp1, srcp2, stride, stride, IN_RGB_F32);
}
Vship_Exception Vship_ComputeSSIMU2Floatv2(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3],
                                           const uint8_t* srcp2[3], int64_t stride, int64_t stride2) {
    return compute_frame(handler, score, srcp1, srcp2, stride, stride2, IN_RGB_F32);
}
Vship_Exception Vship_ComputeSSIMU2Uint16(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3],
                                          const uint8_t* srcp2[3], int64_t stride) {
    return compute_frame(handler, score, srcp1, srcp2, stride, stride, IN_RGB_U16);
}
Vship_Exception Vship_ComputeSSIMU2Uint16v2(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3],
                                            const uint8_t* srcp2[3], int64_t stride, int64_t stride2) {
    return compute_frame(handler, score, srcp1, srcp2, stride, stride2, IN_RGB_U16);
}

Vship_Exception Vship_ButteraugliInit(Vship_ButteraugliHandler* handler, int width, int height,
                                      float intensity_multiplier) {
    return Vship_ButteraugliInitv2(handler, width, height, 2, intensity_multiplier);
}

Vship_Exception Vship_ButteraugliInitv2(Vship_ButteraugliHandler* handler, int width, int height, int Qnorm,
                                        float intensity_multiplier) {
    if (!handler) return Vship_BadPointer;
    // the slot is taken even when initialisation fails, as in the reference
    // (src/VshipLib.cpp:267-280); Vship_ButteraugliFree releases it
    handler->id = butter_table().allocate();
    if (width <= 0 || height <= 0) return Vship_OutOfVRAM;
    return guarded([&] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
        butter_table().set(handler->id,
                           std::make_shared<ButterHandle>(width, height, Qnorm, intensity_multiplier, dev));
    });
}

Vship_Exception Vship_ButteraugliFree(Vship_ButteraugliHandler handler) {
    return butter_table().release(handler.id) ? Vship_NoError : Vship_BadHandler;
}

Vship_Exception Vship_ComputeButteraugliFloat(Vship_ButteraugliHandler handler, Vship_ButteraugliScore* score,
                                              const uint8_t* dstp, int64_t dststride, const uint8_t* srcp1[3],
                                              const uint8_t* srcp2[3], int64_t stride) {
    return compute_butter(handler, score, dstp, dststride, srcp1, srcp2, stride, stride, IN_RGB_F32);
}
Vship_Exception Vship_ComputeButteraugliFloatv2(Vship_ButteraugliHandler handler, Vship_ButteraugliScore* score,
                                                const uint8_t* dstp, int64_t dststride, const uint8_t* srcp1[3],
                                                const uint8_t* srcp2[3], int64_t stride, int64_t stride2) {
    return compute_butter(handler, score, dstp, dststride, srcp1, srcp2, stride, stride2, IN_RGB_F32);
}
Vship_Exception Vship_ComputeButteraugliUint16(Vship_ButteraugliHandler handler, Vship_ButteraugliScore* score,
                                               const uint8_t* dstp, int64_t dststride, const uint8_t* srcp1[3],
                                               const uint8_t* srcp2[3], int64_t stride) {
    return compute_butter(handler, score, dstp, dststride, srcp1, srcp2, stride, stride, IN_RGB_U16);
}
Vship_Exception Vship_ComputeButteraugliUint16v2(Vship_ButteraugliHandler handler, Vship_ButteraugliScore* score,
                                                 const uint8_t* dstp, int64_t dststride, const uint8_t* srcp1[3],
                                                 const uint8_t* srcp2[3], int64_t stride, int64_t stride2) {
    return compute_butter(handler, score, dstp, dststride, srcp1, srcp2, stride, stride2, IN_RGB_U16);
}

// ---------------------------------------------------------------------------
// batched Butteraugli (VshipButteraugliBatch.h)
// ---------------------------------------------------------------------------
uint64_t Vship_ButteraugliPairFootprint(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    return bt::Engine::pair_bytes(width, height);
}

Vship_Exception Vship_ButteraugliBatchInit(Vship_ButteraugliBatchHandler* handler,
                                           const Vship_ButteraugliBatchConfig* cfg) {
    if (!handler || !cfg) return Vship_BadPointer;
    if (cfg->max_pairs < 1 || cfg->max_pairs > 4096 || cfg->width < 1 || cfg->height < 1) return Vship_BadPointer;
    const int id = butter_batch_table().allocate();
    handler->id = id;
    const Vship_Exception e = guarded([&] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
        auto b = std::make_shared<ButterBatchHandle>();
        b->engine = std::make_unique<bt::Engine>(cfg->width, cfg->height, cfg->max_pairs, cfg->qnorm,
                                                 cfg->intensity_multiplier, dev);
        b->engine->set_profile(cfg->profile != 0);
        butter_batch_table().set(id, std::move(b));
    });
    if (e != Vship_NoError) butter_batch_table().release(id);
    return e;
}

Vship_Exception Vship_ButteraugliBatchFree(Vship_ButteraugliBatchHandler handler) {
    return butter_batch_table().release(handler.id) ? Vship_NoError : Vship_BadHandler;
}

Vship_Exception Vship_ButteraugliBatchCompute(Vship_ButteraugliBatchHandler handler, int format,
                                              const Vship_ButteraugliPair* pairs, int npairs, double* norms) {
    std::shared_ptr<ButterBatchHandle> b = butter_batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    if (!pairs || !norms || npairs < 1) return Vship_BadPointer;
    if (format < 0 || format >= IN_NUM_FORMATS) return Vship_BadPointer;
    BusyGuard g(b->busy);
    if (!g.ok) return Vship_BadHandler;
    return guarded([&] {
        std::vector<bt::PairDesc> d(npairs);
        for (int i = 0; i < npairs; i++) {
            for (int c = 0; c < 3; c++) {
                d[i].src[0][c] = pairs[i].ref[c];
                d[i].src[1][c] = pairs[i].dist[c];
                d[i].stride[0][c] = pairs[i].ref_stride[c];
                d[i].stride[1][c] = pairs[i].dist_stride[c];
            }
            d[i].map = pairs[i].distmap;
            d[i].map_stride = pairs[i].distmap_stride;
            // bits counts only for the 16-bit container formats; 8-bit ones are 8
            d[i].bits = fmt_wide(format) ? pairs[i].bits : 8;
            d[i].cloc = pairs[i].chroma_location;
            matrix_coefs(pairs[i].matrix, d[i].kr, d[i].kb);
        }
        b->engine->run(format, d.data(), npairs, norms);
    });
}

Vship_Exception Vship_ButteraugliBatchGetStats(Vship_ButteraugliBatchHandler handler,
                                               Vship_ButteraugliBatchStats* st) {
    std::shared_ptr<ButterBatchHandle> b = butter_batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    if (!st) return Vship_BadPointer;
    // the engine's timing state belongs to a compute in flight on this handle
    BusyGuard g(b->busy);
    if (!g.ok) return Vship_BadHandler;
    std::memset(st, 0, sizeof(*st));
    const bt::Engine& e = *b->engine;
    float ms[bt::ST_COUNT] = {};
    st->timed = e.stage_ms(ms) ? 1 : 0;
    st->input_ms = ms[bt::ST_INPUT];
    st->opsin_ms = ms[bt::ST_OPSIN];
    st->low_ms = ms[bt::ST_LOW];
    st->mid_ms = ms[bt::ST_MID];
    st->high_ms = ms[bt::ST_HIGH];
    st->diff_ms = ms[bt::ST_DIFF];
    st->mask_ms = ms[bt::ST_MASK];
    st->norm_ms = ms[bt::ST_NORM];
    st->npairs = e.last_npairs();
    for (int s = 0; s < bt::ST_COUNT; s++) st->stage_bytes[s] = e.stage_bytes(s) * (uint64_t)e.last_npairs();
    return Vship_NoError;
}

// ---------------------------------------------------------------------------
// batched entry points
// ---------------------------------------------------------------------------
uint64_t Vship_SSIMU2PairFootprint(int width, int height, int fp16_planes) {
    if (width <= 0 || height <= 0) return 0;
    return (uint64_t)pair_bytes(width, height, fp16_planes != 0);
}

Vship_Exception Vship_SSIMU2BatchInit(Vship_SSIMU2BatchHandler* handler, const Vship_SSIMU2BatchConfig* cfg) {
    if (!handler || !cfg) return Vship_BadPointer;
    if (cfg->max_pairs < 1 || cfg->max_width < 1 || cfg->max_height < 1) return Vship_BadPointer;
    const int id = batch_table().allocate();
    handler->id = id;
    const Vship_Exception e = guarded([&] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
        EngineOptions o;
        o.device = dev;
        o.max_pairs = cfg->max_pairs;
        o.fp16 = cfg->fp16_planes != 0;
        o.map_th = cfg->map_tile_height == 48 || cfg->map_tile_height == 64 || cfg->map_tile_height == 112 ||
                           cfg->map_tile_height == kMapMfma
                       ? cfg->map_tile_height
                       : default_map_th(cfg->max_width, cfg->max_height);
        o.persistent_grid = cfg->persistent_grid;
        int tth = 0, tgrid = 0, tgraph = 0;
        if (cfg->map_tile_height == 0 && cfg->persistent_grid == 0 &&
            tuned_shape(cfg->max_width, cfg->max_height, cfg->fp16_planes != 0, &tth, &tgrid, &tgraph)) {
            o.map_th = tth;
            o.persistent_grid = tgrid;
        }
        o.arena_bytes = cfg->arena_bytes ? (int64_t)cfg->arena_bytes
                                         : (int64_t)cfg->max_pairs *
                                               (int64_t)Vship_SSIMU2PairFootprint(cfg->max_width, cfg->max_height,
                                                                                  cfg->fp16_planes);
        o.max_map_tiles = (int64_t)cfg->max_pairs * map_tiles(cfg->max_width, cfg->max_height, 32);
        o.use_graph = cfg->use_graph != 0;  // a captured graph runs a fixed grid (4 x CUs when 0)
        o.profile = cfg->profile != 0;
        o.host_copy_threads = 4;
        auto b = std::make_shared<BatchHandle>();
        b->engine = std::make_unique<Engine>(o);
        batch_table().set(id, std::move(b));
    });
    if (e != Vship_NoError) batch_table().release(id);
    return e;
}

Vship_Exception Vship_SSIMU2BatchFree(Vship_SSIMU2BatchHandler handler) {
    return batch_table().release(handler.id) ? Vship_NoError : Vship_BadHandler;
}

Vship_Exception Vship_SSIMU2BatchSubmit(Vship_SSIMU2BatchHandler handler, const Vship_SSIMU2Pair* pairs,
                                        int npairs) {
    std::shared_ptr<BatchHandle> b = batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    if (!pairs || npairs < 1) return Vship_BadPointer;
    BusyGuard g(b->busy);
    if (!g.ok) return Vship_BadHandler;
    return guarded([&] {
        std::vector<PairInput> in(npairs);
        for (int i = 0; i < npairs; i++) in[i] = to_pair_input(pairs[i]);
        b->engine->submit(in.data(), npairs);
    });
}

Vship_Exception Vship_SSIMU2BatchSubmitHost(Vship_SSIMU2BatchHandler handler, const Vship_SSIMU2Pair* pairs,
                                            int npairs) {
    std::shared_ptr<BatchHandle> b = batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    if (!pairs || npairs < 1) return Vship_BadPointer;
    BusyGuard g(b->busy);
    if (!g.ok) return Vship_BadHandler;
    return guarded([&] {
        std::vector<PairInput> in(npairs);
        for (int i = 0; i < npairs; i++) in[i] = to_pair_input(pairs[i]);
        b->engine->submit_host(in.data(), npairs);
    });
}

Vship_Exception Vship_SSIMU2BatchHostStaging(Vship_SSIMU2BatchHandler handler, uint64_t chunk_bytes, int slots,
                                             int copy_threads) {
    std::shared_ptr<BatchHandle> b = batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    if (chunk_bytes < 4096 || chunk_bytes > (1ull << 30) || slots < 2 || slots > 64 || copy_threads < 1 ||
        copy_threads > 64)
        return Vship_BadPointer;
    BusyGuard g(b->busy);
    if (!g.ok) return Vship_BadHandler;
    return guarded([&] { b->engine->set_host_staging((int64_t)chunk_bytes, slots, copy_threads); });
}

Vship_Exception Vship_SSIMU2BatchWait(Vship_SSIMU2BatchHandler handler, double* scores, double* features) {
    std::shared_ptr<BatchHandle> b = batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    BusyGuard g(b->busy);
    if (!g.ok) return Vship_BadHandler;
    return guarded([&] { b->engine->wait(scores, features); });
}

Vship_Exception Vship_SSIMU2BatchCompute(Vship_SSIMU2BatchHandler handler, const Vship_SSIMU2Pair* pairs,
                                         int npairs, double* scores, double* features) {
    {
        // Wait returns the oldest outstanding batch: with one already in
        // flight, Compute would hand back that batch's results
        std::shared_ptr<BatchHandle> b = batch_table().get(handler.id);
        if (!b) return Vship_BadHandler;
        BusyGuard g(b->busy);
        if (!g.ok || b->engine->outstanding() != 0) return Vship_BadHandler;
    }
    const Vship_Exception e = Vship_SSIMU2BatchSubmit(handler, pairs, npairs);
    if (e != Vship_NoError) return e;
    return Vship_SSIMU2BatchWait(handler, scores, features);
}

Vship_Exception Vship_SSIMU2BatchGetStats(Vship_SSIMU2BatchHandler handler, Vship_SSIMU2BatchStats* st) {
    std::shared_ptr<BatchHandle> b = batch_table().get(handler.id);
    if (!b) return Vship_BadHandler;
    if (!st) return Vship_BadPointer;
    const Engine& e = *b->engine;
    st->prep_ms = e.timings().prep_ms;
    st->map_ms = e.timings().map_ms;
    st->map_fp32_ms = e.timings().map32_ms;
    st->map_fp64_ms = e.timings().map64_ms;
    st->finalize_ms = e.timings().finalize_ms;
    st->total_ms = e.timings().total_ms;
    st->npairs = e.last_npairs();
    st->prep_tiles = e.last_prep_tiles();
    st->map_tiles = e.last_map_tiles();
    st->map_tiles_fp32 = e.last_map_tiles_class(0);
    st->map_tiles_fp64 = e.last_map_tiles_class(1);
    st->map_tile_height = e.options().map_th;
    st->map_bytes = e.last_map_bytes();
    st->map_bytes_fp32 = e.last_map_bytes_class(0);
    st->prep_bytes = e.last_prep_bytes();
    return Vship_NoError;
}

Vship_Exception Vship_DeviceMalloc(void** ptr, uint64_t size) {
    if (!ptr) return Vship_BadPointer;
    if (hipMalloc(ptr, size) != hipSuccess) return Vship_OutOfVRAM;
    return Vship_NoError;
}

Vship_Exception Vship_DeviceFree(void* ptr) { return hipFree(ptr) == hipSuccess ? Vship_NoError : Vship_BadPointer; }

Vship_Exception Vship_MemcpyHtoD(void* dst, const void* src, uint64_t size) {
    return hipMemcpy(dst, src, size, hipMemcpyHostToDevice) == hipSuccess ? Vship_NoError : Vship_BadPointer;
}

Vship_Exception Vship_MemcpyDtoH(void* dst, const void* src, uint64_t size) {
    return hipMemcpy(dst, src, size, hipMemcpyDeviceToHost) == hipSuccess ? Vship_NoError : Vship_BadPointer;
}

Vship_Exception Vship_MemcpyDtoD(void* dst, const void* src, uint64_t size) {
    return hipMemcpy(dst, src, size, hipMemcpyDeviceToDevice) == hipSuccess ? Vship_NoError : Vship_BadPointer;
}

}  // extern "C"
