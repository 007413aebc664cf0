// Allocation-free SSIMULACRA2 entry point, Vship_SSIMU2ComputeNoAlloc.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc). It scores
// one device-resident pair with the kernels and launch shape of
// Vship_SSIMU2BatchCompute at its defaults (matrix-core map kernel, one
// workgroup per tile, FP32 planes), so its results are the batch entry
// point's bit for bit. Instead of an engine's allocations it uses `scratch`,
// one static device array per GPU the library runs on, held until the
// process ends:
//
//   [descriptor block, kDescBytes][map tile partials, kMaxTiles x 18 doubles]
//   [XYB arena, kArenaBytes: one pair of up to 1920 x 1080]
//
// Every call on a GPU uses that GPU's array and leaves its contents behind;
// nothing in it has to start at zero (the descriptor block is written by a
// kernel from its argument at the start of every call, the prep kernel writes
// every XYB plane the map kernels read, each map tile writes its partials).
// Calls that overlap on one GPU race on the array: the caller orders them
// (one stream, or events).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "VshipSSIMU2Batch.h"
#include "engine.hpp"
#include "ssimu2_kernels.hpp"

namespace vship {
namespace {

constexpr int64_t kDescBytes = 2048;
constexpr int64_t kMaxTiles = 4096;
constexpr int64_t kArenaBytes = 64ll << 20;  // 1920 x 1080 needs 66.6 MB (63.5 MiB)
constexpr int64_t kOffPartials = kDescBytes;
constexpr int64_t kOffArena = kOffPartials + kMaxTiles * kFeat * (int64_t)sizeof(double);
constexpr int64_t kScratchBytes = kOffArena + kArenaBytes;
static_assert(kOffArena % 256 == 0, "arena offset keeps the planes' 256-byte alignment");
static_assert(kScratchBytes == VSHIP_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// The descriptor block travels as a kernel argument, so that no host-to-device
// copy (and no pinned staging) is needed.
struct DescArg {
    uint32_t w[kDescBytes / 4];
};

__global__ __launch_bounds__(256) void noalloc_desc_kernel(DescArg a, uint32_t* __restrict__ dst, int nwords) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = a.w[i];
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices] = {};
Taps g_taps{};
bool g_taps_ready = false;

// The device address of this GPU's scratch array, looked up once per device
// with that device current. In host code `scratch` names a host-side stand-in,
// never the device memory; only this address is passed to kernels.
uint8_t* scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) VSHIP_THROW(E_BadDeviceArgument);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_taps_ready) {
        g_taps = make_taps();
        g_taps_ready = true;
    }
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) VSHIP_THROW(E_BadDeviceCode);
        g_scratch[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch[dev];
}

void compute_noalloc(const Vship_SSIMU2Pair& pair, double* score, double* features, hipStream_t st) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) VSHIP_THROW(E_DeviceCountError);
    uint8_t* base = scratch_address(dev);
    // outputs: device memory, checked like the input planes
    require_device_rows(score, sizeof(double), 1, sizeof(double));
    if (features) require_device_rows(features, 108 * sizeof(double), 1, 108 * sizeof(double));
    PairInput in{};
    for (int i = 0; i < 3; i++) {
        in.src[0][i] = pair.ref[i];
        in.src[1][i] = pair.dist[i];
        in.stride[0][i] = pair.ref_stride[i];
        in.stride[1][i] = pair.dist_stride[i];
    }
    in.w = pair.width;
    in.h = pair.height;
    in.fmt = pair.format;
    in.bits = pair.bits;
    in.matrix = pair.matrix;
    in.chroma_loc = pair.chroma_location;
    const DescLayout L = desc_layout(1);
    if ((int64_t)L.bytes > kDescBytes) VSHIP_THROW(E_BadDeviceCode);
    DescArg arg;
    std::memset(&arg, 0, sizeof(arg));
    uint8_t* block = reinterpret_cast<uint8_t*>(arg.w);
    // sizes beyond the scratch (arena or tile table) throw E_OutOfVRAM here,
    // before anything is enqueued
    const PlanInfo P = plan_batch(&in, 1, 1, false, kMapMfma, kArenaBytes, kMaxTiles, block, L);
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        VSHIP_THROW(E_BadDeviceCode);
    noalloc_desc_kernel<<<1, 256, 0, st>>>(arg, reinterpret_cast<uint32_t*>(base), (int)(L.bytes / 4));
    if (hipGetLastError() != hipSuccess) VSHIP_THROW(E_BadDeviceCode);
    LaunchShape S;
    S.fp16 = false;
    S.map_th = kMapMfma;
    S.persistent_grid = 0;
    S.graph = false;
    S.max_pairs = 1;
    S.num_cus = ncu;
    launch_batch(desc_view(base, L), P, S, base + kOffArena, reinterpret_cast<double*>(base + kOffPartials), score,
                 features, g_taps, st, nullptr);
}

}  // namespace
}  // namespace vship

extern "C" {

Vship_Exception Vship_SSIMU2ComputeNoAlloc(const Vship_SSIMU2Pair* pair, double* score, double* features,
                                           void* stream) {
    if (!pair || !score) return Vship_BadPointer;
    try {
        vship::compute_noalloc(*pair, score, features, static_cast<hipStream_t>(stream));
    } catch (const vship::VshipErr& e) {
        return (Vship_Exception)e.code;
    } catch (...) {
        return Vship_BadErrorType;
    }
    return Vship_NoError;
}

}  // extern "C"
