// C-ABI shim around the reference's own SSIMULACRA2 HIP implementation.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile (target "ref") from
// the reference's sources where they lie (src/ssimu2/*.hpp, src/util/*.hpp);
// nothing of the reference is copied here. The resulting
// oracle/_ref/libvship_ref.so runs the reference's kernels unchanged on the
// GPU so the GPU tests can check the CPU oracle and this library against the
// reference's own output on the same inputs.
//
// vref_ssimu2: full score through SSIMU2ComputingImplementation::run
//   (src/ssimu2/main.hpp:153-191), as Vship_ComputeSSIMU2* does.
// vref_ssimu2_features: the same stages (src/ssimu2/main.hpp:106-151,45-98)
//   driven step by step so the 108 features can be read back.
// vref_ssimu2_time: one handle (init once), `iters` calls of run() on the
//   same host planes, as an ABI client of Vship_ComputeSSIMU2* does; returns
//   the wall time of the calls in milliseconds.
#include <chrono>

#include "ssimu2/main.hpp"

namespace {

template <InputMemType FMT>
int features_impl(const uint8_t* p1[3], const uint8_t* p2[3], int64_t s1, int64_t s2, int64_t w,
                  int64_t h, float* feat108) {
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) return 1;
    ssimu2::GaussianHandle gh;
    gh.init();
    int device;
    hipDeviceProp_t prop;
    hipGetDevice(&device);
    hipGetDeviceProperties(&prop, device);
    const int64_t maxshared = prop.sharedMemPerBlock;
    float3* pinned = nullptr;
    if (hipHostMalloc(&pinned, sizeof(float3) * ssimu2::allocsizeScore(w, h, maxshared)) != hipSuccess) return 2;
    const int64_t T = ssimu2::getTotalScaleSize(w, h);
    const int64_t ms = std::max(s1, s2);
    float3* mem = nullptr;
    if (hipMalloc(&mem, sizeof(float3) * T * 2 + std::max((int64_t)sizeof(float3) * T, ms * h * 3)) != hipSuccess)
        return 3;
    float3* a = mem;
    float3* b = mem + T;
    float3* tmp = mem + 2 * T;
    uint8_t* stage[3] = {(uint8_t*)tmp, (uint8_t*)tmp + ms * h, (uint8_t*)tmp + 2 * ms * h};
    for (int i = 0; i < 3; i++) hipMemcpyHtoDAsync(stage[i], (void*)p1[i], s1 * h, stream);
    ssimu2::memoryorganizer<FMT>(a, stage[0], stage[1], stage[2], s1, w, h, stream);
    for (int i = 0; i < 3; i++) hipMemcpyHtoDAsync(stage[i], (void*)p2[i], s2 * h, stream);
    ssimu2::memoryorganizer<FMT>(b, stage[0], stage[1], stage[2], s2, w, h, stream);
    rgb_to_linear(a, T, stream);
    rgb_to_linear(b, T, stream);
    int64_t nw = w, nh = h, idx = 0;
    for (int s = 1; s <= 5; s++) {
        ssimu2::downsample(a + idx, a + idx + nw * nh, nw, nh, stream);
        ssimu2::downsample(b + idx, b + idx + nw * nh, nw, nh, stream);
        idx += nw * nh;
        nw = (nw - 1) / 2 + 1;
        nh = (nh - 1) / 2 + 1;
    }
    rgb_to_positive_xyb(a, T, stream);
    rgb_to_positive_xyb(b, T, stream);
    std::vector<float3> res = ssimu2::allscore_map(a, b, tmp, pinned, w, h, maxshared, gh, stream);
    for (int plane = 0; plane < 3; plane++)
        for (int scale = 0; scale < 6; scale++)
            for (int n = 0; n < 2; n++)
                for (int i = 0; i < 3; i++) {
                    const float3 v = res[scale * 6 + i * 2 + n];
                    feat108[plane * 36 + scale * 6 + n * 3 + i] = plane == 0 ? v.x : (plane == 1 ? v.y : v.z);
                }
    hipStreamSynchronize(stream);
    hipFree(mem);
    hipHostFree(pinned);
    gh.destroy();
    hipStreamDestroy(stream);
    return 0;
}

}  // namespace

extern "C" {

// fmt: 0 = u8, 1 = u16, 2 = f16, 3 = f32 planar RGB (src/util/preprocessor.hpp:94)
int vref_ssimu2(int fmt, const uint8_t* p1[3], const uint8_t* p2[3], int64_t s1, int64_t s2, int w, int h,
                double* score) {
    try {
        ssimu2::SSIMU2ComputingImplementation impl;
        impl.init(w, h);
        double r = 0;
        switch (fmt) {
        case 0: r = impl.run<InputMemType::UINT8>(p1, p2, s1, s2); break;
        case 1: r = impl.run<InputMemType::UINT16>(p1, p2, s1, s2); break;
        case 2: r = impl.run<InputMemType::HALF>(p1, p2, s1, s2); break;
        default: r = impl.run<InputMemType::FLOAT>(p1, p2, s1, s2); break;
        }
        impl.destroy();
        *score = r;
    } catch (const VshipError& e) {
        return (int)e.type;
    }
    return 0;
}

int vref_ssimu2_features(int fmt, const uint8_t* p1[3], const uint8_t* p2[3], int64_t s1, int64_t s2, int w,
                         int h, float* feat108) {
    try {
        switch (fmt) {
        case 0: return features_impl<InputMemType::UINT8>(p1, p2, s1, s2, w, h, feat108);
        case 1: return features_impl<InputMemType::UINT16>(p1, p2, s1, s2, w, h, feat108);
        case 2: return features_impl<InputMemType::HALF>(p1, p2, s1, s2, w, h, feat108);
        default: return features_impl<InputMemType::FLOAT>(p1, p2, s1, s2, w, h, feat108);
        }
    } catch (const VshipError& e) {
        return (int)e.type;
    }
}

int vref_ssimu2_time(int fmt, const uint8_t* p1[3], const uint8_t* p2[3], int64_t s1, int64_t s2, int w, int h,
                     int iters, double* score, double* ms) {
    try {
        ssimu2::SSIMU2ComputingImplementation impl;
        impl.init(w, h);
        double r = 0;
        auto once = [&]() {
            switch (fmt) {
            case 0: r = impl.run<InputMemType::UINT8>(p1, p2, s1, s2); break;
            case 1: r = impl.run<InputMemType::UINT16>(p1, p2, s1, s2); break;
            case 2: r = impl.run<InputMemType::HALF>(p1, p2, s1, s2); break;
            default: r = impl.run<InputMemType::FLOAT>(p1, p2, s1, s2); break;
            }
        };
        once();  // warm-up
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < iters; i++) once();
        const auto t1 = std::chrono::steady_clock::now();
        impl.destroy();
        *score = r;
        *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    } catch (const VshipError& e) {
        return (int)e.type;
    }
    return 0;
}

}  // extern "C"
