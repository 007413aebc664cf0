// C-ABI shim around the reference's own Butteraugli HIP implementation.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile (target "ref") from
// the reference's sources where they lie (src/butter/*.hpp, src/util/*.hpp);
// nothing of the reference is copied here. The resulting
// oracle/_ref/libvship_ref_butter.so runs the reference's kernels unchanged on
// the GPU so that the tests can check the CPU oracle and this library against
// the reference's own output on the same inputs.
//
// vref_butter: the three norms (Q-norm, 3-norm, inf-norm) and, if diffmap is
//   not NULL, the w x h distortion map. It makes the calls of butterprocess
//   (src/butter/main.hpp:141-195) in the same order, with two differences in
//   plumbing only:
//   * the work buffer comes from hipMalloc, freed after a device-wide
//     synchronisation, instead of hipMallocAsync/hipFreeAsync on a private
//     stream;
//   * every launch goes to the null stream. MaltaDiffMap
//     (src/butter/maltaDiff.hpp:670) launches there regardless of the stream
//     it is given, so one stream for all keeps its reads ordered with the rest.
//
// Sizes the reference's tiled blur cannot take. TiledGaussianBlur_Kernel
// (src/butter/gaussianblur.hpp:127-130, 154-157) reads
// gaussiankernel_integral[end2] with end2 as low as r - 23, where r in 1..32
// is the width (height) left for the last 32-column (row) block. For the
// sigma 1.2 table, whose integral starts 17 floats into the Gaussian buffer,
// r <= 5 reads before that buffer's start. vref_butter refuses (returns 3)
// any width or height, at either scale, whose remainder modulo 32 is 1..5;
// nothing the kernel writes depends on those reads otherwise.
#include <cstdio>

#include "butter/main.hpp"

namespace {

bool tiled_blur_safe(int64_t n) {
    const int64_t r = n % 32;
    return r == 0 || r > 5;
}

}  // namespace

extern "C" {

int vref_butter_size_ok(int w, int h) {
    const int64_t nw = (w - 1) / 2 + 1, nh = (h - 1) / 2 + 1;
    return w > 0 && h > 0 && tiled_blur_safe(w) && tiled_blur_safe(h) && tiled_blur_safe(nw) && tiled_blur_safe(nh);
}

int vref_butter(int fmt, const uint8_t* p1[3], const uint8_t* p2[3], int64_t s1, int64_t s2, int w, int h, int qnorm,
                float intensity, float* norms3, float* diffmap) {
    if (!vref_butter_size_ok(w, h)) {
        fprintf(stderr, "vref_butter: %dx%d is a size the reference's tiled blur reads out of bounds for\n", w, h);
        return 3;
    }
    if (fmt != 0 && fmt != 1) return 4;
    const int64_t width = w, height = h, wh = width * height;
    const hipStream_t stream = 0;
    butter::GaussianHandle gh;
    float* pinned = nullptr;
    float* mem = nullptr;
    try {
        gh.init();
        if (hipHostMalloc(&pinned, sizeof(float) * butter::allocsizeScore(width, height)) != hipSuccess) {
            gh.destroy();
            return 5;
        }
        const size_t bytes =
            std::max((size_t)(std::max(s1, s2) * height + 6 * sizeof(float) * wh), sizeof(float) * wh * 31);
        if (hipMalloc(&mem, bytes) != hipSuccess) {
            hipHostFree(pinned);
            gh.destroy();
            return 6;
        }
        float* src1[3] = {mem, mem + wh, mem + 2 * wh};
        float* src2[3] = {mem + 3 * wh, mem + 4 * wh, mem + 5 * wh};
        float* stage = mem + 6 * wh;
        for (int c = 0; c < 3; c++) {
            hipMemcpyHtoD(stage, (void*)p1[c], s1 * height);
            if (fmt == 0)
                butter::strideEliminator<InputMemType::FLOAT>(src1[c], stage, s1, width, height, stream);
            else
                butter::strideEliminator<InputMemType::UINT16>(src1[c], stage, s1, width, height, stream);
            hipDeviceSynchronize();
            hipMemcpyHtoD(stage, (void*)p2[c], s2 * height);
            if (fmt == 0)
                butter::strideEliminator<InputMemType::FLOAT>(src2[c], stage, s2, width, height, stream);
            else
                butter::strideEliminator<InputMemType::UINT16>(src2[c], stage, s2, width, height, stream);
            hipDeviceSynchronize();
        }
        butter::linearRGB(src1, width, height, stream);
        butter::linearRGB(src2, width, height, stream);
        int device = 0;
        hipDeviceProp_t prop;
        hipGetDevice(&device);
        hipGetDeviceProperties(&prop, device);
        float* dm = butter::getmultiscalediffmap(src1, src2, stage, width, height, intensity, prop.sharedMemPerBlock,
                                                 gh, stream);
        if (diffmap != nullptr) {
            butter::strideAdder(dm, stage, width * 4, width, height, stream);
            hipDeviceSynchronize();
            hipMemcpyDtoH(diffmap, stage, sizeof(float) * wh);
        }
        std::tuple<float, float, float> r =
            butter::diffmapscore(dm, mem + 9 * wh, mem + 10 * wh, pinned, wh, qnorm, stream);
        hipDeviceSynchronize();
        norms3[0] = std::get<0>(r);
        norms3[1] = std::get<1>(r);
        norms3[2] = std::get<2>(r);
    } catch (const VshipError& e) {
        fprintf(stderr, "vref_butter: %s", e.getErrorMessage().c_str());
        hipDeviceSynchronize();
        if (mem) hipFree(mem);
        if (pinned) hipHostFree(pinned);
        return 1;
    } catch (...) {
        fprintf(stderr, "vref_butter: unknown exception\n");
        return 2;
    }
    hipFree(mem);
    hipHostFree(pinned);
    gh.destroy();
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        fprintf(stderr, "vref_butter: %s\n", hipGetErrorString(err));
        return 7;
    }
    return 0;
}

}
