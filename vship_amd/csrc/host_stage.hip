// Host-to-device uploads through a ring of page-locked chunks (HostStager).
#include "host_stage.hpp"

#include <algorithm>
#include <cstring>

#include "engine.hpp"

namespace vship {

HostMem host_memory_kind(const void* p, size_t bytes) {
    if (!p || bytes == 0) return HostMem::Pageable;
    const uint8_t* b = static_cast<const uint8_t*>(p);
    bool locked = true;
    for (const uint8_t* q : {b, b + bytes - 1}) {
        hipPointerAttribute_t attr;
        if (hipPointerGetAttributes(&attr, q) != hipSuccess) {
            (void)hipGetLastError();  // pageable memory: the query fails and leaves its error behind
            locked = false;
            continue;
        }
        if (q == b && attr.type == hipMemoryTypeDevice) return HostMem::Device;
        if (attr.type != hipMemoryTypeHost) locked = false;
    }
    return locked ? HostMem::PageLocked : HostMem::Pageable;
}

UploadPlan plan_upload(const uint8_t* const* src, const size_t* bytes, int n) {
    constexpr size_t kAlign = 256, kMaxGap = 64 << 10;
    UploadPlan P;
    P.off.resize(n);
    size_t total = 0;
    for (int i = 0; i < n;) {
        if (!src[i] || bytes[i] == 0) VSHIP_THROW(E_BadPointer);
        const HostMem kind = host_memory_kind(src[i], bytes[i]);
        if (kind == HostMem::Device) VSHIP_THROW(E_BadPointer);
        UploadPlan::Run r{src[i], bytes[i], (total + kAlign - 1) / kAlign * kAlign, kind == HostMem::PageLocked};
        P.off[i] = r.off;
        int j = i + 1;
        if (r.locked) {
            hipDeviceptr_t base = nullptr;
            size_t size = 0;
            if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)src[i]) != hipSuccess) {
                (void)hipGetLastError();
                size = 0;
            }
            const uint8_t* lo = static_cast<const uint8_t*>(base);
            const uint8_t* hi = lo + size;
            if (size == 0 || r.src < lo || r.src + r.bytes > hi) hi = lo = nullptr;
            for (; lo && j < n; j++) {
                const uint8_t* s = src[j];
                const uint8_t* prev_end = r.src + r.bytes;
                if (!s || bytes[j] == 0 || s < prev_end || (size_t)(s - prev_end) > kMaxGap ||
                    (size_t)(s - r.src) % kAlign != 0 || s + bytes[j] > hi)
                    break;
                P.off[j] = r.off + (size_t)(s - r.src);
                r.bytes = (size_t)(s + bytes[j] - r.src);
            }
        }
        P.runs.push_back(r);
        total = r.off + r.bytes;
        i = j;
    }
    P.total = total;
    return P;
}

HostStager::HostStager(int device, size_t chunk_bytes, int slots, int threads) {
    HIP_OR_THROW(hipSetDevice(device), E_BadDeviceArgument);
    chunk_ = std::max<size_t>(chunk_bytes / 4096 * 4096, 4096);
    slots = std::max(slots, 2);
    try {
        for (int i = 0; i < slots; i++) {
            void* p = nullptr;
            HIP_OR_THROW(hipHostMalloc(&p, chunk_, hipHostMallocDefault), E_OutOfRAM);
            ring_.push_back(static_cast<uint8_t*>(p));
            hipEvent_t e = nullptr;
            HIP_OR_THROW(hipEventCreateWithFlags(&e, hipEventDisableTiming), E_OutOfRAM);
            ev_.push_back(e);
        }
        if (threads > 1) pool_ = new ParallelCopy(threads);
    } catch (...) {
        release();
        throw;
    }
}

HostStager::~HostStager() { release(); }

void HostStager::release() {
    for (hipEvent_t e : ev_) {
        (void)hipEventSynchronize(e);
        (void)hipEventDestroy(e);
    }
    for (uint8_t* p : ring_) (void)hipHostFree(p);
    ev_.clear();
    ring_.clear();
    delete pool_;
    pool_ = nullptr;
}

void HostStager::upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    if (!dst || !src) VSHIP_THROW(E_BadPointer);
    const uint8_t* s = static_cast<const uint8_t*>(src);
    const UploadPlan P = plan_upload(&s, &bytes, 1);
    upload(static_cast<uint8_t*>(dst), P, st);
}

void HostStager::upload(uint8_t* dst_base, const UploadPlan& plan, hipStream_t st) {
    if (!dst_base) VSHIP_THROW(E_BadPointer);
    // the open chunk: slot k holds device bytes [c0, c0 + fill)
    int k = -1;
    size_t c0 = 0, fill = 0;
    std::vector<CopySeg> segs;
    auto flush = [&] {
        if (k < 0) return;
        if (pool_)
            pool_->copy(segs.data(), (int)segs.size());
        else
            copy_seg_range(segs.data(), (int)segs.size(), 0, SIZE_MAX);
        HIP_OR_THROW(hipMemcpyAsync(dst_base + c0, ring_[k], fill, hipMemcpyHostToDevice, st), E_BadPointer);
        HIP_OR_THROW(hipEventRecord(ev_[k], st), E_BadDeviceCode);
        k = -1;
        segs.clear();
    };
    for (const UploadPlan::Run& r : plan.runs) {
        if (r.locked) {
            flush();
            HIP_OR_THROW(hipMemcpyAsync(dst_base + r.off, r.src, r.bytes, hipMemcpyHostToDevice, st), E_BadPointer);
            direct_bytes_ += r.bytes;
            continue;
        }
        for (size_t done = 0; done < r.bytes;) {
            const size_t dpos = r.off + done;
            if (k >= 0 && dpos >= c0 + chunk_) flush();
            if (k < 0) {
                k = next_;
                next_ = (next_ + 1) % (int)ring_.size();
                // the slot's previous DMA has to have read it before it is refilled
                HIP_OR_THROW(hipEventSynchronize(ev_[k]), E_BadDeviceCode);
                c0 = dpos;
                fill = 0;
            }
            const size_t n = std::min(r.bytes - done, c0 + chunk_ - dpos);
            segs.push_back({ring_[k] + (dpos - c0), r.src + done, n});
            fill = dpos - c0 + n;
            done += n;
            if (fill == chunk_) flush();
        }
        pageable_bytes_ += r.bytes;
    }
    flush();
}

void HostStager::download_rows(void* dst, int64_t dst_stride, const void* src, int64_t src_stride,
                               size_t row_bytes, int rows, hipStream_t st) {
    if (rows < 1 || row_bytes == 0) return;
    if (!dst || !src || dst_stride < (int64_t)row_bytes || src_stride < (int64_t)row_bytes)
        VSHIP_THROW(E_BadPointer);
    uint8_t* d = static_cast<uint8_t*>(dst);
    const uint8_t* s = static_cast<const uint8_t*>(src);
    const HostMem kind = host_memory_kind(d, (size_t)dst_stride * (rows - 1) + row_bytes);
    if (kind == HostMem::Device) VSHIP_THROW(E_BadPointer);
    if (kind == HostMem::PageLocked) {
        HIP_OR_THROW(hipMemcpy2DAsync(d, dst_stride, s, src_stride, row_bytes, rows, hipMemcpyDeviceToHost, st),
                     E_BadPointer);
        HIP_OR_THROW(hipStreamSynchronize(st), E_BadPointer);
        direct_bytes_ += row_bytes * rows;
        return;
    }
    if (row_bytes > chunk_) VSHIP_THROW(E_BadPointer);
    const int per = (int)std::min<size_t>(chunk_ / row_bytes, (size_t)rows);
    const int nchunks = (rows + per - 1) / per;
    const int nslots = (int)ring_.size();
    // chunk c goes through slot (next_ + c) % nslots; up to nslots DMAs ahead
    auto enqueue = [&](int c) {
        const int k = (next_ + c) % nslots;
        const int r0 = c * per, n = std::min(per, rows - r0);
        HIP_OR_THROW(hipEventSynchronize(ev_[k]), E_BadDeviceCode);
        HIP_OR_THROW(hipMemcpy2DAsync(ring_[k], row_bytes, s + (int64_t)r0 * src_stride, src_stride, row_bytes, n,
                                      hipMemcpyDeviceToHost, st),
                     E_BadPointer);
        HIP_OR_THROW(hipEventRecord(ev_[k], st), E_BadDeviceCode);
    };
    for (int c = 0; c < std::min(nchunks, nslots); c++) enqueue(c);
    for (int c = 0; c < nchunks; c++) {
        const int k = (next_ + c) % nslots;
        const int r0 = c * per, n = std::min(per, rows - r0);
        HIP_OR_THROW(hipEventSynchronize(ev_[k]), E_BadDeviceCode);
        if (dst_stride == (int64_t)row_bytes)
            std::memcpy(d + (int64_t)r0 * dst_stride, ring_[k], row_bytes * n);
        else
            for (int r = 0; r < n; r++) std::memcpy(d + (int64_t)(r0 + r) * dst_stride, ring_[k] + row_bytes * r, row_bytes);
        if (c + nslots < nchunks) enqueue(c + nslots);
    }
    next_ = (next_ + nchunks) % nslots;
    pageable_bytes_ += row_bytes * rows;
}

}  // namespace vship
