// Host-to-device uploads through a ring of page-locked chunks (HostStager).
#include "host_stage.hpp"

#include <algorithm>
#include <cstring>

#include "engine.hpp"

namespace vship {

bool host_page_locked(const void* p, size_t bytes) {
    if (!p || bytes == 0) return false;
    const uint8_t* b = static_cast<const uint8_t*>(p);
    for (const uint8_t* q : {b, b + bytes - 1}) {
        hipPointerAttribute_t attr;
        if (hipPointerGetAttributes(&attr, q) != hipSuccess) {
            (void)hipGetLastError();  // pageable memory: the query fails and leaves its error behind
            return false;
        }
        if (attr.type != hipMemoryTypeHost) return false;
    }
    return true;
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
    if (host_page_locked(src, bytes)) {
        HIP_OR_THROW(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), E_BadPointer);
        direct_bytes_ += bytes;
        return;
    }
    uint8_t* d = static_cast<uint8_t*>(dst);
    const uint8_t* s = static_cast<const uint8_t*>(src);
    for (size_t off = 0; off < bytes; off += chunk_) {
        const size_t n = std::min(chunk_, bytes - off);
        const int k = next_;
        next_ = (next_ + 1) % (int)ring_.size();
        // the slot's previous DMA has to have read it before it is refilled
        HIP_OR_THROW(hipEventSynchronize(ev_[k]), E_BadDeviceCode);
        if (pool_)
            pool_->copy(ring_[k], s + off, n);
        else
            std::memcpy(ring_[k], s + off, n);
        HIP_OR_THROW(hipMemcpyAsync(d + off, ring_[k], n, hipMemcpyHostToDevice, st), E_BadPointer);
        HIP_OR_THROW(hipEventRecord(ev_[k], st), E_BadDeviceCode);
    }
    pageable_bytes_ += bytes;
}

void HostStager::download_rows(void* dst, int64_t dst_stride, const void* src, int64_t src_stride,
                               size_t row_bytes, int rows, hipStream_t st) {
    if (rows < 1 || row_bytes == 0) return;
    if (!dst || !src || dst_stride < (int64_t)row_bytes || src_stride < (int64_t)row_bytes)
        VSHIP_THROW(E_BadPointer);
    uint8_t* d = static_cast<uint8_t*>(dst);
    const uint8_t* s = static_cast<const uint8_t*>(src);
    if (host_page_locked(d, (size_t)dst_stride * (rows - 1) + row_bytes)) {
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
