// Host-to-device uploads through a ring of page-locked chunks.
//
// A copy from pageable memory has to pass through page-locked memory before
// the DMA engine can read it. HostStager owns that memory: a ring of `slots`
// page-locked chunks with one event each. upload() copies a chunk of the
// caller's bytes into the next free slot on the CPU (split over `threads`
// threads), enqueues the slot's DMA on the given stream and records the
// slot's event, so that the CPU copy of chunk k+1 runs while chunk k crosses
// PCIe, and a slot is refilled only after its DMA has finished. A source that
// is already page-locked (hipHostMalloc) is handed to one DMA directly, and
// upload() returns at once.
//
// One stager serves one caller thread at a time.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "parallel_copy.hpp"

namespace vship {

// True when [p, p + bytes) starts and ends in page-locked host memory.
// Clears the runtime's last-error state that a query on pageable memory sets.
bool host_page_locked(const void* p, size_t bytes);

class HostStager {
public:
    HostStager(int device, size_t chunk_bytes, int slots, int threads);
    ~HostStager();
    HostStager(const HostStager&) = delete;
    HostStager& operator=(const HostStager&) = delete;

    // Copies `bytes` from host `src` to device `dst`, in stream order on `st`.
    // Returns once the caller's memory may be reused: at once for a
    // page-locked source, after the last chunk has been copied into the ring
    // otherwise. The DMA itself completes in stream order.
    void upload(void* dst, const void* src, size_t bytes, hipStream_t st);
    // Copies `rows` rows of `row_bytes` bytes from device `src` (pitch
    // `src_stride`) to host `dst` (pitch `dst_stride`) after the work already
    // on `st`, and returns when they have arrived. A page-locked destination
    // takes one 2-D DMA; a pageable one receives the rows chunk by chunk from
    // the ring, each chunk's DMA running while the previous one is copied out.
    void download_rows(void* dst, int64_t dst_stride, const void* src, int64_t src_stride, size_t row_bytes,
                       int rows, hipStream_t st);

    size_t chunk_bytes() const { return chunk_; }
    int slots() const { return (int)ring_.size(); }
    uint64_t pageable_bytes() const { return pageable_bytes_; }
    uint64_t direct_bytes() const { return direct_bytes_; }

private:
    void release();

    size_t chunk_ = 0;
    std::vector<uint8_t*> ring_;  // page-locked chunks
    std::vector<hipEvent_t> ev_;  // slot i's last DMA
    int next_ = 0;
    ParallelCopy* pool_ = nullptr;
    uint64_t pageable_bytes_ = 0, direct_bytes_ = 0;
};

}  // namespace vship
