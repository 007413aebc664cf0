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

enum class HostMem { Pageable, PageLocked, Device };

// What [p, p + bytes) is: page-locked when it starts and ends in page-locked
// host memory, Device when it starts in device memory (an error for an
// upload), pageable otherwise. Clears the runtime's last-error state that a
// query on pageable memory sets.
HostMem host_memory_kind(const void* p, size_t bytes);

// The device layout of a list of host pieces, and the runs one upload moves.
struct UploadPlan {
    struct Run {
        const uint8_t* src;
        size_t bytes;
        size_t off;   // device offset
        bool locked;  // page-locked: one DMA straight from src
    };
    std::vector<size_t> off;  // device offset of each piece
    std::vector<Run> runs;    // in device order
    size_t total = 0;         // device bytes
};

// Lays out pieces (src[i], bytes[i] > 0) in one device buffer, in order, each
// at a 256-byte-aligned offset. A page-locked piece that follows a page-locked
// piece of the same allocation, at a host distance that is a multiple of 256
// and with a gap of at most 64 KiB, keeps that host distance on the device,
// so that one DMA carries both (and the gap). Throws E_BadPointer for a piece
// in device memory.
UploadPlan plan_upload(const uint8_t* const* src, const size_t* bytes, int n);

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
    // Uploads the runs of `plan` to dst_base + their offsets, in stream order
    // on `st`: page-locked runs by one DMA each, pageable runs packed in
    // device order into the ring's chunks (a chunk carries as many runs as
    // fit, copied by the CPU threads, then one DMA). Returns once the callers'
    // memory may be reused.
    void upload(uint8_t* dst_base, const UploadPlan& plan, hipStream_t st);
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
