// A fork-join group of worker threads for large memcpy calls (no HIP
// dependency): copy() splits the bytes of one range, or of a list of
// segments, into equal parts cut at 4 KiB boundaries, and the calling thread
// copies the first part itself.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace vship {

// [begin, end) of part `idx` of `parts` of a `bytes`-long copy; the parts
// cover [0, bytes) without overlap.
inline void copy_part_range(size_t bytes, int idx, int parts, size_t* begin, size_t* end) {
    const size_t per = ((bytes + parts - 1) / parts + 4095) / 4096 * 4096;
    *begin = std::min(bytes, per * idx);
    *end = std::min(bytes, *begin + per);
}

// One contiguous piece of a gathered copy.
struct CopySeg {
    uint8_t* dst;
    const uint8_t* src;
    size_t bytes;
};

// Copies the bytes [b, e) of the concatenation of `segs`.
inline void copy_seg_range(const CopySeg* segs, int n, size_t b, size_t e) {
    size_t s0 = 0;
    for (int i = 0; i < n && s0 < e; i++) {
        const size_t s1 = s0 + segs[i].bytes;
        const size_t lo = std::max(b, s0), hi = std::min(e, s1);
        if (hi > lo) std::memcpy(segs[i].dst + (lo - s0), segs[i].src + (lo - s0), hi - lo);
        s0 = s1;
    }
}

class ParallelCopy {
public:
    explicit ParallelCopy(int threads) : nthreads_(std::max(1, std::min(threads, 64))) {
        for (int i = 1; i < nthreads_; i++) workers_.emplace_back(&ParallelCopy::worker, this, i);
    }
    ~ParallelCopy() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        go_.notify_all();
        for (auto& t : workers_) t.join();
    }
    ParallelCopy(const ParallelCopy&) = delete;
    ParallelCopy& operator=(const ParallelCopy&) = delete;

    int threads() const { return nthreads_; }

    void copy(void* dst, const void* src, size_t bytes) {
        const CopySeg s{static_cast<uint8_t*>(dst), static_cast<const uint8_t*>(src), bytes};
        copy(&s, 1);
    }

    // Copies every segment; the threads split the segments' total bytes.
    void copy(const CopySeg* segs, int n) {
        size_t total = 0;
        for (int i = 0; i < n; i++) total += segs[i].bytes;
        // below ~1 MiB per extra thread the wake-ups cost more than they save
        if (workers_.empty() || total < ((size_t)1 << 20) * workers_.size()) {
            copy_seg_range(segs, n, 0, total);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            segs_ = segs;
            nseg_ = n;
            total_ = total;
            remaining_ = (int)workers_.size();
            gen_++;
        }
        go_.notify_all();
        part(segs, n, total, 0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return remaining_ == 0; });
    }

private:
    void part(const CopySeg* segs, int n, size_t total, int idx) const {
        size_t b, e;
        copy_part_range(total, idx, nthreads_, &b, &e);
        if (e > b) copy_seg_range(segs, n, b, e);
    }

    // Each worker takes every generation: copy() starts the next one only
    // after all workers have finished the last.
    void worker(int idx) {
        uint64_t seen = 0;
        for (;;) {
            const CopySeg* segs;
            int n;
            size_t total;
            {
                std::unique_lock<std::mutex> lk(mu_);
                go_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                segs = segs_;
                n = nseg_;
                total = total_;
            }
            part(segs, n, total, idx);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--remaining_ == 0) done_.notify_one();
            }
        }
    }

    const int nthreads_;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    uint64_t gen_ = 0;
    int remaining_ = 0;
    bool stop_ = false;
    const CopySeg* segs_ = nullptr;
    int nseg_ = 0;
    size_t total_ = 0;
};

}  // namespace vship
