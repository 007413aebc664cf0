// A fork-join group of worker threads for large memcpy calls (no HIP
// dependency): copy() splits the bytes into equal parts cut at 4 KiB
// boundaries, and the calling thread copies the first part itself.
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
        // below ~1 MiB per extra thread the wake-ups cost more than they save
        if (workers_.empty() || bytes < ((size_t)1 << 20) * workers_.size()) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<uint8_t*>(dst);
            src_ = static_cast<const uint8_t*>(src);
            bytes_ = bytes;
            remaining_ = (int)workers_.size();
            gen_++;
        }
        go_.notify_all();
        part(static_cast<uint8_t*>(dst), static_cast<const uint8_t*>(src), bytes, 0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return remaining_ == 0; });
    }

private:
    void part(uint8_t* d, const uint8_t* s, size_t bytes, int idx) const {
        size_t b, e;
        copy_part_range(bytes, idx, nthreads_, &b, &e);
        if (e > b) std::memcpy(d + b, s + b, e - b);
    }

    // Each worker takes every generation: copy() starts the next one only
    // after all workers have finished the last.
    void worker(int idx) {
        uint64_t seen = 0;
        for (;;) {
            uint8_t* d;
            const uint8_t* s;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                go_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                d = dst_;
                s = src_;
                n = bytes_;
            }
            part(d, s, n, idx);
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
    uint8_t* dst_ = nullptr;
    const uint8_t* src_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace vship
