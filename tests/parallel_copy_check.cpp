// CPU check of vship_amd/csrc/parallel_copy.hpp: the parts of every split
// tile [0, bytes) exactly, and ParallelCopy reproduces memcpy for sizes
// around the part boundaries and the single-thread cut-off, for 1..9 threads,
// and serial copies for random segment lists.
// Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../vship_amd/csrc/parallel_copy.hpp"

int main() {
    std::mt19937_64 rng(7);
    // part ranges: contiguous, disjoint, covering
    for (int parts = 1; parts <= 17; parts++)
        for (size_t bytes : {(size_t)1, (size_t)4095, (size_t)4096, (size_t)4097, (size_t)16385, (size_t)65536 * 3 + 1,
                             (size_t)(1 << 20) * 5 + 123}) {
            size_t expect = 0;
            for (int i = 0; i < parts; i++) {
                size_t b, e;
                vship::copy_part_range(bytes, i, parts, &b, &e);
                if (b != expect && e > b) {
                    std::printf("gap: bytes %zu parts %d part %d begins %zu, expected %zu\n", bytes, parts, i, b, expect);
                    return 1;
                }
                if (e > b) expect = e;
            }
            if (expect != bytes) {
                std::printf("cover: bytes %zu parts %d ends at %zu\n", bytes, parts, expect);
                return 1;
            }
        }
    std::vector<uint8_t> src((size_t)40 << 20), dst(src.size());
    for (auto& v : src) v = (uint8_t)rng();
    for (int threads = 1; threads <= 9; threads++) {
        vship::ParallelCopy pc(threads);
        for (size_t bytes : {(size_t)0, (size_t)1, (size_t)4097, ((size_t)1 << 20) * (threads - 1),
                             ((size_t)1 << 20) * (threads - 1) + 1, ((size_t)1 << 20) * threads + 4095,
                             (size_t)(16 << 20) + 3, src.size() - 5}) {
            const size_t off = bytes % 7;  // unaligned source
            if (bytes + off > src.size()) continue;
            std::fill(dst.begin(), dst.end(), 0xA5);
            pc.copy(dst.data(), src.data() + off, bytes);
            for (size_t i = 0; i < bytes; i++)
                if (dst[i] != src[off + i]) {
                    std::printf("mismatch: threads %d bytes %zu at %zu\n", threads, bytes, i);
                    return 1;
                }
            for (size_t i = bytes; i < std::min(dst.size(), bytes + 8192); i++)
                if (dst[i] != 0xA5) {
                    std::printf("overrun: threads %d bytes %zu at %zu\n", threads, bytes, i);
                    return 1;
                }
        }
    }
    // gathered copies: random segment lists (empty segments included) into
    // disjoint destinations with gaps that must stay untouched
    for (int threads = 1; threads <= 9; threads += 2) {
        vship::ParallelCopy pc(threads);
        for (int trial = 0; trial < 20; trial++) {
            std::vector<vship::CopySeg> segs;
            std::vector<std::pair<size_t, size_t>> ranges;  // (dst offset, bytes)
            size_t dpos = 0;
            const int n = 1 + (int)(rng() % 12);
            for (int i = 0; i < n; i++) {
                const size_t bytes = trial % 5 == 0 ? rng() % 5000 : rng() % ((size_t)3 << 20);
                dpos += rng() % 300;  // gap
                const size_t soff = rng() % (src.size() - bytes);
                if (dpos + bytes > dst.size()) break;
                segs.push_back({dst.data() + dpos, src.data() + soff, bytes});
                ranges.push_back({dpos, soff});
                dpos += bytes;
            }
            std::fill(dst.begin(), dst.end(), 0x5A);
            pc.copy(segs.data(), (int)segs.size());
            std::vector<uint8_t> want(dst.size(), 0x5A);
            for (size_t i = 0; i < segs.size(); i++)
                std::memcpy(want.data() + ranges[i].first, src.data() + ranges[i].second, segs[i].bytes);
            if (std::memcmp(want.data(), dst.data(), dst.size()) != 0) {
                std::printf("segments mismatch: threads %d trial %d\n", threads, trial);
                return 1;
            }
        }
    }
    std::printf("ok\n");
    return 0;
}
