// ffvship: the command line of the reference's FFVship (src/FFVship.cpp,
// src/ffvship_utility/ffmpegmain.hpp:790-840, doc/FFVship.md) over this
// library's batched SSIMULACRA2 and Butteraugli engines, for YUV4MPEG2 (.y4m)
// inputs.
//
// Same options and meanings for what a y4m input allows (--source/-s,
// --encoded/-e, --metric/-m SSIMULACRA2|Butteraugli, --intensity-target,
// --qnorm, --start, --end, --every, --encoded-offset, --source-indices,
// --encoded-indices, --json, --live-score-output, --threads/-t,
// --gpu-threads/-g, --gpu-id, --list-gpu, --version), the same frame
// selection (src/FFVship.cpp:262-330), JSON file ([[score], ...], or
// [[qnorm, 3-norm, inf-norm], ...] for Butteraugli, :439-468), live lines
// (first the frame count, then "index score" or "index q 3 inf", :113-140)
// and console statistics (:142-200, one block per norm for Butteraugli,
// :473-499). Decoding goes away: y4m
// frames are raw planar YUV, read straight into pinned host memory, copied
// to the device and scored in batches of 8 x --gpu-threads pairs, the next
// batch read from disk while the GPU scores the current one.
//
// Colour handling follows FFVship's defaults for what a y4m header leaves
// open (src/ffvship_utility/ffmpegToZimgFormat.hpp:4-15,223-246): matrix
// BT.709 above 650 rows, else BT.470BG (the BT.601 coefficients); limited
// range; chroma location from the C tag (420jpeg and 420: center, 420mpeg2:
// left, 420paldv: top-left), else left. XCOLORRANGE=FULL, mono and alpha
// streams are refused. Source and encoded must have the same size and
// layout (FFVship would crop or resize; this build does not).
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "VshipButteraugliBatch.h"
#include "VshipSSIMU2Batch.h"

namespace ffv {

// ---------------------------------------------------------------------------
// YUV4MPEG2
// ---------------------------------------------------------------------------
struct Y4mFormat {
    int width = 0, height = 0;
    int sw = 1, sh = 1;      // log2 chroma subsampling
    int bits = 8;            // sample depth; > 8 means 16-bit little-endian containers
    int chroma_location = Vship_Chroma_Center;
    std::string error;       // non-empty: unsupported or malformed
};

// Layout of a y4m C tag (without the leading C). The YUV4MPEG2 default is
// 420jpeg.
bool parse_colorspace(const std::string& c, Y4mFormat& f) {
    f.bits = 8;
    f.chroma_location = Vship_Chroma_Left;
    std::string base = c;
    const size_t p = c.find('p');
    if (p != std::string::npos && p + 1 < c.size() && isdigit((unsigned char)c[p + 1])) {
        base = c.substr(0, p);
        f.bits = atoi(c.c_str() + p + 1);
        if (f.bits < 9 || f.bits > 16) return false;
    }
    if (base == "420jpeg" || base == "420") {
        f.sw = 1, f.sh = 1;
        if (f.bits == 8) f.chroma_location = Vship_Chroma_Center;
    } else if (base == "420mpeg2") {
        f.sw = 1, f.sh = 1;
    } else if (base == "420paldv") {
        f.sw = 1, f.sh = 1;
        f.chroma_location = Vship_Chroma_TopLeft;
    } else if (base == "422") {
        f.sw = 1, f.sh = 0;
    } else if (base == "444") {
        f.sw = 0, f.sh = 0;
    } else if (base == "411") {
        f.sw = 2, f.sh = 0;
    } else {
        return false;
    }
    return true;
}

// The stream header line (without the trailing newline).
Y4mFormat parse_header(const std::string& line) {
    Y4mFormat f;
    std::istringstream in(line);
    std::string tok;
    in >> tok;
    if (tok != "YUV4MPEG2") {
        f.error = "not a YUV4MPEG2 stream";
        return f;
    }
    std::string cs = "420jpeg";
    while (in >> tok) {
        switch (tok[0]) {
        case 'W': f.width = atoi(tok.c_str() + 1); break;
        case 'H': f.height = atoi(tok.c_str() + 1); break;
        case 'C': cs = tok.substr(1); break;
        case 'X':
            if (tok == "XCOLORRANGE=FULL") f.error = "full-range YUV is not supported (limited range only)";
            break;
        default: break;  // F, I, A and other X tags do not change the samples
        }
    }
    if (f.error.empty() && !parse_colorspace(cs, f)) f.error = "unsupported y4m colorspace C" + cs;
    if (f.error.empty() && (f.width < 1 || f.height < 1)) f.error = "missing or invalid W/H in the y4m header";
    return f;
}

// Vship_PixelFormat of a layout and depth.
int pixel_format(const Y4mFormat& f) {
    int layout;
    if (f.sw == 1 && f.sh == 1) layout = 0;
    else if (f.sw == 1 && f.sh == 0) layout = 1;
    else if (f.sw == 0 && f.sh == 0) layout = 2;
    else if (f.sw == 0 && f.sh == 1) layout = 3;
    else if (f.sw == 2 && f.sh == 0) layout = 4;
    else layout = 5;
    return Vship_PF_YUV420_U8 + 2 * layout + (f.bits > 8 ? 1 : 0);
}

int64_t plane_width(const Y4mFormat& f, int pl) { return pl ? (f.width + (1 << f.sw) - 1) >> f.sw : f.width; }
int64_t plane_height(const Y4mFormat& f, int pl) { return pl ? (f.height + (1 << f.sh) - 1) >> f.sh : f.height; }
int64_t sample_bytes(const Y4mFormat& f) { return f.bits > 8 ? 2 : 1; }
int64_t plane_bytes(const Y4mFormat& f, int pl) { return plane_width(f, pl) * plane_height(f, pl) * sample_bytes(f); }
int64_t frame_bytes(const Y4mFormat& f) { return plane_bytes(f, 0) + plane_bytes(f, 1) + plane_bytes(f, 2); }

// A y4m file: its format and the byte offset of every frame's samples.
struct Y4mFile {
    std::string path;
    Y4mFormat fmt;
    std::vector<int64_t> frames;
    int fd = -1;
    std::string error;

    bool open_file(const std::string& p) {
        path = p;
        fd = ::open(p.c_str(), O_RDONLY);
        if (fd < 0) {
            error = "cannot open " + p;
            return false;
        }
        std::string line;
        int64_t pos = 0;
        if (!read_line(pos, line, 4096)) {
            error = p + ": no y4m header";
            return false;
        }
        fmt = parse_header(line);
        if (!fmt.error.empty()) {
            error = p + ": " + fmt.error;
            return false;
        }
        const int64_t fb = frame_bytes(fmt);
        const int64_t size = lseek(fd, 0, SEEK_END);
        // walk the FRAME headers (each may carry parameters)
        while (pos < size) {
            if (!read_line(pos, line, 4096) || line.compare(0, 5, "FRAME") != 0) {
                error = p + ": malformed FRAME header at byte " + std::to_string(pos);
                return false;
            }
            if (pos + fb > size) break;  // a truncated last frame is dropped
            frames.push_back(pos);
            pos += fb;
        }
        return true;
    }

    // The line starting at pos (advanced past its newline).
    bool read_line(int64_t& pos, std::string& line, int max_len) {
        line.clear();
        char buf[256];
        while ((int)line.size() < max_len) {
            const ssize_t n = pread(fd, buf, sizeof(buf), pos + (int64_t)line.size());
            if (n <= 0) return false;
            const char* nl = (const char*)memchr(buf, '\n', (size_t)n);
            if (nl) {
                line.append(buf, nl - buf);
                pos += (int64_t)line.size() + 1;
                return true;
            }
            line.append(buf, (size_t)n);
        }
        return false;
    }

    bool read_frame(int index, uint8_t* dst) const {
        const int64_t fb = frame_bytes(fmt);
        int64_t done = 0;
        while (done < fb) {
            const ssize_t n = pread(fd, dst + done, (size_t)(fb - done), frames[index] + done);
            if (n <= 0) return false;
            done += n;
        }
        return true;
    }

    ~Y4mFile() {
        if (fd >= 0) ::close(fd);
    }
};

// ---------------------------------------------------------------------------
// frame selection (src/FFVship.cpp:262-330)
// ---------------------------------------------------------------------------
struct Selection {
    std::vector<int> source, encoded;
    std::string error;
};

Selection select_frames(int total_source, int total_encoded, int start, int end, int every, int offset,
                        const std::vector<int>& source_indices, const std::vector<int>& encoded_indices) {
    Selection s;
    if (every < 1) {
        s.error = "--every must be at least 1";
        return s;
    }
    if (start < 0) start = 0;
    if (end < 0) end = total_source;
    end = std::min(end, total_source);
    start = std::min(start, total_source);
    if (end < start) end = start;
    start = std::max(-offset, start);
    end = start + std::min(total_encoded - (start + offset), end - start);
    if (end < start) {
        s.error = "encoded_offset " + std::to_string(offset) + " does not allow comparing both videos";
        return s;
    }
    if (source_indices.empty()) {
        for (int i = start; i < end; i += every) s.source.push_back(i);
    } else {
        for (int el : source_indices) {
            if ((int64_t)el * every + start >= total_source || el < 0) {
                s.error = "Source Invalid frame index found " + std::to_string(el) + " which will be placed at " +
                          std::to_string(el * every + start) + " after computed end " + std::to_string(total_source);
                return s;
            }
            s.source.push_back(start + el * every);
        }
    }
    if (encoded_indices.empty() && source_indices.empty()) {
        for (int i = start; i < end; i += every) s.encoded.push_back(i + offset);
    } else if (encoded_indices.empty()) {
        for (int el : source_indices) {
            if ((int64_t)el * every + start + offset >= total_encoded) {
                s.error = "Encoded Invalid frame index found " + std::to_string(el) + " which will be placed at " +
                          std::to_string(el * every + start + offset) + " after computed end " +
                          std::to_string(total_encoded);
                return s;
            }
            s.encoded.push_back(start + el * every + offset);
        }
    } else {
        for (int el : encoded_indices) {
            if ((int64_t)el * every + start >= end || el < 0) {
                s.error = "Encoded Invalid frame index found " + std::to_string(el) + " which will be placed at " +
                          std::to_string(el * every + start + offset) + " after computed end " +
                          std::to_string(end + offset);
                return s;
            }
            s.encoded.push_back(start + el * every + offset);
        }
    }
    if (s.encoded.size() != s.source.size()) s.error = "Source indices and Encoded indices are of different sizes, aborting";
    return s;
}

bool parse_index_list(const std::string& text, std::vector<int>& out) {
    out.clear();
    std::stringstream ss(text);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        char* endp = nullptr;
        const long v = strtol(item.c_str(), &endp, 10);
        if (*endp != '\0') return false;
        out.push_back((int)v);
    }
    return true;
}

// ---------------------------------------------------------------------------
// output formats
// ---------------------------------------------------------------------------
// JSON file of per-frame scores: [[s0], [s1], ...] (src/FFVship.cpp:439-468).
std::string json_scores(const std::vector<float>& scores) {
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < scores.size(); i++) o << "[" << scores[i] << (i + 1 == scores.size() ? "]" : "], ");
    o << "]";
    return o.str();
}

// Butteraugli: [[q, 3, inf], ...], values with the stream's default
// formatting, separated by ", " (src/FFVship.cpp:448-452).
std::string json_norms(const std::vector<float>& s3) {
    std::ostringstream o;
    const size_t n = s3.size() / 3;
    o << "[";
    for (size_t i = 0; i < n; i++)
        o << "[" << s3[3 * i] << ", " << s3[3 * i + 1] << ", " << s3[3 * i + 2] << (i + 1 == n ? "]" : "], ");
    o << "]";
    return o.str();
}

// Metric names as FFVship parses them (ffmpegmain.hpp:770-780), lowercased:
// 0 SSIMULACRA2, 1 Butteraugli, 2 CVVDP (not built here), -1 unknown.
int metric_of(const std::string& name) {
    std::string l = name;
    for (char& c : l) c = (char)std::tolower((unsigned char)c);
    if (l == "ssimulacra2" || l == "ssimu2") return 0;
    if (l == "butteraugli" || l == "butter") return 1;
    if (l == "cvvdp") return 2;
    return -1;
}

// Statistics block (src/FFVship.cpp:142-200): the label centred in a dashed
// rule as wide as "name : value", then the seven statistics right-aligned.
std::string stats_block(const std::vector<float>& data, const std::string& label) {
    if (data.empty()) return "";
    std::vector<float> v = data;
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    double mean = 0.0;
    for (float x : v) mean += x;
    mean /= (double)n;
    double ss = 0.0;
    for (float x : v) ss += (x - mean) * (x - mean);
    const std::pair<const char*, double> rows[] = {{"Average", mean},
                                                   {"Standard Deviation", std::sqrt(ss / (double)n)},
                                                   {"Median", v[n / 2]},
                                                   {"5th percentile", v[n / 20]},
                                                   {"95th percentile", v[19 * n / 20]},
                                                   {"Minimum", v.front()},
                                                   {"Maximum", v.back()}};
    size_t name_w = 0;
    for (const auto& r : rows) name_w = std::max(name_w, strlen(r.first));
    const int value_w = 12, total = (int)name_w + 3 + value_w;
    const int pad = std::max(0, (total - (int)label.size()) / 2);
    std::ostringstream o;
    o << std::string(pad, '-') << label << std::string(std::max(0, total - pad - (int)label.size()), '-') << "\n";
    for (const auto& r : rows)
        o << std::setw((int)name_w) << std::right << r.first << " : " << std::setw(value_w) << std::right << std::fixed
          << std::setprecision(6) << r.second << "\n";
    o << "\n";
    return o.str();
}

// ---------------------------------------------------------------------------
// command line
// ---------------------------------------------------------------------------
struct Options {
    std::string source, encoded, metric = "SSIMULACRA2", json, source_indices, encoded_indices;
    int start = 0, end = -1, every = 1, offset = 0, threads = 1, gpu_threads = 8, gpu_id = 0;
    int qnorm = 2, intensity = 203;  // Butteraugli (ffmpegmain.hpp:725-727)
    bool live = false, list_gpu = false, version = false, help = false;
};

const char* kUsage =
    "usage: ffvship [-h] --source SOURCE.y4m --encoded ENCODED.y4m [-m SSIMULACRA2|Butteraugli]\n"
    "               [--intensity-target NITS] [--qnorm Q]\n"
    "               [--start N] [--end N] [--every N] [--encoded-offset N]\n"
    "               [--source-indices i,j,..] [--encoded-indices i,j,..]\n"
    "               [-t THREADS] [-g GPU_THREADS] [--gpu-id ID] [--json OUTPUT]\n"
    "               [--live-score-output] [--list-gpu] [--version]\n";

// Returns "" or an error message.
std::string parse_args(int argc, char** argv, Options& o) {
    auto need = [&](int& i, const std::string& name) -> const char* {
        if (i + 1 >= argc) return nullptr;
        (void)name;
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        const char* v = nullptr;
        auto as_int = [&](int& dst) -> bool {
            if (!(v = need(i, a))) return false;
            char* endp = nullptr;
            const long x = strtol(v, &endp, 10);
            if (*endp != '\0') return false;
            dst = (int)x;
            return true;
        };
        if (a == "-h" || a == "--help") o.help = true;
        else if (a == "--live-score-output") o.live = true;
        else if (a == "--list-gpu") o.list_gpu = true;
        else if (a == "--version") o.version = true;
        else if (a == "--source" || a == "-s" || a == "--reference") {
            if (!(v = need(i, a))) return "missing value for " + a;
            o.source = v;
        } else if (a == "--encoded" || a == "-e") {
            if (!(v = need(i, a))) return "missing value for " + a;
            o.encoded = v;
        } else if (a == "--metric" || a == "-m") {
            if (!(v = need(i, a))) return "missing value for " + a;
            o.metric = v;
        } else if (a == "--json") {
            if (!(v = need(i, a))) return "missing value for " + a;
            o.json = v;
        } else if (a == "--source-indices") {
            if (!(v = need(i, a))) return "missing value for " + a;
            o.source_indices = v;
        } else if (a == "--encoded-indices") {
            if (!(v = need(i, a))) return "missing value for " + a;
            o.encoded_indices = v;
        } else if (a == "--start") {
            if (!as_int(o.start)) return "invalid value for " + a;
        } else if (a == "--end") {
            if (!as_int(o.end)) return "invalid value for " + a;
        } else if (a == "--every") {
            if (!as_int(o.every)) return "invalid value for " + a;
        } else if (a == "--encoded-offset") {
            if (!as_int(o.offset)) return "invalid value for " + a;
        } else if (a == "--threads" || a == "-t") {
            if (!as_int(o.threads)) return "invalid value for " + a;
        } else if (a == "--gpu-threads" || a == "-g") {
            if (!as_int(o.gpu_threads)) return "invalid value for " + a;
        } else if (a == "--gpu-id") {
            if (!as_int(o.gpu_id)) return "invalid value for " + a;
        } else if (a == "--intensity-target") {
            if (!as_int(o.intensity)) return "invalid value for " + a;
        } else if (a == "--qnorm") {
            if (!as_int(o.qnorm)) return "invalid value for " + a;
        } else {
            return "unknown option " + a;
        }
    }
    return "";
}

std::string vship_error(Vship_Exception e) {
    char buf[1024];
    Vship_GetErrorMessage(e, buf, (int)sizeof(buf));
    return buf;
}

// JSON file, then (unless live lines were asked for) the console summary.
// scores: one per frame, or three (q, 3, inf) for Butteraugli.
int report(const Options& o, const std::vector<float>& scores, double secs) {
    const bool butter = metric_of(o.metric) == 1;
    if (!o.json.empty()) {
        std::ofstream js(o.json);
        if (!js) {
            std::cerr << "Failed to open output file" << std::endl;
            return 1;
        }
        js << (butter ? json_norms(scores) : json_scores(scores));
    }
    if (o.live) return 0;
    const size_t n = butter ? scores.size() / 3 : scores.size();
    const double fps = secs > 0 ? (double)n / secs : 0.0;
    std::cout << (butter ? "Butteraugli" : "SSIMU2") << " Result between " << o.source << " and " << o.encoded
              << std::endl;
    std::cout << "Computed " << n << " frames at " << fps << " fps\n" << std::endl;
    if (!butter) {
        std::cout << stats_block(scores, "SSIMULACRA2");
        return 0;
    }
    std::vector<float> nq(n), n3(n), ninf(n);
    for (size_t i = 0; i < n; i++) nq[i] = scores[3 * i], n3[i] = scores[3 * i + 1], ninf[i] = scores[3 * i + 2];
    std::cout << stats_block(nq, std::to_string(o.qnorm) + "-Norm");
    std::cout << stats_block(n3, "3-Norm");
    std::cout << stats_block(ninf, "INF-Norm");
    return 0;
}

// One batch slot: pinned host frames, device frames, descriptors, scores.
struct Slot {
    uint8_t* host = nullptr;
    uint8_t* dev = nullptr;
    std::vector<Vship_SSIMU2Pair> pairs;
    std::vector<Vship_ButteraugliPair> bpairs;
    std::vector<int> task;  // task index of each pair
    int n = 0;
};

int run(const Options& o) {
    const int metric = metric_of(o.metric);
    if (metric != 0 && metric != 1) {
        std::cerr << "metric " << o.metric << " is not supported by this build (SSIMULACRA2, Butteraugli)"
                  << std::endl;
        return 1;
    }
    const bool butter = metric == 1;
    const int per = butter ? 3 : 1;  // values per frame
    Y4mFile src, enc;
    if (!src.open_file(o.source)) {
        std::cerr << src.error << std::endl;
        return 1;
    }
    if (!enc.open_file(o.encoded)) {
        std::cerr << enc.error << std::endl;
        return 1;
    }
    const Y4mFormat& f = src.fmt;
    if (enc.fmt.width != f.width || enc.fmt.height != f.height || pixel_format(enc.fmt) != pixel_format(f) ||
        enc.fmt.bits != f.bits) {
        std::cerr << "source and encoded differ in size, layout or depth (no resize in this build)" << std::endl;
        return 1;
    }
    std::vector<int> sidx, eidx;
    if (!parse_index_list(o.source_indices, sidx) || !parse_index_list(o.encoded_indices, eidx)) {
        std::cerr << "indices are integers separated by commas" << std::endl;
        return 1;
    }
    if (!eidx.empty() && sidx.empty()) {
        std::cerr << "--encoded-indices needs --source-indices" << std::endl;
        return 1;
    }
    const Selection sel = select_frames((int)src.frames.size(), (int)enc.frames.size(), o.start, o.end, o.every,
                                        o.offset, sidx, eidx);
    if (!sel.error.empty()) {
        std::cerr << sel.error << std::endl;
        return 1;
    }
    const int nframes = (int)sel.source.size();
    if (o.live) std::cout << nframes << std::endl;
    std::vector<float> scores((size_t)nframes * per, 0.0f);
    if (nframes == 0) return report(o, scores, 0.0);

    Vship_Exception e = Vship_SetDevice(o.gpu_id);
    if (e != Vship_NoError) {
        std::cerr << vship_error(e) << std::endl;
        return 1;
    }
    const int batch = std::max(1, std::min(8 * std::max(1, o.gpu_threads), std::max(1, nframes)));
    Vship_SSIMU2BatchConfig cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.max_pairs = batch;
    cfg.max_width = f.width;
    cfg.max_height = f.height;
    Vship_SSIMU2BatchHandler h{};
    Vship_ButteraugliBatchHandler bh{};
    if (butter) {
        Vship_ButteraugliBatchConfig bc;
        std::memset(&bc, 0, sizeof(bc));
        bc.width = f.width;
        bc.height = f.height;
        bc.max_pairs = batch;
        bc.qnorm = o.qnorm;
        bc.intensity_multiplier = (float)o.intensity;
        e = Vship_ButteraugliBatchInit(&bh, &bc);
    } else {
        e = Vship_SSIMU2BatchInit(&h, &cfg);
    }
    if (e != Vship_NoError) {
        std::cerr << vship_error(e) << std::endl;
        return 1;
    }
    const int64_t fb = frame_bytes(f);
    const int64_t slot_bytes = 2 * fb * batch;
    Slot slots[2];
    int rc = 0;
    for (Slot& s : slots) {
        if ((e = Vship_PinnedMalloc((void**)&s.host, (uint64_t)slot_bytes)) != Vship_NoError ||
            (e = Vship_DeviceMalloc((void**)&s.dev, (uint64_t)slot_bytes)) != Vship_NoError) {
            std::cerr << vship_error(e) << std::endl;
            rc = 1;
        }
        s.pairs.resize(batch);
        s.bpairs.resize(batch);
        s.task.resize(batch);
    }
    const int matrix = f.height > 650 ? Vship_Matrix_BT709 : Vship_Matrix_BT601;
    const int fmt = pixel_format(f);
    const int nthreads = std::max(1, o.threads);
    std::vector<double> out((size_t)batch * per);

    // read tasks [first, first + n) into slot s (frames pinned, descriptors set)
    auto fill = [&](Slot& s, int first, int n) -> bool {
        s.n = n;
        std::vector<char> ok(n, 1);
        auto work = [&](int t0) {
            for (int k = t0; k < n; k += nthreads) {
                uint8_t* r = s.host + 2 * fb * k;
                ok[k] = src.read_frame(sel.source[first + k], r) && enc.read_frame(sel.encoded[first + k], r + fb);
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; t++) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
        for (int k = 0; k < n; k++) {
            if (!ok[k]) return false;
            Vship_SSIMU2Pair& p = s.pairs[k];
            std::memset(&p, 0, sizeof(p));
            int64_t off = 2 * fb * k;
            for (int img = 0; img < 2; img++)
                for (int pl = 0; pl < 3; pl++) {
                    const uint8_t* ptr = s.dev + off;
                    const int64_t stride = plane_width(f, pl) * sample_bytes(f);
                    if (img == 0) {
                        p.ref[pl] = ptr;
                        p.ref_stride[pl] = stride;
                    } else {
                        p.dist[pl] = ptr;
                        p.dist_stride[pl] = stride;
                    }
                    off += plane_bytes(f, pl);
                }
            p.width = f.width;
            p.height = f.height;
            p.format = fmt;
            p.bits = f.bits;
            p.matrix = matrix;
            p.chroma_location = f.chroma_location;
            Vship_ButteraugliPair& b = s.bpairs[k];
            std::memset(&b, 0, sizeof(b));
            for (int pl = 0; pl < 3; pl++) {
                b.ref[pl] = p.ref[pl];
                b.dist[pl] = p.dist[pl];
                b.ref_stride[pl] = p.ref_stride[pl];
                b.dist_stride[pl] = p.dist_stride[pl];
            }
            b.bits = f.bits;
            b.matrix = matrix;
            b.chroma_location = f.chroma_location;
            s.task[k] = first + k;
        }
        return true;
    };
    // Butteraugli's batch call is synchronous: it runs on a thread of its
    // own, so that reading the next batch still overlaps the GPU work
    std::thread butter_worker;
    Vship_Exception butter_rc = Vship_NoError;
    auto submit = [&](Slot& s) -> Vship_Exception {
        Vship_Exception x = Vship_MemcpyHtoD(s.dev, s.host, (uint64_t)(2 * fb * s.n));
        if (x != Vship_NoError) return x;
        if (!butter) return Vship_SSIMU2BatchSubmit(h, s.pairs.data(), s.n);
        butter_worker = std::thread([&, sp = &s] {
            butter_rc = Vship_ButteraugliBatchCompute(bh, fmt, sp->bpairs.data(), sp->n, out.data());
        });
        return Vship_NoError;
    };
    auto wait = [&]() -> Vship_Exception {
        if (!butter) return Vship_SSIMU2BatchWait(h, out.data(), nullptr);
        if (butter_worker.joinable()) butter_worker.join();
        return butter_rc;
    };

    const auto t0 = std::chrono::steady_clock::now();
    int next = 0, cur = 0;
    bool pending = false;
    if (rc == 0 && nframes > 0) {
        const int n0 = std::min(batch, nframes);
        if (!fill(slots[0], 0, n0)) {
            std::cerr << "failed to read frames" << std::endl;
            rc = 1;
        } else if ((e = submit(slots[0])) != Vship_NoError) {
            std::cerr << vship_error(e) << std::endl;
            rc = 1;
        } else {
            pending = true;
            next = n0;
        }
    }
    while (rc == 0 && pending) {
        // read the next batch while the GPU scores the current one
        Slot& nx = slots[cur ^ 1];
        const int nn = std::min(batch, nframes - next);
        bool have_next = nn > 0;
        if (have_next && !fill(nx, next, nn)) {
            std::cerr << "failed to read frames" << std::endl;
            rc = 1;
        }
        Slot& s = slots[cur];
        if ((e = wait()) != Vship_NoError) {
            std::cerr << vship_error(e) << std::endl;
            rc = 1;
            break;
        }
        pending = false;
        for (int k = 0; k < s.n; k++) {
            for (int m = 0; m < per; m++) scores[(size_t)s.task[k] * per + m] = (float)out[(size_t)k * per + m];
            if (!o.live) continue;
            if (butter)
                std::cout << s.task[k] << " " << (float)out[3 * k] << " " << (float)out[3 * k + 1] << " "
                          << (float)out[3 * k + 2] << "\n";
            else
                std::cout << s.task[k] << " " << (float)out[k] << "\n";
        }
        if (o.live) std::cout.flush();
        if (rc == 0 && have_next) {
            if ((e = submit(nx)) != Vship_NoError) {
                std::cerr << vship_error(e) << std::endl;
                rc = 1;
                break;
            }
            pending = true;
            next += nn;
            cur ^= 1;
        }
    }
    if (butter_worker.joinable()) butter_worker.join();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (Slot& s : slots) {
        if (s.host) Vship_PinnedFree(s.host);
        if (s.dev) Vship_DeviceFree(s.dev);
    }
    if (butter)
        Vship_ButteraugliBatchFree(bh);
    else
        Vship_SSIMU2BatchFree(h);
    if (rc != 0) return rc;
    return report(o, scores, secs);
}

int list_gpus() {
    int n = 0;
    Vship_Exception e = Vship_GetDeviceCount(&n);
    if (e != Vship_NoError) {
        std::cerr << vship_error(e) << std::endl;
        return 1;
    }
    for (int i = 0; i < n; i++) {
        Vship_DeviceInfo info;
        if ((e = Vship_GetDeviceInfo(&info, i)) != Vship_NoError) {
            std::cerr << vship_error(e) << std::endl;
            return 1;
        }
        std::cout << "GPU " << i << ": " << info.name << std::endl;
    }
    return 0;
}

}  // namespace ffv

#ifdef FFVSHIP_TEST_HOOKS
// Host logic of the command line for CPU tests (no GPU calls).
extern "C" {
int ffv_parse_header(const char* line, int* out6, char* err, int err_len) {
    const ffv::Y4mFormat f = ffv::parse_header(line);
    snprintf(err, (size_t)err_len, "%s", f.error.c_str());
    if (!f.error.empty()) return 1;
    out6[0] = f.width, out6[1] = f.height, out6[2] = ffv::pixel_format(f), out6[3] = f.bits;
    out6[4] = f.chroma_location, out6[5] = (int)ffv::frame_bytes(f);
    return 0;
}
int ffv_select(int total_src, int total_enc, int start, int end, int every, int offset, const char* sidx,
               const char* eidx, int* src_out, int* enc_out, int cap, char* err, int err_len) {
    std::vector<int> s, e;
    if (!ffv::parse_index_list(sidx, s) || !ffv::parse_index_list(eidx, e)) return -2;
    const ffv::Selection sel = ffv::select_frames(total_src, total_enc, start, end, every, offset, s, e);
    snprintf(err, (size_t)err_len, "%s", sel.error.c_str());
    if (!sel.error.empty()) return -1;
    const int n = (int)sel.source.size();
    for (int i = 0; i < n && i < cap; i++) src_out[i] = sel.source[i], enc_out[i] = sel.encoded[i];
    return n;
}
int ffv_stats(const float* data, int n, const char* label, char* out, int out_len) {
    const std::string s = ffv::stats_block(std::vector<float>(data, data + n), label);
    snprintf(out, (size_t)out_len, "%s", s.c_str());
    return (int)s.size();
}
int ffv_json(const float* data, int n, char* out, int out_len) {
    const std::string s = ffv::json_scores(std::vector<float>(data, data + n));
    snprintf(out, (size_t)out_len, "%s", s.c_str());
    return (int)s.size();
}
int ffv_json_norms(const float* data, int nframes, char* out, int out_len) {
    const std::string s = ffv::json_norms(std::vector<float>(data, data + 3 * nframes));
    snprintf(out, (size_t)out_len, "%s", s.c_str());
    return (int)s.size();
}
int ffv_metric(const char* name) { return ffv::metric_of(name); }
}
#else
int main(int argc, char** argv) {
    ffv::Options o;
    const std::string err = ffv::parse_args(argc, argv, o);
    if (!err.empty()) {
        std::cerr << err << "\n" << ffv::kUsage;
        return 1;
    }
    if (o.help) {
        std::cout << ffv::kUsage;
        return 0;
    }
    if (o.version) {
        std::cout << "FFVship 3.1.0-b" << std::endl;
        std::cout << "vship_amd: SSIMULACRA2 and Butteraugli on MI355X, YUV4MPEG2 inputs" << std::endl;
        return 0;
    }
    if (o.list_gpu) return ffv::list_gpus();
    if (o.source.empty() || o.encoded.empty()) {
        std::cerr << "--source and --encoded are required\n" << ffv::kUsage;
        return 1;
    }
    return ffv::run(o);
}
#endif
