// Batch engine (host side). See engine.hpp.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "host_stage.hpp"
#include "ssimu2_kernels.hpp"

namespace vship {

namespace {

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

int input_bytes_per_sample(int fmt) {
    if (fmt_yuv(fmt)) return fmt_wide(fmt) ? 2 : 1;
    switch (fmt) {
    case IN_RGB_U8: return 1;
    case IN_RGB_F32: return 4;
    default: return 2;
    }
}

// Samples per row and rows of plane pl (chroma planes of a YUV format are
// ceil(w / 2^sw) x ceil(h / 2^sh)).
int plane_w(const PairInput& p, int pl) { return pl > 0 ? (p.w + (1 << fmt_csw(p.fmt)) - 1) >> fmt_csw(p.fmt) : p.w; }
int plane_h(const PairInput& p, int pl) { return pl > 0 ? (p.h + (1 << fmt_csh(p.fmt)) - 1) >> fmt_csh(p.fmt) : p.h; }

// Every plane must lie inside one device allocation with rows that hold the
// image: a bad pointer or stride becomes an error here instead of a GPU fault.
void validate_planes(const PairInput& p) {
    const int bps = input_bytes_per_sample(p.fmt);
    for (int im = 0; im < 2; im++)
        for (int pl = 0; pl < 3; pl++) {
            const int64_t row = (int64_t)plane_w(p, pl) * bps;
            if (!p.src[im][pl] || p.stride[im][pl] < row) VSHIP_THROW(E_BadPointer);
            require_device_rows(p.src[im][pl], p.stride[im][pl], plane_h(p, pl), row);
        }
}

}  // namespace

// Normalised 17-tap Gaussian, sigma 1.5, radius 8 (src/ssimu2/gaussianblur.hpp:7-18,
// src/util/preprocessor.hpp:89-92), recomputed in double and checked against
// the compile-time taps the kernels use.
Taps make_taps() {
    double g[17], sum = 0.0;
    for (int i = 0; i < 17; i++) {
        const double k = i - 8;
        g[i] = std::exp(-k * k / (2.0 * 1.5 * 1.5));
        sum += g[i];
    }
    Taps t;
    for (int i = 0; i < 17; i++) {
        t.g[i] = (float)(g[i] / sum);
        if (t.g[i] != tap(i)) VSHIP_THROW(E_BadDeviceCode);
    }
    return t;
}

void require_device_rows(const void* ptr, int64_t stride, int rows, int64_t row_bytes) {
    if (!ptr || rows < 1 || row_bytes < 1 || stride < row_bytes) VSHIP_THROW(E_BadPointer);
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) VSHIP_THROW(E_BadPointer);
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) VSHIP_THROW(E_BadPointer);
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr) != hipSuccess) VSHIP_THROW(E_BadPointer);
    const uint8_t* b = reinterpret_cast<const uint8_t*>(base);
    const uint8_t* p = reinterpret_cast<const uint8_t*>(ptr);
    const int64_t extent = (int64_t)(rows - 1) * stride + row_bytes;
    // the kernels address a plane with 32-bit byte offsets
    if (extent >= (int64_t(1) << 32)) VSHIP_THROW(E_BadPointer);
    const uint8_t* end = p + extent;
    if (p < b || end > b + size) VSHIP_THROW(E_BadPointer);
}

void matrix_coefs(int matrix, float& kr, float& kb) {
    switch (matrix) {
    case 1: kr = 0.299f; kb = 0.114f; break;       // BT.601
    case 2: kr = 0.2627f; kb = 0.0593f; break;     // BT.2020 non-constant luminance
    default: kr = 0.2126f; kb = 0.0722f; break;    // BT.709
    }
}

PairLayout pair_layout(int w, int h, bool fp16) {
    PairLayout L;
    int64_t off = 0;
    int ww = w, hh = h;
    for (int s = 0; s < kScales; s++) {
        L.sw[s] = ww;
        L.sh[s] = hh;
        L.pitch[s] = (int)round_up(ww, 64);
        L.plane_elems[s] = (int64_t)L.pitch[s] * hh;
        L.scale_off[s] = off;
        off += 3 * L.plane_elems[s] * scale_elem_bytes(s, fp16);
        ww = (ww - 1) / 2 + 1;
        hh = (hh - 1) / 2 + 1;
    }
    L.img_bytes = off;
    return L;
}

int64_t pair_bytes(int w, int h, bool fp16) { return 2 * pair_layout(w, h, fp16).img_bytes; }

int prep_tiles(int w, int h) { return 2 * ((w + kPrepTile - 1) / kPrepTile) * ((h + kPrepTile - 1) / kPrepTile); }

int map_tiles(int w, int h, int th) {
    int n = 0, ww = w, hh = h;
    for (int s = 0; s < kScales; s++) {
        n += ((ww + map_tile_w(th) - 1) / map_tile_w(th)) * ((hh + map_tile_h(th) - 1) / map_tile_h(th));
        ww = (ww - 1) / 2 + 1;
        hh = (hh - 1) / 2 + 1;
    }
    return n;
}

uint64_t map_algorithmic_bytes(int w, int h, bool fp16) {
    uint64_t t = 0;
    int ww = w, hh = h;
    for (int s = 0; s < kScales; s++) {
        t += (uint64_t)ww * hh * scale_elem_bytes(s, fp16);
        ww = (ww - 1) / 2 + 1;
        hh = (hh - 1) / 2 + 1;
    }
    return t * 2 /*images*/ * 3 /*planes*/;
}

uint64_t prep_algorithmic_bytes(const PairInput& p, bool fp16) {
    const uint64_t bps = input_bytes_per_sample(p.fmt);
    uint64_t in = 0;
    for (int pl = 0; pl < 3; pl++) in += bps * (uint64_t)plane_w(p, pl) * plane_h(p, pl);
    return 2 * in + map_algorithmic_bytes(p.w, p.h, fp16);  // read inputs, write every XYB plane once
}

Engine::Engine(const EngineOptions& opt) : opt_(opt) {
    if (opt_.max_pairs < 1) VSHIP_THROW(E_BadPointer);
    if (opt_.map_th != 48 && opt_.map_th != 64 && opt_.map_th != 112 && opt_.map_th != kMapMfma) opt_.map_th = 48;
    HIP_OR_THROW(hipSetDevice(opt_.device), E_BadDeviceArgument);
    HIP_OR_THROW(hipDeviceGetAttribute(&num_cus_, hipDeviceAttributeMultiprocessorCount, opt_.device),
                 E_BadDeviceCode);
    taps_ = make_taps();
    try {
        HIP_OR_THROW(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), E_OutOfVRAM);
        desc_ = desc_layout(opt_.max_pairs);
        const size_t out_bytes = sizeof(double) * 109 * opt_.max_pairs;
        for (int k = 0; k < kDepth; k++) {
            HIP_OR_THROW(hipHostMalloc(reinterpret_cast<void**>(&desc_h_[k]), desc_.bytes, hipHostMallocDefault),
                         E_OutOfRAM);
            std::memset(desc_h_[k], 0, desc_.bytes);
            HIP_OR_THROW(hipHostMalloc(reinterpret_cast<void**>(&out_h_[k]), out_bytes, hipHostMallocDefault),
                         E_OutOfRAM);
            for (auto& e : ev_[k]) HIP_OR_THROW(hipEventCreate(&e), E_OutOfRAM);
            HIP_OR_THROW(hipEventCreateWithFlags(&done_[k], hipEventDisableTiming), E_OutOfRAM);
        }
        HIP_OR_THROW(hipMalloc(reinterpret_cast<void**>(&desc_d_), desc_.bytes), E_OutOfVRAM);
        arena_bytes_ = opt_.arena_bytes / 256 * 256;
        if (arena_bytes_ <= 0) VSHIP_THROW(E_OutOfVRAM);
        HIP_OR_THROW(hipMalloc(&arena_d_, (size_t)arena_bytes_), E_OutOfVRAM);
        max_map_tiles_ = opt_.max_map_tiles;
        if (max_map_tiles_ <= 0) VSHIP_THROW(E_OutOfVRAM);
        HIP_OR_THROW(hipMalloc(reinterpret_cast<void**>(&partials_d_), sizeof(double) * kFeat * max_map_tiles_),
                     E_OutOfVRAM);
        HIP_OR_THROW(hipMalloc(reinterpret_cast<void**>(&out_d_), out_bytes), E_OutOfVRAM);
    } catch (...) {
        release();
        throw;
    }
}

Engine::~Engine() { release(); }

void Engine::release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (copy_) (void)hipStreamSynchronize(copy_);
    stager_.reset();
    for (int k = 0; k < kDepth; k++) {
        if (graph_exec_[k]) (void)hipGraphExecDestroy(graph_exec_[k]);
        if (graph_[k]) (void)hipGraphDestroy(graph_[k]);
        graph_exec_[k] = nullptr;
        graph_[k] = nullptr;
        for (hipEvent_t* e : {&done_[k], &in_ready_[k], &in_free_[k]})
            if (*e) {
                (void)hipEventDestroy(*e);
                *e = nullptr;
            }
        for (auto& e : ev_[k])
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        if (out_h_[k]) (void)hipHostFree(out_h_[k]);
        if (desc_h_[k]) (void)hipHostFree(desc_h_[k]);
        if (in_d_[k]) (void)hipFree(in_d_[k]);
        out_h_[k] = nullptr;
        desc_h_[k] = nullptr;
        in_d_[k] = nullptr;
        in_cap_[k] = 0;
    }
    if (out_d_) (void)hipFree(out_d_);
    if (partials_d_) (void)hipFree(partials_d_);
    if (arena_d_) (void)hipFree(arena_d_);
    if (desc_d_) (void)hipFree(desc_d_);
    if (copy_) (void)hipStreamDestroy(copy_);
    if (stream_) (void)hipStreamDestroy(stream_);
    out_d_ = nullptr;
    partials_d_ = nullptr;
    arena_d_ = nullptr;
    desc_d_ = nullptr;
    copy_ = nullptr;
    stream_ = nullptr;
}

DescLayout desc_layout(int max_pairs) {
    DescLayout L;
    const size_t nseg = (size_t)kScales * max_pairs;
    L.off_pairs = round_up(sizeof(BatchHdr), 64);
    L.off_prep = L.off_pairs + round_up(sizeof(PairDev) * max_pairs, 64);
    L.off_segstart = L.off_prep + round_up(sizeof(int32_t) * (max_pairs + 1), 64);
    L.off_segcount = L.off_segstart + round_up(sizeof(int32_t) * nseg, 64);
    L.off_order = L.off_segcount + round_up(sizeof(int32_t) * nseg, 64);
    L.off_ostart = L.off_order + round_up(sizeof(int32_t) * nseg, 64);
    L.bytes = L.off_ostart + round_up(sizeof(int32_t) * (nseg + 1), 64);
    return L;
}

DescView desc_view(const uint8_t* base, const DescLayout& L) {
    DescView d;
    d.hdr = reinterpret_cast<const BatchHdr*>(base);
    d.pairs = reinterpret_cast<const PairDev*>(base + L.off_pairs);
    d.prep_start = reinterpret_cast<const int32_t*>(base + L.off_prep);
    d.seg_start = reinterpret_cast<const int32_t*>(base + L.off_segstart);
    d.seg_count = reinterpret_cast<const int32_t*>(base + L.off_segcount);
    d.order_seg = reinterpret_cast<const int32_t*>(base + L.off_order);
    d.order_start = reinterpret_cast<const int32_t*>(base + L.off_ostart);
    return d;
}

PlanInfo plan_batch(const PairInput* in, int n, int max_pairs, bool fp16, int map_th, int64_t arena_bytes,
                    int64_t max_map_tiles, uint8_t* block, const DescLayout& DL) {
    if (n < 1 || n > max_pairs || !in) VSHIP_THROW(E_BadPointer);
    PlanInfo P;
    BatchHdr* hdr = reinterpret_cast<BatchHdr*>(block);
    PairDev* pd = reinterpret_cast<PairDev*>(block + DL.off_pairs);
    int32_t* prep_start = reinterpret_cast<int32_t*>(block + DL.off_prep);
    int32_t* seg_start = reinterpret_cast<int32_t*>(block + DL.off_segstart);
    int32_t* seg_count = reinterpret_cast<int32_t*>(block + DL.off_segcount);
    int32_t* order_seg = reinterpret_cast<int32_t*>(block + DL.off_order);
    int32_t* order_start = reinterpret_cast<int32_t*>(block + DL.off_ostart);
    int64_t arena_off = 0;
    int64_t ptiles = 0;
    uint64_t mbytes = 0, pbytes = 0;
    for (int i = 0; i < n; i++) {
        const PairInput& p = in[i];
        if (p.w < 1 || p.h < 1 || p.fmt < IN_RGB_U8 || p.fmt >= IN_NUM_FORMATS) VSHIP_THROW(E_BadPointer);
        if (fmt_wide(p.fmt) && (p.bits < 9 || p.bits > 16)) VSHIP_THROW(E_BadPointer);
        if (p.chroma_loc < CLOC_LEFT || p.chroma_loc > CLOC_TOP) VSHIP_THROW(E_BadPointer);
        validate_planes(p);
        const PairLayout L = pair_layout(p.w, p.h, fp16);
        PairDev& d = pd[i];
        std::memset(&d, 0, sizeof(d));
        for (int im = 0; im < 2; im++)
            for (int pl = 0; pl < 3; pl++) {
                d.src[im][pl] = p.src[im][pl];
                d.stride[im][pl] = p.stride[im][pl];
            }
        d.fmt = p.fmt;
        d.bits = fmt_yuv(p.fmt) ? (fmt_wide(p.fmt) ? p.bits : 8) : 0;
        d.csw = fmt_csw(p.fmt);
        d.csh = fmt_csh(p.fmt);
        d.cloc = fmt_yuv(p.fmt) ? p.chroma_loc : CLOC_LEFT;
        matrix_coefs(p.matrix, d.kr, d.kb);
        d.ikg = 1.0f / (1.0f - d.kr - d.kb);
        d.w = p.w;
        d.h = p.h;
        d.arena_off = arena_off;
        d.img_bytes = L.img_bytes;
        for (int s = 0; s < kScales; s++) {
            d.scale_off[s] = L.scale_off[s];
            d.plane_elems[s] = L.plane_elems[s];
            d.pitch[s] = L.pitch[s];
            d.sw[s] = L.sw[s];
            d.sh[s] = L.sh[s];
        }
        arena_off += 2 * L.img_bytes;
        if (arena_off > arena_bytes) VSHIP_THROW(E_OutOfVRAM);
        prep_start[i] = (int32_t)ptiles;
        ptiles += prep_tiles(p.w, p.h);
        mbytes += map_algorithmic_bytes(p.w, p.h, fp16);
        pbytes += prep_algorithmic_bytes(p, fp16);
    }
    // map segments, group by group: class 0 (FP32 moments) before class 1
    // (FP64), FP32 planes before FP16 planes within a class
    int64_t mtiles = 0;
    int k = 0;
    int nseg_grp[kMapGroups] = {};
    const bool mfma = map_th == kMapMfma;
    for (int g = 0; g < kMapGroups; g++) {
        const int cls = g < 4 ? g / 2 : 0;
        const bool half = g & 1;
        const bool exact = g >= 4;
        for (int i = 0; i < n; i++)
            for (int s = 0; s < kScales; s++) {
                const int sw = pd[i].sw[s], sh = pd[i].sh[s];
                const bool float_in = pd[i].fmt == IN_RGB_F16 || pd[i].fmt == IN_RGB_F32;
                if ((mfma && cls == 0 && float_in) != exact) continue;
                if (((int64_t)sw * sh < (mfma ? kSmallSegmentMx : kSmallSegment) ? 1 : 0) != cls) continue;
                if ((fp16 && s < kHalfScales) != half) continue;
                const int th = cls ? kSmallTH : map_th;
                const int tw = map_tile_w(th), tht = map_tile_h(th);
                const int64_t cnt = (int64_t)((sw + tw - 1) / tw) * ((sh + tht - 1) / tht);
                const int seg = i * kScales + s;
                P.cls_bytes[cls] += (uint64_t)sw * sh * 2 * 3 * scale_elem_bytes(s, fp16);
                seg_start[seg] = (int32_t)mtiles;
                seg_count[seg] = (int32_t)cnt;
                order_seg[k] = seg;
                order_start[k] = (int32_t)mtiles;
                k++;
                mtiles += cnt;
                nseg_grp[g]++;
                P.ntiles_grp[g] += (int)cnt;
                if (mtiles > max_map_tiles || mtiles > INT32_MAX) VSHIP_THROW(E_OutOfVRAM);
            }
    }
    order_start[k] = (int32_t)mtiles;
    if (ptiles > INT32_MAX) VSHIP_THROW(E_OutOfVRAM);
    prep_start[n] = (int32_t)ptiles;
    int ngen = 0;
    for (int i = 0; i < n; i++) ngen += prep_generic(pd[i].fmt, pd[i].csw, pd[i].csh, pd[i].cloc) ? 1 : 0;
    hdr->npairs = n;
    hdr->prep_tiles = (int32_t)ptiles;
    hdr->prep_generic_pairs = ngen;
    hdr->map_tiles = (int32_t)mtiles;
    hdr->map_th = map_th;
    for (int g = 0; g < kMapGroups; g++) {
        hdr->nseg[g] = nseg_grp[g];
        hdr->tiles[g] = P.ntiles_grp[g];
        P.ntiles_cls[g < 4 ? g / 2 : 0] += P.ntiles_grp[g];
    }
    P.nprep_generic = ngen;
    P.npairs = n;
    P.ntiles_prep = (int)ptiles;
    P.ntiles_map = (int)mtiles;
    P.map_bytes = mbytes;
    P.prep_bytes = pbytes;
    return P;
}

void launch_batch(const DescView& d, const PlanInfo& P, const LaunchShape& S, void* arena, double* partials,
                  double* scores, double* features, const Taps& taps, hipStream_t st, hipEvent_t* ev) {
    const bool graph = S.graph;
    const int fixed = S.persistent_grid > 0 ? S.persistent_grid : 4 * S.num_cus;
    auto grid_for = [&](int work, int cap_in_graph) {
        if (graph) return cap_in_graph;
        if (S.persistent_grid > 0) return std::max(1, std::min(work, fixed));
        return std::max(1, work);
    };
    const int g_prep = grid_for(P.ntiles_prep, fixed);
    const int g_fin = grid_for(P.npairs, std::min(S.max_pairs, fixed));
    HIP_OR_THROW(launch_prep(S.fp16, false, g_prep, d.pairs, d.prep_start, d.hdr, arena, st), E_BadDeviceCode);
    // pairs on the generic chroma path have their own prep kernel; a captured
    // graph always holds it (it returns at once when the batch has none)
    if (graph || P.nprep_generic > 0)
        HIP_OR_THROW(launch_prep(S.fp16, true, g_prep, d.pairs, d.prep_start, d.hdr, arena, st), E_BadDeviceCode);
    if (ev) HIP_OR_THROW(hipEventRecord(ev[0], st), E_BadDeviceCode);
    // one launch per map group; FP16-plane groups exist only with FP16 storage,
    // and an eager launch is skipped when its group is empty
    const bool mfma = S.map_th == kMapMfma;
    for (int cls = 0; cls < 2; cls++) {
        // on the matrix cores, class 0 also has the exact-split groups of float-input pairs
        const int nexact = (cls == 0 && mfma) ? 2 : 1;
        for (int exact = 0; exact < nexact; exact++)
            for (int half = 0; half < 2; half++) {
                const int g = map_group(cls, half, exact);
                if (half && !S.fp16) continue;
                if (!graph && P.ntiles_grp[g] == 0) continue;
                // an eager launch of few tiles gives each channel of a tile its
                // own workgroup (the same partials): fewer than two tiles per
                // workgroup slot on the matrix-core kernel (4 per CU), and for
                // the vector kernels (as few as 1 workgroup per CU in the FP64
                // class) when the split workgroups fit in two rounds, as for
                // one pair per call
                const bool small = (cls == 0 && mfma) ? P.ntiles_grp[g] < 8 * S.num_cus
                                                      : 3 * P.ntiles_grp[g] <= 2 * S.num_cus;
                const int csplit = (!graph && S.persistent_grid == 0 && small) ? 3 : 1;
                const int grid = grid_for(P.ntiles_grp[g] * csplit, cls == 0 ? fixed : std::max(1, fixed / 4));
                if (cls == 0 && mfma)
                    HIP_OR_THROW(launch_map_mx(half, exact, grid, d.pairs, d.order_seg, d.order_start, d.hdr, arena,
                                               partials, csplit, st),
                                 E_BadDeviceCode);
                else
                    HIP_OR_THROW(launch_map(half, cls, cls ? kSmallTH : S.map_th, grid, d.pairs, d.order_seg,
                                            d.order_start, d.hdr, arena, partials, taps, csplit, st),
                                 E_BadDeviceCode);
            }
        if (ev) HIP_OR_THROW(hipEventRecord(ev[1 + cls], st), E_BadDeviceCode);
    }
    HIP_OR_THROW(launch_finalize(g_fin, d.pairs, d.seg_start, d.seg_count, d.hdr, partials, scores, features, st),
                 E_BadDeviceCode);
    if (ev) HIP_OR_THROW(hipEventRecord(ev[3], st), E_BadDeviceCode);
}

void Engine::enqueue_body(int k) {
    const bool prof = opt_.profile && !opt_.use_graph;
    LaunchShape S;
    S.fp16 = opt_.fp16;
    S.map_th = opt_.map_th;
    S.persistent_grid = opt_.persistent_grid;
    S.graph = opt_.use_graph;
    S.max_pairs = opt_.max_pairs;
    S.num_cus = num_cus_;
    hipEvent_t* ev = ev_[k];
    if (prof) HIP_OR_THROW(hipEventRecord(ev[0], stream_), E_BadDeviceCode);
    HIP_OR_THROW(hipMemcpyAsync(desc_d_, desc_h_[k], desc_.bytes, hipMemcpyHostToDevice, stream_),
                 E_BadDeviceCode);
    if (prof) HIP_OR_THROW(hipEventRecord(ev[1], stream_), E_BadDeviceCode);
    double* scores_d = out_d_;
    double* feats_d = out_d_ + opt_.max_pairs;
    launch_batch(desc_view(desc_d_, desc_), plan_, S, arena_d_, partials_d_, scores_d, feats_d, taps_, stream_,
                 prof ? ev + 2 : nullptr);
    // scores [0, max_pairs) and features [max_pairs, ...) are one block on
    // both sides: one copy brings back the scores and the batch's features
    const int np = opt_.use_graph ? opt_.max_pairs : plan_.npairs;
    HIP_OR_THROW(hipMemcpyAsync(out_h_[k], scores_d, sizeof(double) * (opt_.max_pairs + 108 * (size_t)np),
                                hipMemcpyDeviceToHost, stream_),
                 E_BadDeviceCode);
    if (prof) HIP_OR_THROW(hipEventRecord(ev[6], stream_), E_BadDeviceCode);
}

void Engine::submit(const PairInput* pairs, int n) {
    if (outstanding() >= kDepth) VSHIP_THROW(E_BadHandler);
    HIP_OR_THROW(hipSetDevice(opt_.device), E_BadDeviceArgument);
    // slot k was last used by batch head_ - kDepth, which has been waited for,
    // so its descriptor and result buffers are free
    const int k = (int)(head_ % kDepth);
    plan_ = plan_batch(pairs, n, opt_.max_pairs, opt_.fp16, opt_.map_th, arena_bytes_, max_map_tiles_, desc_h_[k],
                       desc_);
    if (opt_.use_graph) {
        if (!graph_exec_[k]) {
            HIP_OR_THROW(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), E_BadDeviceCode);
            try {
                enqueue_body(k);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_OR_THROW(hipStreamEndCapture(stream_, &graph_[k]), E_BadDeviceCode);
            HIP_OR_THROW(hipGraphInstantiate(&graph_exec_[k], graph_[k], nullptr, nullptr, 0), E_BadDeviceCode);
        }
        HIP_OR_THROW(hipGraphLaunch(graph_exec_[k], stream_), E_BadDeviceCode);
    } else {
        enqueue_body(k);
    }
    HIP_OR_THROW(hipEventRecord(done_[k], stream_), E_BadDeviceCode);
    npairs_[k] = plan_.npairs;
    head_++;
}

void Engine::submit_host(const PairInput* pairs, int n) {
    if (outstanding() >= kDepth) VSHIP_THROW(E_BadHandler);
    if (!pairs || n < 1 || n > opt_.max_pairs) VSHIP_THROW(E_BadPointer);
    HIP_OR_THROW(hipSetDevice(opt_.device), E_BadDeviceArgument);
    // one block per plane: (rows - 1) strides plus one row
    std::vector<const uint8_t*> src((size_t)n * 6);
    std::vector<size_t> blk((size_t)n * 6);
    for (int i = 0; i < n; i++) {
        const PairInput& p = pairs[i];
        if (p.w < 1 || p.h < 1 || p.fmt < 0 || p.fmt > IN_YUV410_U16) VSHIP_THROW(E_BadPointer);
        const int64_t bps = input_bytes_per_sample(p.fmt);
        for (int im = 0; im < 2; im++)
            for (int pl = 0; pl < 3; pl++) {
                const int64_t row = (int64_t)plane_w(p, pl) * bps;
                const int64_t st = p.stride[im][pl];
                if (!p.src[im][pl] || st < row) VSHIP_THROW(E_BadPointer);
                const size_t j = (size_t)i * 6 + im * 3 + pl;
                src[j] = p.src[im][pl];
                blk[j] = (size_t)st * (plane_h(p, pl) - 1) + (size_t)row;
            }
    }
    if (!copy_) {
        HIP_OR_THROW(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), E_OutOfVRAM);
        for (int k = 0; k < kDepth; k++) {
            HIP_OR_THROW(hipEventCreateWithFlags(&in_ready_[k], hipEventDisableTiming), E_OutOfRAM);
            HIP_OR_THROW(hipEventCreateWithFlags(&in_free_[k], hipEventDisableTiming), E_OutOfRAM);
        }
        stager_ = std::make_unique<HostStager>(opt_.device, (size_t)opt_.host_chunk_bytes, opt_.host_slots,
                                               opt_.host_copy_threads);
    }
    // device layout: page-locked planes of one allocation keep their host
    // spacing (one DMA for the lot), pageable ones are packed
    const UploadPlan U = plan_upload(src.data(), blk.data(), n * 6);
    const int k = (int)(head_ % kDepth);
    if (U.total > in_cap_[k]) {
        // the batch that last read this buffer has been waited for
        if (in_d_[k]) (void)hipFree(in_d_[k]);
        in_d_[k] = nullptr;
        in_cap_[k] = 0;
        HIP_OR_THROW(hipMalloc(reinterpret_cast<void**>(&in_d_[k]), U.total), E_OutOfVRAM);
        in_cap_[k] = U.total;
    }
    // the upload into buffer k starts after the kernels that read it last
    HIP_OR_THROW(hipStreamWaitEvent(copy_, in_free_[k], 0), E_BadDeviceCode);
    stager_->upload(in_d_[k], U, copy_);
    std::vector<PairInput> dev(pairs, pairs + n);
    for (int i = 0; i < n; i++)
        for (int im = 0; im < 2; im++)
            for (int pl = 0; pl < 3; pl++) dev[i].src[im][pl] = in_d_[k] + U.off[(size_t)i * 6 + im * 3 + pl];
    HIP_OR_THROW(hipEventRecord(in_ready_[k], copy_), E_BadDeviceCode);
    HIP_OR_THROW(hipStreamWaitEvent(stream_, in_ready_[k], 0), E_BadDeviceCode);
    submit(dev.data(), n);
    HIP_OR_THROW(hipEventRecord(in_free_[k], stream_), E_BadDeviceCode);
}

void Engine::set_host_staging(int64_t chunk_bytes, int slots, int copy_threads) {
    if (stager_) VSHIP_THROW(E_BadHandler);
    opt_.host_chunk_bytes = chunk_bytes;
    opt_.host_slots = slots;
    opt_.host_copy_threads = copy_threads;
}

uint64_t Engine::host_pageable_bytes() const { return stager_ ? stager_->pageable_bytes() : 0; }
uint64_t Engine::host_direct_bytes() const { return stager_ ? stager_->direct_bytes() : 0; }

void Engine::wait(double* scores, double* features) {
    if (outstanding() < 1) VSHIP_THROW(E_BadHandler);
    const int k = (int)(tail_ % kDepth);
    tail_++;
    HIP_OR_THROW(hipEventSynchronize(done_[k]), E_BadDeviceCode);
    const int np = npairs_[k];
    if (scores) std::memcpy(scores, out_h_[k], sizeof(double) * np);
    if (features) std::memcpy(features, out_h_[k] + opt_.max_pairs, sizeof(double) * 108 * np);
    if (opt_.profile && !opt_.use_graph) {
        const hipEvent_t* ev = ev_[k];
        (void)hipEventElapsedTime(&timings_.prep_ms, ev[1], ev[2]);
        (void)hipEventElapsedTime(&timings_.map32_ms, ev[2], ev[3]);
        (void)hipEventElapsedTime(&timings_.map64_ms, ev[3], ev[4]);
        (void)hipEventElapsedTime(&timings_.map_ms, ev[2], ev[4]);
        (void)hipEventElapsedTime(&timings_.finalize_ms, ev[4], ev[5]);
        (void)hipEventElapsedTime(&timings_.total_ms, ev[0], ev[6]);
    }
}

}  // namespace vship
