// Kernel plugin: the permute / pack / unpack kernels of a transpose on gfx950.
//
// Drop-in for kernel_device (src/dtfft_kernel_device.F90): create follows
// abstract_kernel%create (src/dtfft_abstract_kernel.F90:219-288) -- zero dims make a
// dummy kernel, 2D BACKWARD kernels are remapped to FORWARD -- and kernel_device%create
// (src/dtfft_kernel_device.F90:61-102) -- COPY kernels become device copies.  execute
// follows kernel_device%execute (src/dtfft_kernel_device.F90:104-178): permutes are one
// launch, pipelined kernels one launch for the requested neighbour, the non-pipelined
// PACK / UNPACK / PERMUTE_BACKWARD_END one launch per neighbour.
//
// Every kernel type is expressed as a move of a 3D element box (box_desc) whose index
// maps are those of the host routines in src/include/_dtfft_kernel_host_routines.inc;
// the box is then executed by either a row copy or an LDS-tiled transpose.  Unlike
// the CUDA path of the reference, the host-only UNPACK_FORWARD / UNPACK_BACKWARD kinds
// (routines :757-1021) are supported too.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>

#include "internal.h"

namespace dtfft_hip {

// ---- movers ----------------------------------------------------------------------
// Fallback tile of the 16-byte-lane transpose, for boxes or pointers not aligned to its vectors.
static const int kScalarTile = 32, kScalarRows = 8;

int32_t mover::init(int es_, bool transpose_, int tile_, int rows_, int pad_, int vec_) {
    es = es_; transpose = transpose_; tile = tile_; rows = rows_; pad = pad_; vec = vec_;
    std::string expr;
    if (transpose && vec > 1) {
        expr = "box_transpose_v<" + std::to_string(es) + ", " + std::to_string(vec) + ", " + std::to_string(tile) +
               ", " + std::to_string(rows) + ", " + std::to_string(pad) + ">";
        TRY(rtc_get("transpose", "box_transpose<" + std::to_string(es) + ", " + std::to_string(kScalarTile) + ", " +
                                     std::to_string(kScalarRows) + ", 1>", &k_scalar));
    } else if (transpose) {
        expr = "box_transpose<" + std::to_string(es) + ", " + std::to_string(tile) + ", " +
               std::to_string(rows) + ", " + std::to_string(pad) + ">";
    } else {
        expr = "box_copy<" + std::to_string(es) + ", " + std::to_string(vec) + ">";
        TRY(rtc_get("transpose", "box_copy<" + std::to_string(es) + ", 1>", &k_scalar));
    }
    return rtc_get("transpose", expr, &k);
}

static bool aligned_for(const box_desc& d, int vec, const void* in, const void* out, int es);

// Rows of both sides of a transposed box start on whole vectors: the input rows along axis
// 0 (offsets din + i1 is1 + i2 is2), the output rows along axis 1 (dout + i0 os0 + i2 os2).
static bool aligned_for_transpose(const box_desc& d, int vec, const void* in, const void* out, int es) {
    auto m = [&](int64_t v) { return v % vec == 0; };
    return m(d.din) && m(d.dout) && m(d.is1) && m(d.is2) && m(d.os0) && m(d.os2) && d.os1 == 1 &&
           ((uintptr_t)in % (size_t)(vec * es)) == 0 && ((uintptr_t)out % (size_t)(vec * es)) == 0;
}

int32_t mover::run(const box_desc& din, const void* in, void* out, hipStream_t s) const {
    box_desc d = din;
    unsigned grid, block;
    if (d.e0 <= 0 || d.e1 <= 0 || d.e2 <= 0) return DTFFT_HIP_SUCCESS;
    if (transpose) {
        // 16-byte lanes need every row start on both sides, and both pointers, aligned to
        // the vector (the output is contiguous along axis 1: os1 == 1)
        const bool v = vec > 1 && aligned_for_transpose(d, vec, in, out, es);
        const int t = v || vec == 1 ? tile : kScalarTile;
        d.t0 = (d.e0 + t - 1) / t;
        d.t1 = (d.e1 + t - 1) / t;
        d.nblocks = d.t0 * d.t1 * d.e2;
        block = (unsigned)(v ? (tile / vec) * rows : vec == 1 ? tile * rows : kScalarTile * kScalarRows);
        grid = (unsigned)std::min<int64_t>(d.nblocks, 1 << 22);
        void* args[] = {&out, &in, &d};
        return launch(v || vec == 1 ? k : k_scalar, grid, block, s, args);
    } else {
        const int v = aligned_for(d, vec, in, out, es) ? vec : 1;
        const int64_t chunks = (d.e0 + v * 64 - 1) / (v * 64);
        const int64_t items = chunks * d.e1 * d.e2;  // one wave per item, 4 waves per block
        d.t0 = d.t1 = 0;
        d.nblocks = (items + 3) / 4;
        block = 256;
        grid = (unsigned)std::min<int64_t>(d.nblocks, 1 << 22);
        void* args[] = {&out, &in, &d};
        return launch(v == vec ? k : k_scalar, grid, block, s, args);
    }
}

static bool aligned_for(const box_desc& d, int vec, const void* in, const void* out, int es) {
    if (vec == 1) return true;
    auto m = [&](int64_t v) { return v % vec == 0; };
    return m(d.din) && m(d.dout) && m(d.is1) && m(d.is2) && m(d.os1) && m(d.os2) &&
           ((uintptr_t)in % (size_t)(vec * es)) == 0 && ((uintptr_t)out % (size_t)(vec * es)) == 0;
}

// Default launch shapes; the autotuner may replace them (tune_mover).  Transposes of 4- and
// 8-byte elements move 16-byte vectors per lane (box_transpose_v) over 64 x 64 tiles; 16-byte
// elements use 32 x 32 tiles with 16 rows of threads.  Measured on MI355X over 512^3 permutes
// (forward / backward): 4-byte elements 195 / 203 us with 32 rows of 16-byte lanes against
// 250 us for the best one-element-per-lane shape (32 x 32, 2 rows); 8-byte elements 380 / 393
// us with 16 rows against 400 us (32 x 32, 8 rows).
static void default_shape(int es, bool transpose, int& tile, int& rows, int& pad, int& vec) {
    pad = 1;
    vec = std::max(1, 16 / es);
    if (transpose) {
        tile = es == 16 ? 32 : 64;
        rows = es == 4 ? 32 : 16;
    } else {
        tile = rows = 0;
    }
}

int32_t tune_mover(const box_move& m, int effort, bool force_effort, hipStream_t stream, mover* out) {
    int tile, rows, pad, vec;
    default_shape(m.es, m.transpose, tile, rows, pad, vec);
    int tvec = vec;  // transposes: 16-byte lanes (16 / es) or one element per lane
    const auto& d = m.d;
    const std::string key = std::string("mover|") + (m.transpose ? "T" : "C") + "|" + std::to_string(m.es) +
                            "|" + std::to_string(d.e0) + "x" + std::to_string(d.e1) + "x" + std::to_string(d.e2) +
                            "|" + std::to_string(d.is1) + "," + std::to_string(d.os0);
    std::string rec;
    // A forced ESTIMATE takes the default shape even where a tuning record exists, as
    // get_kernel does with force_effort (src/dtfft_kernel_device.F90:326-335).
    const bool estimate_only = force_effort && effort == DTFFT_HIP_ESTIMATE;
    const char* source = "default";
    if (m.transpose && !estimate_only && tune_lookup(key, rec)) {
        int t, r, p, v = 1;
        const int got = std::sscanf(rec.c_str(), "%d %d %d %d", &t, &r, &p, &v);
        if (got >= 3 && (v == 1 || v * m.es == 16)) { tile = t; rows = r; pad = p; tvec = v; source = "record"; }
    } else if (m.transpose && effort >= DTFFT_HIP_MEASURE && !compile_only()) {
        // time candidate shapes on scratch buffers of the box's extent
        const int64_t span_in = d.din + (d.e0 - 1) + (d.e1 - 1) * d.is1 + (d.e2 - 1) * d.is2 + 1;
        const int64_t span_out = d.dout + (d.e0 - 1) * d.os0 + (d.e1 - 1) * d.os1 + (d.e2 - 1) * d.os2 + 1;
        void *bi = nullptr, *bo = nullptr;
        HIP_TRY(hipMalloc(&bi, (size_t)span_in * m.es));
        if (hipMalloc(&bo, (size_t)span_out * m.es) != hipSuccess) { hipFree(bi); return DTFFT_HIP_ERROR_ALLOC_FAILED; }
        if (fill_tuning_scratch(bi, (size_t)span_in * m.es, false, stream) != DTFFT_HIP_SUCCESS) {
            hipFree(bi);
            hipFree(bo);
            return DTFFT_HIP_ERROR_HIP_RUNTIME;
        }
        double best = 1e30;
        const int tiles[] = {16, 32, 64};
        const int rowss[] = {1, 2, 4, 8, 16, 32};
        const int pads[] = {0, 1, 2};
        // one element per lane, and 16-byte lanes for 4- and 8-byte elements where the box's
        // rows start on whole vectors (the scratch buffers are 16-byte aligned)
        const int V = 16 / m.es;
        const bool vec_ok = V > 1 && aligned_for_transpose(d, V, bi, bo, m.es);
        std::vector<int> vecs = {1};
        if (vec_ok) vecs.push_back(V);
        std::vector<mover> cands;
        std::vector<std::array<int, 4>> shapes;
        for (int v : vecs)
            for (int t : tiles)
                for (int r : rowss)
                    for (int p : pads) {
                        const int threads = (t / v) * r;
                        if (threads > 1024 || threads < 64 || r > t || t % r != 0) continue;
                        if ((size_t)t * (t + p) * m.es > 64 * 1024) continue;
                        if (effort < DTFFT_HIP_EXHAUSTIVE && p == 2) continue;
                        mover cand;
                        if (cand.init(m.es, true, t, r, p, v)) continue;
                        cands.push_back(cand);
                        shapes.push_back({t, r, p, v});
                    }
        std::vector<std::function<int32_t()>> runs;
        for (auto& mv : cands) runs.push_back([&mv, &d, bi, bo, stream]() { return mv.run(d, bi, bo, stream); });
        const std::vector<double> ms = time_candidates(runs, stream);
        for (size_t i = 0; i < cands.size(); i++)
            if (ms[i] < best) {
                best = ms[i];
                tile = shapes[i][0]; rows = shapes[i][1]; pad = shapes[i][2]; tvec = shapes[i][3];
            }
        hipStreamSynchronize(stream);
        hipFree(bi);
        hipFree(bo);
        char buf[64];
        std::snprintf(buf, sizeof buf, "%d %d %d %d", tile, rows, pad, tvec);
        tune_store(key, buf);
        source = "timed";
    }
    if (m.transpose && std::getenv("DTFFT_HIP_DEBUG"))
        std::fprintf(stderr, "[dtfft_hip] %s: tile %d rows %d pad %d vec %d (%s)\n", key.c_str(), tile, rows, pad,
                     tvec, source);
    return out->init(m.es, m.transpose, tile, rows, pad, m.transpose ? tvec : vec);
}

// ---- kernel type -> box -------------------------------------------------------------
// Index maps of _dtfft_kernel_host_routines.inc (0-based).  nd = (nxx,nyy,nzz,din,dout).
static int32_t make_box(int kt, int ndims, const int* dims, const int* nd, box_move& m) {
    const int64_t nx = dims[0], ny = dims[1], nz = ndims == 3 ? dims[2] : 1;
    box_desc& d = m.d;
    d = box_desc{};
    const int64_t nxx = nd ? nd[0] : 0, nyy = nd ? nd[1] : 0, nzz = nd ? (ndims == 3 ? nd[2] : 1) : 0;
    const int64_t din = nd ? nd[3] : 0, dout = nd ? nd[4] : 0;
    switch (kt) {
        case DTFFT_HIP_KERNEL_PERMUTE_FORWARD:  // :140-196
            m.transpose = true;
            if (ndims == 2) { d.e0 = nx; d.e1 = ny; d.e2 = 1; d.is1 = nx; d.os0 = ny; d.os1 = 1; }
            else { d.e0 = nx; d.e1 = ny; d.e2 = nz; d.is1 = nx; d.is2 = nx * ny; d.os0 = ny * nz; d.os1 = 1; d.os2 = ny; }
            return 0;
        case DTFFT_HIP_KERNEL_PERMUTE_BACKWARD:  // :262-302, out(z + x*nz + y*nz*nx)
            m.transpose = true;
            d.e0 = nx; d.e1 = nz; d.e2 = ny; d.is1 = nx * ny; d.is2 = nx; d.os0 = nz; d.os1 = 1; d.os2 = nz * nx;
            return 0;
        case DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_START:  // :351-390, out(z + y*nz + x*nz*ny)
            m.transpose = true;
            d.e0 = nx; d.e1 = nz; d.e2 = ny; d.is1 = nx * ny; d.is2 = nx; d.os0 = nz * ny; d.os1 = 1; d.os2 = nz;
            return 0;
        case DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END:
        case DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END_PIPELINED:  // :439-482
            m.transpose = false;
            d.e0 = nxx; d.e1 = nyy; d.e2 = nzz; d.is1 = nzz * nxx; d.is2 = nxx; d.os0 = 1; d.os1 = nx; d.os2 = nx * ny;
            d.din = din; d.dout = dout;
            return 0;
        case DTFFT_HIP_KERNEL_UNPACK:
        case DTFFT_HIP_KERNEL_UNPACK_PIPELINED:  // :575-636
            m.transpose = false;
            d.e0 = nxx; d.e1 = nyy; d.e2 = nzz; d.is1 = nxx; d.is2 = nxx * nyy; d.os0 = 1; d.os1 = nx; d.os2 = nx * ny;
            d.din = din; d.dout = dout;
            return 0;
        case DTFFT_HIP_KERNEL_PACK:
        case DTFFT_HIP_KERNEL_PACK_PIPELINED:  // :657-721
            m.transpose = false;
            d.e0 = nxx; d.e1 = nyy; d.e2 = nzz; d.is1 = nx; d.is2 = nx * ny; d.os0 = 1; d.os1 = nxx; d.os2 = nxx * nyy;
            d.din = din; d.dout = dout;
            return 0;
        case DTFFT_HIP_KERNEL_PACK_FORWARD:  // :1026-1086
            m.transpose = true;
            if (ndims == 2) { d.e0 = nxx; d.e1 = nyy; d.e2 = 1; d.is1 = nx; d.os0 = nyy; d.os1 = 1; }
            else { d.e0 = nxx; d.e1 = nyy; d.e2 = nzz; d.is1 = nx; d.is2 = nx * ny; d.os0 = nyy * nzz; d.os1 = 1; d.os2 = nyy; }
            d.din = din; d.dout = dout;
            return 0;
        case DTFFT_HIP_KERNEL_PACK_BACKWARD:  // :1155-1198, out(z + x*nzz + y*nzz*nxx)
            m.transpose = true;
            d.e0 = nxx; d.e1 = nzz; d.e2 = nyy; d.is1 = nx * ny; d.is2 = nx; d.os0 = nzz; d.os1 = 1; d.os2 = nzz * nxx;
            d.din = din; d.dout = dout;
            return 0;
        case DTFFT_HIP_KERNEL_UNPACK_FORWARD:
        case DTFFT_HIP_KERNEL_UNPACK_FORWARD_PIPELINED:  // :757-811
            m.transpose = true;
            if (ndims == 2) {  // in(x*nyy + y) -> out(y*nx + x): contiguous input axis is y
                d.e0 = nyy; d.e1 = nxx; d.e2 = 1; d.is1 = nyy; d.os0 = nx; d.os1 = 1;
            } else {  // in(z + x*nzz + y*nzz*nxx) -> out(x + y*nx + z*nx*ny)
                d.e0 = nzz; d.e1 = nxx; d.e2 = nyy; d.is1 = nzz; d.is2 = nzz * nxx; d.os0 = nx * ny; d.os1 = 1; d.os2 = nx;
            }
            d.din = din; d.dout = dout;
            return 0;
        case DTFFT_HIP_KERNEL_UNPACK_BACKWARD:
        case DTFFT_HIP_KERNEL_UNPACK_BACKWARD_PIPELINED:  // :907-946, in(y + z*nyy + x*nzz*nyy)
            m.transpose = true;
            d.e0 = nyy; d.e1 = nxx; d.e2 = nzz; d.is1 = nzz * nyy; d.is2 = nyy; d.os0 = nx; d.os1 = 1; d.os2 = nx * ny;
            d.din = din; d.dout = dout;
            return 0;
        default:
            return DTFFT_HIP_ERROR_INVALID_KERNEL_TYPE;
    }
}

static bool is_pipelined_only(int kt) {
    return kt == DTFFT_HIP_KERNEL_UNPACK_PIPELINED || kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END_PIPELINED ||
           kt == DTFFT_HIP_KERNEL_COPY_PIPELINED || kt == DTFFT_HIP_KERNEL_PACK_BACKWARD ||
           kt == DTFFT_HIP_KERNEL_PACK_FORWARD || kt == DTFFT_HIP_KERNEL_PACK_PIPELINED ||
           kt == DTFFT_HIP_KERNEL_UNPACK_FORWARD_PIPELINED || kt == DTFFT_HIP_KERNEL_UNPACK_BACKWARD_PIPELINED;
}
static bool needs_neighbors(int kt) {
    return !(kt == DTFFT_HIP_KERNEL_PERMUTE_FORWARD || kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD ||
             kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_START || kt == DTFFT_HIP_KERNEL_COPY);
}

struct kernel_handle {
    int kernel_type = 0;
    int ndims = 0;
    int dims[3] = {1, 1, 1};
    int es = 0;
    bool dummy = false;
    std::vector<std::array<int, 5>> nd;
    std::vector<box_move> boxes;   // one per neighbour (or one for permutes)
    std::vector<mover> movers;     // matching boxes
};

}  // namespace dtfft_hip

using namespace dtfft_hip;

extern "C" {

int32_t dtfft_hip_kernel_create(int32_t kernel_type, int32_t ndims, const int32_t* dims, int64_t base_storage,
                                const int32_t* neighbor_data, int32_t n_peers, int32_t effort,
                                int32_t force_effort, void** handle) {
    if (!handle || !dims) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *handle = nullptr;
    if (ndims != 2 && ndims != 3) return DTFFT_HIP_ERROR_INVALID_N_DIMENSIONS;
    if (base_storage != 4 && base_storage != 8 && base_storage != 16) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (effort < DTFFT_HIP_ESTIMATE || effort > DTFFT_HIP_EXHAUSTIVE) return DTFFT_HIP_ERROR_INVALID_EFFORT_FLAG;
    auto h = std::make_unique<kernel_handle>();
    h->ndims = ndims;
    h->es = (int)base_storage;
    int64_t total = 1;
    for (int i = 0; i < ndims; i++) {
        if (dims[i] < 0) return DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE;
        h->dims[i] = dims[i];
        total *= dims[i];
        if (dims[i] == 0) h->dummy = true;  // src/dtfft_abstract_kernel.F90:233-237
    }
    if (total > INT32_MAX) return DTFFT_HIP_ERROR_INDEX_OVERFLOW;  // check_if_overflow
    // 2D remap, src/dtfft_abstract_kernel.F90:272-283
    int kt = kernel_type;
    if (ndims == 2) {
        if (kt == DTFFT_HIP_KERNEL_PACK_BACKWARD) kt = DTFFT_HIP_KERNEL_PACK_FORWARD;
        else if (kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD) kt = DTFFT_HIP_KERNEL_PERMUTE_FORWARD;
        else if (kt == DTFFT_HIP_KERNEL_UNPACK_BACKWARD) kt = DTFFT_HIP_KERNEL_UNPACK_FORWARD;
        else if (kt == DTFFT_HIP_KERNEL_UNPACK_BACKWARD_PIPELINED) kt = DTFFT_HIP_KERNEL_UNPACK_FORWARD_PIPELINED;
    }
    if (ndims == 2 && (kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_START || kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END ||
                       kt == DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END_PIPELINED))
        return DTFFT_HIP_ERROR_INVALID_KERNEL_TYPE;  // 2-step permute is 3D only
    h->kernel_type = kt;
    if (kt == DTFFT_HIP_KERNEL_DUMMY) h->dummy = true;
    if (h->dummy) { *handle = h.release(); return DTFFT_HIP_SUCCESS; }
    if (needs_neighbors(kt)) {
        if (!neighbor_data || n_peers <= 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
        for (int p = 0; p < n_peers; p++) {
            std::array<int, 5> a;
            for (int r = 0; r < 5; r++) a[r] = neighbor_data[p * 5 + r];
            h->nd.push_back(a);
        }
    }
    if (kt == DTFFT_HIP_KERNEL_COPY || kt == DTFFT_HIP_KERNEL_COPY_PIPELINED) {
        *handle = h.release();
        return DTFFT_HIP_SUCCESS;  // device copies, src/dtfft_kernel_device.F90:77-81
    }
    hipStream_t s = nullptr;
    const size_t nb = needs_neighbors(kt) ? h->nd.size() : 1;
    for (size_t p = 0; p < nb; p++) {
        box_move m;
        m.es = h->es;
        int32_t rc = make_box(kt, ndims, h->dims, needs_neighbors(kt) ? h->nd[p].data() : nullptr, m);
        if (rc) return rc;
        mover mv;
        rc = tune_mover(m, effort, force_effort != 0, s, &mv);
        if (rc) return rc;
        h->boxes.push_back(m);
        h->movers.push_back(mv);
    }
    *handle = h.release();
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_kernel_execute(void* handle, const void* in, void* out, void* stream, int32_t neighbor,
                                 int32_t sync) {
    auto* h = (kernel_handle*)handle;
    if (!h) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (h->dummy) return DTFFT_HIP_SUCCESS;
    hipStream_t s = (hipStream_t)stream;
    const int kt = h->kernel_type;
    const int np = (int)h->nd.size();
    if (is_pipelined_only(kt) && (neighbor < 1 || neighbor > np)) return DTFFT_HIP_ERROR_INVALID_NEIGHBOR;
    if (neighbor < 0 || neighbor > np) return DTFFT_HIP_ERROR_INVALID_NEIGHBOR;
    if (kt == DTFFT_HIP_KERNEL_COPY) {
        size_t bytes = (size_t)h->es * (size_t)h->dims[0] * h->dims[1] * h->dims[2];
        HIP_TRY(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, s));
    } else if (kt == DTFFT_HIP_KERNEL_COPY_PIPELINED) {
        auto& a = h->nd[neighbor - 1];
        size_t bytes = (size_t)h->es * (size_t)a[0] * a[1] * a[2];
        HIP_TRY(hipMemcpyAsync((char*)out + (size_t)h->es * a[4], (const char*)in + (size_t)h->es * a[3], bytes,
                               hipMemcpyDeviceToDevice, s));
    } else if (!needs_neighbors(kt)) {
        TRY(h->movers[0].run(h->boxes[0].d, in, out, s));
    } else if (neighbor >= 1) {
        TRY(h->movers[neighbor - 1].run(h->boxes[neighbor - 1].d, in, out, s));
    } else {
        for (int p = 0; p < np; p++) TRY(h->movers[p].run(h->boxes[p].d, in, out, s));
    }
    if (sync) HIP_TRY(hipStreamSynchronize(s));
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_kernel_destroy(void* handle) {
    delete (kernel_handle*)handle;
    return DTFFT_HIP_SUCCESS;
}

}  // extern "C"
