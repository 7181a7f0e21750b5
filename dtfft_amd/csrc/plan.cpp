// Single-rank plans: the dtFFT plan API (include/dtfft.h) for one process and one GPU.
//
// A plan owns the FFT stages and the transposes of the pencil cycle.  The sequences
// follow src/dtfft_plan.F90:
//   execute_2d      (:877-913)   FFT_X, X->Y, FFT_Y
//   execute_z_slab  (:974-1016)  2D FFT over X,Y, X->Z, FFT_Z
//   execute_generic (:1057-1101) FFT_X, X->Y, FFT_Y, Y->Z, FFT_Z
// with the reference's buffer ping-pong between in / aux / out.  On one rank every
// transpose is a single permute kernel (src/dtfft_reshape_handle_generic.F90:232-253):
// X->Y, Y->Z, Z->X are PERMUTE_FORWARD, the reverse directions PERMUTE_BACKWARD.
//
// With enable_fused (default for C2C plans) each transpose is fused into the FFT pass
// that follows it: the kernel reads the pencil with the FFT axis strided and writes the
// next pencil with the FFT axis contiguous, so a 3D direction is three HBM passes
// (in -> out -> aux -> out) instead of five.  R2C plans run the plain sequence.
#include <algorithm>
#include <cstring>

#include "internal.h"

extern "C" int32_t dtfft_hip_kernel_create(int32_t, int32_t, const int32_t*, int64_t, const int32_t*, int32_t,
                                           int32_t, int32_t, void**);
extern "C" int32_t dtfft_hip_kernel_execute(void*, const void*, void*, void*, int32_t, int32_t);
extern "C" int32_t dtfft_hip_kernel_destroy(void*);

namespace dtfft_hip {

static dtfft_hip_config_t g_config = {0, 1, 0, 2, 5, nullptr, 1, 1, 1};

const dtfft_hip_config_t& global_config() { return g_config; }

struct permute {
    void* h = nullptr;
    permute() = default;
    permute(const permute&) = delete;
    permute& operator=(const permute&) = delete;
    ~permute() { if (h) dtfft_hip_kernel_destroy(h); }
    int32_t create(int kt, int ndims, int64_t a, int64_t b, int64_t c, int es, int effort) {
        int32_t d[3] = {(int32_t)a, (int32_t)b, (int32_t)c};
        return dtfft_hip_kernel_create(kt, ndims, d, es, nullptr, 0, effort, 0, &h);
    }
    int32_t run(const void* in, void* out, hipStream_t s) const {
        if (!h) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
        return dtfft_hip_kernel_execute(h, in, out, (void*)s, 0, 0);
    }
};

struct plan_t {
    int ndims = 0;
    int64_t nx = 1, ny = 1, nz = 1;  // global dims, X fastest
    int64_t nc = 1;                  // complex extent of X (nx/2+1 for R2C)
    bool r2c = false, dbl = false, fused = false, z_slab = false;
    bool transpose_only = false;
    int effort = 0;
    int lds_dma = 1;
    hipStream_t stream = nullptr;
    size_t elem_size = 8;            // dtfft_get_element_size: complex bytes, reals for R2C
    size_t cplx = 8;                 // bytes of one complex element
    size_t alloc_elems = 0;
    void* aux = nullptr;             // allocated lazily, src/dtfft_plan.F90:2286-2331

    permute t_xy, t_yx, t_yz, t_zy, t_xz, t_zx;
    fft_stage fx, fx_c2r, fy, fz, fy_cols;           // plain stages
    fft_stage u_xy, u_yz, u_zx, u_yx;                // fused stages (u_yx: 2D only)
    fft_stage fy_zc, u_zx_c2r, u_yx_c2r;             // fused R2C backward (3D, 3D, 2D)
    // chunk-blocked pairs (row pass -> column-in pass) of the fused C2C sequences:
    // forward fx -> u_xy, backward fz -> u_zx (3D) / fy -> u_yx (2D)
    fft_stage fx_blk, u_xy_blk, fb_blk, ub_blk;
    bool blk_fwd = false, blk_bwd = false;
    bool r2r = false;
    int kind[3] = {0, 0, 0};                         // R2R kinds along X, Y, Z
    r2r_stage rx, ry, rz, rxb, ryb, rzb;             // R2R forward / backward stages
    // fused R2R sequence: the transposes folded into the stages reading across lines
    r2r_stage ryc, rzc, rxbc, rybc;
    bool r2r_fused = false;
    // chunk-blocked pairs of the fused R2R sequence: forward rx -> ryc, backward rzb -> rxbc
    // (3D) / ryb -> rxbc (2D)
    r2r_stage rx_blk, ryc_blk, rb_blk, rbc_blk;
    bool r2r_blk_fwd = false, r2r_blk_bwd = false;
    void* work = nullptr;                            // complex work lines of the R2R stages
    // forward + backward captured as one hipGraph for one set of pointers
    hipStream_t gstream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    void* gptr[4] = {nullptr, nullptr, nullptr, nullptr};

    plan_t() = default;
    plan_t(const plan_t&) = delete;
    ~plan_t() {
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        if (gstream) hipStreamDestroy(gstream);
        if (aux) hipFree(aux);
        if (work) hipFree(work);
    }
};

static int32_t ensure_aux(plan_t* p, void*& aux) {
    if (aux) return DTFFT_HIP_SUCCESS;
    if (!p->aux) {
        if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
        if (hipMalloc(&p->aux, std::max<size_t>(1, p->alloc_elems * p->elem_size)) != hipSuccess) {
            p->aux = nullptr;
            return DTFFT_HIP_ERROR_ALLOC_FAILED;
        }
    }
    aux = p->aux;
    return DTFFT_HIP_SUCCESS;
}

// Contiguous C2C stage over `lines` lines of length n.
static int32_t contig(plan_t* p, int64_t n, int64_t lines, fft_stage* st) {
    return plan_fft_stage((int)n, p->dbl, lines, 1, n, 0, 1, n, 0, 1, false, p->effort, p->stream, st);
}

// Fused transpose + FFT: lines (l, b) with l contiguous in the input, FFT axis stride ns;
// output written with the FFT axis contiguous.
static int32_t fused(plan_t* p, int64_t n, int64_t nl, int64_t nb, int64_t in_bs, int64_t in_ns, int64_t out_ls,
                     int64_t out_bs, fft_stage* st) {
    return plan_fft_stage((int)n, p->dbl, nl, nb, 1, in_bs, in_ns, out_ls, out_bs, 1, false, p->effort, p->stream,
                          st, p->lds_dma);
}

// Chunk-blocked variant of a (row pass, column-in pass) pair.  `cols` reads the natural output
// of `rows`: nl = rows' length lines side by side (in_ls 1), its FFT axis rows' lines
// (in_ns = rows' length), its batches whole planes of rows' output.  The blocked pair writes
// each plane as chunks of B = cols' lpb lines: element (line x, row y) of a plane at
// (x / B) * (ny * B) + y * B + x % B, ny = cols' FFT length.  A column tile is then one
// contiguous block of ny * B elements (B * ny * 8 bytes, 64 KiB at 512 FP32 points), which the
// LDS-DMA column kernel streams instead of gathering 128-byte pieces of ny rows; the row pass
// writes B-element pieces of its lines, each continuing the previous line's piece (runs of
// LPB * B elements per block).  The pair is built where cols runs a column-in kernel (LDS-DMA,
// whose tile row is B elements = 128 bytes, or the plain one) and B divides its lines.
static int32_t blocked_pair(plan_t* p, const fft_stage& rows, const fft_stage& cols, fft_stage* rows_blk,
                            fft_stage* cols_blk, bool* ok) {
    *ok = false;
    const fft_config& rc = rows.cfg;
    const fft_config& cc = cols.cfg;
    const int64_t es = (int64_t)p->cplx;
    if (!global_config().blocked || !rows.k || !cols.k) return DTFFT_HIP_SUCCESS;
    if (rc.kind != 0 || rc.any || rc.dma || rc.seg || rc.in_cols || rc.out_cols || rc.n <= 1) return DTFFT_HIP_SUCCESS;
    if (cc.kind != 0 || cc.any || !cc.in_cols || cc.out_cols || cc.seg || cc.n <= 1) return DTFFT_HIP_SUCCESS;
    const int64_t B = cc.lpb, nrow = rc.n, ny = cc.n;
    const lines_desc& cd = cols.d;
    if ((cc.dma && B * es != 128) || cd.nl != nrow || cd.nl % B != 0 || cd.in_ls != 1 || cd.in_ns != nrow) return DTFFT_HIP_SUCCESS;
    if (cd.nb > 1 && cd.in_bs != nrow * ny) return DTFFT_HIP_SUCCESS;
    if (rows.d.nl * rows.d.nb != ny * cd.nb || rows.d.in_ls != nrow || rows.d.in_ns != 1) return DTFFT_HIP_SUCCESS;
    fft_config r = rc;
    r.oseg = (int)B;
    lines_desc rd = make_lines(ny, cd.nb, nrow, nrow * ny, 1, B, nrow * ny, 1, r.lpb);
    rd.out_seg_stride = ny * B;
    lines_desc bd = cd;
    bd.in_ns = B;
    bd.in_ts = ny * B;
    if (cc.dma && !dma_ok(cc, bd)) return DTFFT_HIP_SUCCESS;
    // On the blocked input the persistent double-buffer kernel is the faster LDS-DMA form:
    // measured on MI355X at 512^3 FP32, 386-395 us per pass against 397-401 us with one tile
    // per block (and 414-418 us for the plain column kernel); on gathered input it is the
    // other way round.  Under the default rule (lds_dma 1) the blocked stage takes it.
    fft_config c = cc;
    if (c.dma == 2 && p->lds_dma == 1) {
        fft_config c2 = c;
        c2.dma = 1;
        if (dma_ok(c2, bd)) c = c2;
    }
    TRY(rows_blk->init(r, rd));
    TRY(cols_blk->init(c, bd));
    *ok = true;
    return DTFFT_HIP_SUCCESS;
}

// Chunk-blocked pair of the fused 3D R2C forward: the R2C rows write the complex X-pencil in
// chunks of B lines (element (x, y) of plane z at z * ncp * ny + (x / B) * ny * B + y * B +
// x % B), padded to ncp = ceil(nc / B) * B lines, so that the X->Y pass reads whole contiguous
// tiles (FP32: by LDS DMA) although nc = nx / 2 + 1 lines do not divide into them (nor do rows
// of nc complex values align to 16 bytes).  The padding lines of the last chunk are never
// written; the X->Y pass transforms them into planes x >= nc of the Y-pencil in aux, which the
// Y->Z pass does not read.  Built where the padding costs at most 1/16 of the lines (the
// plan's buffers grow to ncp * ny * nz complex values).
static int32_t blocked_r2c(plan_t* p, bool* ok) {
    *ok = false;
    const int64_t ny = p->ny, nz = p->nz, nc = p->nc, B = 16;
    const fft_config& rc = p->fx.cfg;
    if (!global_config().blocked || !p->fx.k || rc.kind != 1 || rc.any || rc.dma)
        return DTFFT_HIP_SUCCESS;  // the half-length R2C row kernel (fft_r2c_kernel) only
    const int64_t ncp = (nc + B - 1) / B * B;
    if ((ncp - nc) * 16 > nc) return DTFFT_HIP_SUCCESS;
    fft_stage nat;  // the X->Y pass over the padded rows in the natural layout: its configuration
    TRY(fused(p, ny, ncp, nz, ncp * ny, ncp, ny * nz, ny, &nat));
    // the column pass's tile is B lines: FP32 the LDS-DMA kernel (128-byte pieces), FP64 the
    // plain column kernel (256-byte pieces)
    const fft_config& cc = nat.cfg;
    if (!nat.k || cc.kind != 0 || cc.any || !cc.in_cols || cc.out_cols || cc.seg || cc.lpb != B)
        return DTFFT_HIP_SUCCESS;
    lines_desc bd = nat.d;
    bd.in_ns = B;
    bd.in_ts = ny * B;  // planes stay ncp * ny apart
    fft_config c = cc;
    // the persistent double-buffer LDS-DMA kernel is the faster form on a blocked input
    if (c.dma == 2 && p->lds_dma == 1) {
        fft_config c2 = c;
        c2.dma = 1;
        if (dma_ok(c2, bd)) c = c2;
    }
    if (c.dma && !dma_ok(c, bd)) return DTFFT_HIP_SUCCESS;
    fft_config r = rc;
    r.oseg = (int)B;
    const int64_t hl = p->fx.d.in_ls;  // real row length in complex units (nx / 2)
    lines_desc rd = make_lines(ny, nz, hl, hl * ny, 1, B, ncp * ny, 1, r.lpb);
    rd.out_seg_stride = ny * B;
    TRY(p->fx_blk.init(r, rd));
    TRY(p->u_xy_blk.init(c, bd));
    p->alloc_elems = std::max(p->alloc_elems, (size_t)(2 * ncp * ny * nz));
    *ok = true;
    return DTFFT_HIP_SUCCESS;
}

// R2R stages along X, Y (and Z), each over the rows of its pencil; the backward stages
// use the inverse kinds (src/dtfft_utils.F90:447-471).
static int32_t build_r2r(plan_t* p) {
    const int64_t nx = p->nx, ny = p->ny, nz = p->nz;
    const int E = p->effort;
    auto rows = [&](r2r_stage& st, int kind, int64_t n, int64_t lines) {
        return st.init(kind, n, p->dbl, lines, 1, n, 0, 1, n, 0, 1, E, p->stream);
    };
    TRY(rows(p->rx, p->kind[0], nx, ny * nz));
    TRY(rows(p->ry, p->kind[1], ny, nz * nx));
    TRY(rows(p->rxb, r2r_inverse_kind(p->kind[0]), nx, ny * nz));
    TRY(rows(p->ryb, r2r_inverse_kind(p->kind[1]), ny, nz * nx));
    size_t wb = std::max({p->rx.work_bytes(), p->ry.work_bytes(), p->rxb.work_bytes(), p->ryb.work_bytes()});
    if (p->ndims == 3) {
        TRY(rows(p->rz, p->kind[2], nz, nx * ny));
        TRY(rows(p->rzb, r2r_inverse_kind(p->kind[2]), nz, nx * ny));
        wb = std::max({wb, p->rz.work_bytes(), p->rzb.work_bytes()});
    }
    if (!compile_only()) HIP_TRY(hipMalloc(&p->work, std::max<size_t>(wb, 16)));
    // Fused sequence, as the fused C2C one: forward X rows, X->Y + Y, Y->Z + Z; backward Z
    // rows, Z->X + X, Y in place on the X-pencil (the axes of a separable R2R commute).  Taken
    // when every stage that reads across lines has the fused form (types 2 / 3 of even length,
    // lines in whole 128-byte tile rows); otherwise the reference's sequence with permutes.
    if (p->fused) {
        const int k0 = p->kind[0], k1 = p->kind[1];
        const int i0 = r2r_inverse_kind(k0), i1 = r2r_inverse_kind(k1);
        bool a = false, b = false, c = false, e = false;
        if (p->ndims == 3) {
            const int k2 = p->kind[2];
            TRY(p->ryc.init_cols(k1, ny, p->dbl, nx, nz, nx * ny, nx, ny * nz, ny, 1, false, &a));
            TRY(p->rzc.init_cols(k2, nz, p->dbl, ny, nx, ny * nz, ny, nz * nx, nz, 1, false, &b));
            TRY(p->rxbc.init_cols(i0, nx, p->dbl, nz, ny, nz * nx, nz, nx * ny, nx, 1, false, &c));
            TRY(p->rybc.init_cols(i1, ny, p->dbl, nx, nz, nx * ny, nx, 1, nx * ny, nx, true, &e));
            p->r2r_fused = a && b && c && e;
            if (p->r2r_fused) {
                TRY(r2r_stage::blocked_pair(p->rx, p->ryc, &p->rx_blk, &p->ryc_blk, &p->r2r_blk_fwd));
                TRY(r2r_stage::blocked_pair(p->rzb, p->rxbc, &p->rb_blk, &p->rbc_blk, &p->r2r_blk_bwd));
            }
        } else {
            TRY(p->ryc.init_cols(k1, ny, p->dbl, nx, 1, 0, nx, ny, 0, 1, false, &a));
            TRY(p->rxbc.init_cols(i0, nx, p->dbl, ny, 1, 0, ny, nx, 0, 1, false, &c));
            p->r2r_fused = a && c;
            if (p->r2r_fused) {
                TRY(r2r_stage::blocked_pair(p->rx, p->ryc, &p->rx_blk, &p->ryc_blk, &p->r2r_blk_fwd));
                TRY(r2r_stage::blocked_pair(p->ryb, p->rxbc, &p->rb_blk, &p->rbc_blk, &p->r2r_blk_bwd));
            }
        }
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t build(plan_t* p) {
    const int es = (int)p->cplx;
    const int64_t nx = p->nx, ny = p->ny, nz = p->nz, nc = p->nc;
    const int E = p->effort;
    if (p->ndims == 2) {
        TRY(p->t_xy.create(DTFFT_HIP_KERNEL_PERMUTE_FORWARD, 2, nc, ny, 1, es, E));
        TRY(p->t_yx.create(DTFFT_HIP_KERNEL_PERMUTE_BACKWARD, 2, ny, nc, 1, es, E));
        if (p->transpose_only) return DTFFT_HIP_SUCCESS;
        if (p->r2r) return build_r2r(p);
        if (p->r2c) {
            TRY(plan_real_stage(1, (int)nx, p->dbl, ny, nx, nc, E, p->stream, &p->fx));
            TRY(plan_real_stage(2, (int)nx, p->dbl, ny, nc, nx, E, p->stream, &p->fx_c2r));
        } else {
            TRY(contig(p, nx, ny, &p->fx));
        }
        TRY(contig(p, ny, nc, &p->fy));
        if (p->fused) {
            TRY(fused(p, ny, nc, 1, 0, nc, ny, 0, &p->u_xy));   // X (nc,ny) -> Y (ny,nc)
            if (p->r2c)  // Y (ny,nc) -> X real (nx,ny) + C2R along X
                TRY(plan_c2r_cols_stage((int)nx, p->dbl, ny, 1, 0, ny, nx / 2, 0, &p->u_yx_c2r, p->lds_dma));
            else
                TRY(fused(p, nx, ny, 1, 0, ny, nx, 0, &p->u_yx));   // Y (ny,nx) -> X (nx,ny)
            if (!p->r2c) {
                TRY(blocked_pair(p, p->fx, p->u_xy, &p->fx_blk, &p->u_xy_blk, &p->blk_fwd));
                TRY(blocked_pair(p, p->fy, p->u_yx, &p->fb_blk, &p->ub_blk, &p->blk_bwd));
            }
        }
        return DTFFT_HIP_SUCCESS;
    }
    // 3D pencils: X = (nc, ny, nz), Y = (ny, nz, nc), Z = (nz, nc, ny)
    if (p->z_slab) {
        TRY(p->t_xz.create(DTFFT_HIP_KERNEL_PERMUTE_BACKWARD, 3, nc, ny, nz, es, E));
        TRY(p->t_zx.create(DTFFT_HIP_KERNEL_PERMUTE_FORWARD, 3, nz, nc, ny, es, E));
    } else {
        TRY(p->t_xy.create(DTFFT_HIP_KERNEL_PERMUTE_FORWARD, 3, nc, ny, nz, es, E));
        TRY(p->t_yz.create(DTFFT_HIP_KERNEL_PERMUTE_FORWARD, 3, ny, nz, nc, es, E));
        TRY(p->t_zy.create(DTFFT_HIP_KERNEL_PERMUTE_BACKWARD, 3, nz, nc, ny, es, E));
        TRY(p->t_yx.create(DTFFT_HIP_KERNEL_PERMUTE_BACKWARD, 3, ny, nz, nc, es, E));
    }
    if (p->transpose_only) return DTFFT_HIP_SUCCESS;
    if (p->r2r) return build_r2r(p);
    if (p->r2c) {
        TRY(plan_real_stage(1, (int)nx, p->dbl, ny * nz, nx, nc, E, p->stream, &p->fx));
        TRY(plan_real_stage(2, (int)nx, p->dbl, ny * nz, nc, nx, E, p->stream, &p->fx_c2r));
    } else {
        TRY(contig(p, nx, ny * nz, &p->fx));
    }
    TRY(contig(p, nz, nc * ny, &p->fz));
    if (p->z_slab || p->fused) {
        // Y along the strided axis of the X-pencil (second half of the 2D FFT), in place
        TRY(plan_fft_stage((int)ny, p->dbl, nc, nz, 1, nc * ny, nc, 1, nc * ny, nc, true, E, p->stream, &p->fy_cols,
                           p->lds_dma));
    }
    if (!p->z_slab) TRY(contig(p, ny, nz * nc, &p->fy));
    if (p->fused) {
        // On one rank the generic and the Z-slab cycles end in the same Z-pencil, so the
        // fused sequence picks the axis order whose strided reads stay short (one X or Y
        // row apart): forward X, X->Y+Y, Y->Z+Z; backward Z, Z->X+X, Y in place.
        TRY(fused(p, ny, nc, nz, nc * ny, nc, ny * nz, ny, &p->u_xy));      // X -> Y + FFT_Y
        TRY(fused(p, nz, ny, nc, ny * nz, ny, nz * nc, nz, &p->u_yz));      // Y -> Z + FFT_Z
        if (p->r2c) {
            // backward: Z, then Y in place on the Z-pencil (X must come last), then
            // Z -> X fused with the C2R along X into the real X-pencil rows
            TRY(plan_fft_stage((int)ny, p->dbl, nz * nc, 1, 1, 0, nz * nc, 1, 0, nz * nc, true, E, p->stream,
                               &p->fy_zc, p->lds_dma));
            TRY(plan_c2r_cols_stage((int)nx, p->dbl, nz, ny, nz * nc, nz, ny * (nx / 2), nx / 2, &p->u_zx_c2r,
                                    p->lds_dma));
            TRY(blocked_r2c(p, &p->blk_fwd));
        } else {
            TRY(fused(p, nx, nz, ny, nz * nc, nz, nc * ny, nc, &p->u_zx));  // Z -> X + FFT_X
            TRY(blocked_pair(p, p->fx, p->u_xy, &p->fx_blk, &p->u_xy_blk, &p->blk_fwd));
            TRY(blocked_pair(p, p->fz, p->u_zx, &p->fb_blk, &p->ub_blk, &p->blk_bwd));
        }
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t create_plan(int8_t ndims, const int32_t* dims, int32_t precision, int32_t effort, int32_t executor,
                           bool r2c, dtfft_hip_plan_t* out, bool r2r = false, const int32_t* kinds = nullptr) {
    if (!out) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (ndims != 2 && ndims != 3) return DTFFT_HIP_ERROR_INVALID_N_DIMENSIONS;
    if (!dims) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    for (int i = 0; i < ndims; i++) if (dims[i] <= 0) return DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE;
    if (precision != DTFFT_HIP_SINGLE && precision != DTFFT_HIP_DOUBLE) return DTFFT_HIP_ERROR_INVALID_PRECISION;
    if (effort < DTFFT_HIP_ESTIMATE || effort > DTFFT_HIP_EXHAUSTIVE) return DTFFT_HIP_ERROR_INVALID_EFFORT_FLAG;
    if (executor != DTFFT_HIP_EXECUTOR_NONE && executor != DTFFT_HIP_EXECUTOR_HIP)
        return DTFFT_HIP_ERROR_INVALID_EXECUTOR_TYPE;
    if (r2c && executor == DTFFT_HIP_EXECUTOR_NONE) return DTFFT_HIP_ERROR_R2C_TRANSPOSE_PLAN;
    auto p = std::make_unique<plan_t>();
    p->ndims = ndims;
    // dims are given in reverse order (slowest first), include/dtfft.h:436
    p->nx = dims[ndims - 1];
    p->ny = dims[ndims - 2];
    p->nz = ndims == 3 ? dims[0] : 1;
    p->r2c = r2c;
    p->dbl = precision == DTFFT_HIP_DOUBLE;
    p->transpose_only = executor == DTFFT_HIP_EXECUTOR_NONE;
    p->effort = g_config.enable_kernel_autotune && effort == DTFFT_HIP_ESTIMATE ? DTFFT_HIP_MEASURE : effort;
    p->stream = (hipStream_t)g_config.stream;
    p->lds_dma = g_config.lds_dma;
    p->z_slab = ndims == 3 && g_config.enable_z_slab != 0;
    p->fused = g_config.enable_fused != 0 && !p->transpose_only && !(r2c && (p->nx % 2 != 0 || p->nx < 4));
    p->nc = r2c ? p->nx / 2 + 1 : p->nx;
    p->cplx = p->dbl ? 16 : 8;
    p->elem_size = r2c ? p->cplx / 2 : p->cplx;
    if (r2r) {
        // real elements throughout; one rank, so the generic order (Z-slab gives the same
        // Z-pencil); fused transposes where every stage has the fused form (build_r2r)
        p->r2r = !p->transpose_only;
        p->cplx = p->elem_size = p->dbl ? 8 : 4;
        p->z_slab = false;
        p->fused = p->r2r && g_config.enable_fused != 0;
        if (p->r2r) {
            if (!kinds) return DTFFT_HIP_ERROR_MISSING_R2R_KINDS;
            for (int i = 0; i < ndims; i++) {
                if (!r2r_kind_valid(kinds[i])) return DTFFT_HIP_ERROR_INVALID_R2R_KINDS;
                p->kind[i] = kinds[ndims - 1 - i];  // kinds come in the reversed order of dims
            }
        }
    }
    const int64_t ncplx = p->nc * p->ny * p->nz;
    const int64_t nreal = p->nx * p->ny * p->nz;
    if (ncplx > INT32_MAX || nreal > INT32_MAX) return DTFFT_HIP_ERROR_INDEX_OVERFLOW;
    p->alloc_elems = r2c ? (size_t)std::max(nreal, 2 * ncplx) : (size_t)ncplx;
    TRY(build(p.get()));
    *out = p.release();
    return DTFFT_HIP_SUCCESS;
}

// ---- execution -------------------------------------------------------------------------
// Step runner: executes every step of a sequence, or only step `only` (profiling entry
// point dtfft_hip_plan_execute_stage); counts the steps either way.
struct steps {
    int only = -1;
    int count = 0;
};
#define STEP(r, expr)                                       \
    do {                                                    \
        if ((r).only < 0 || (r).only == (r).count) TRY(expr); \
        (r).count++;                                        \
    } while (0)

static int32_t exec_transpose_only(plan_t* p, void* in, void* out, bool fwd, void* aux, steps& r) {
    hipStream_t s = p->stream;
    if (p->ndims == 2) { STEP(r, fwd ? p->t_xy.run(in, out, s) : p->t_yx.run(in, out, s)); return 0; }
    if (p->z_slab) { STEP(r, fwd ? p->t_xz.run(in, out, s) : p->t_zx.run(in, out, s)); return 0; }
    TRY(ensure_aux(p, aux));
    if (fwd) {
        STEP(r, p->t_xy.run(in, aux, s));
        STEP(r, p->t_yz.run(aux, out, s));
    } else {
        STEP(r, p->t_zy.run(in, aux, s));
        STEP(r, p->t_yx.run(aux, out, s));
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t exec_plain(plan_t* p, void* in, void* out, bool fwd, void* aux, steps& r) {
    hipStream_t s = p->stream;
    const int F = DTFFT_HIP_FFT_FORWARD, B = DTFFT_HIP_FFT_BACKWARD;
    TRY(ensure_aux(p, aux));
    const fft_stage& fxb = p->r2c ? p->fx_c2r : p->fx;
    if (p->ndims == 2) {  // execute_2d, src/dtfft_plan.F90:904-912
        if (fwd) {
            STEP(r, p->fx.run(in, aux, F, s));
            STEP(r, p->t_xy.run(aux, out, s));
            STEP(r, p->fy.run(out, out, F, s));
        } else {
            STEP(r, p->fy.run(in, in, B, s));
            STEP(r, p->t_yx.run(in, aux, s));
            STEP(r, fxb.run(aux, out, B, s));
        }
        return DTFFT_HIP_SUCCESS;
    }
    if (p->z_slab) {  // execute_z_slab, src/dtfft_plan.F90:1007-1015
        if (fwd) {
            STEP(r, p->fx.run(in, aux, F, s));
            STEP(r, p->fy_cols.run(aux, aux, F, s));
            STEP(r, p->t_xz.run(aux, out, s));
            STEP(r, p->fz.run(out, out, F, s));
        } else {
            STEP(r, p->fz.run(in, in, B, s));
            STEP(r, p->t_zx.run(in, aux, s));
            STEP(r, p->fy_cols.run(aux, aux, B, s));
            STEP(r, fxb.run(aux, out, B, s));
        }
        return DTFFT_HIP_SUCCESS;
    }
    if (fwd) {  // execute_generic, src/dtfft_plan.F90:1088-1100
        STEP(r, p->fx.run(in, aux, F, s));
        STEP(r, p->t_xy.run(aux, out, s));
        STEP(r, p->fy.run(out, out, F, s));
        STEP(r, p->t_yz.run(out, aux, s));
        STEP(r, p->fz.run(aux, out, F, s));
    } else {
        STEP(r, p->fz.run(in, aux, B, s));
        STEP(r, p->t_zy.run(aux, in, s));
        STEP(r, p->fy.run(in, in, B, s));
        STEP(r, p->t_yx.run(in, aux, s));
        STEP(r, fxb.run(aux, out, B, s));
    }
    return DTFFT_HIP_SUCCESS;
}

// R2R: the reference's plain sequences (execute_2d / execute_generic) with R2R stages.
static int32_t exec_r2r(plan_t* p, void* in, void* out, bool fwd, void* aux, steps& r) {
    hipStream_t s = p->stream;
    void* w = p->work;
    TRY(ensure_aux(p, aux));
    if (p->r2r_fused) {
        // the chunk-blocked pairs write their intermediate into aux, or into out (3D forward,
        // when out is not the input)
        const bool bf = p->r2r_blk_fwd, bb = p->r2r_blk_bwd;
        if (p->ndims == 2) {
            if (fwd) {
                STEP(r, (bf ? p->rx_blk : p->rx).run(in, aux, w, s));
                STEP(r, (bf ? p->ryc_blk : p->ryc).run(aux, out, w, s));
            } else {
                STEP(r, (bb ? p->rb_blk : p->ryb).run(in, aux, w, s));
                STEP(r, (bb ? p->rbc_blk : p->rxbc).run(aux, out, w, s));
            }
        } else if (fwd) {
            const bool b3 = bf && in != out;
            STEP(r, (b3 ? p->rx_blk : p->rx).run(in, out, w, s));
            STEP(r, (b3 ? p->ryc_blk : p->ryc).run(out, aux, w, s));
            STEP(r, p->rzc.run(aux, out, w, s));
        } else {
            STEP(r, (bb ? p->rb_blk : p->rzb).run(in, aux, w, s));
            STEP(r, (bb ? p->rbc_blk : p->rxbc).run(aux, out, w, s));
            STEP(r, p->rybc.run(out, out, w, s));
        }
        return DTFFT_HIP_SUCCESS;
    }
    if (p->ndims == 2) {
        if (fwd) {
            STEP(r, p->rx.run(in, aux, w, s));
            STEP(r, p->t_xy.run(aux, out, s));
            STEP(r, p->ry.run(out, out, w, s));
        } else {
            STEP(r, p->ryb.run(in, in, w, s));
            STEP(r, p->t_yx.run(in, aux, s));
            STEP(r, p->rxb.run(aux, out, w, s));
        }
        return DTFFT_HIP_SUCCESS;
    }
    if (fwd) {
        STEP(r, p->rx.run(in, aux, w, s));
        STEP(r, p->t_xy.run(aux, out, s));
        STEP(r, p->ry.run(out, out, w, s));
        STEP(r, p->t_yz.run(out, aux, s));
        STEP(r, p->rz.run(aux, out, w, s));
    } else {
        STEP(r, p->rzb.run(in, aux, w, s));
        STEP(r, p->t_zy.run(aux, in, s));
        STEP(r, p->ryb.run(in, in, w, s));
        STEP(r, p->t_yx.run(in, aux, s));
        STEP(r, p->rxb.run(aux, out, w, s));
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t exec_fused(plan_t* p, void* in, void* out, bool fwd, void* aux, steps& r) {
    hipStream_t s = p->stream;
    const int F = DTFFT_HIP_FFT_FORWARD, B = DTFFT_HIP_FFT_BACKWARD;
    TRY(ensure_aux(p, aux));
    // the chunk-blocked pairs write their intermediate into aux, or into out when out is not
    // the input (the forward 3D row pass writes out, which the column pass then reads)
    if (p->ndims == 2) {
        if (fwd) {
            const bool blk = p->blk_fwd;
            STEP(r, (blk ? p->fx_blk : p->fx).run(in, aux, F, s));
            STEP(r, (blk ? p->u_xy_blk : p->u_xy).run(aux, out, F, s));
        } else {
            const bool blk = p->blk_bwd && !p->r2c;
            STEP(r, (blk ? p->fb_blk : p->fy).run(in, aux, B, s));
            STEP(r, (p->r2c ? p->u_yx_c2r : blk ? p->ub_blk : p->u_yx).run(aux, out, B, s));
        }
        return DTFFT_HIP_SUCCESS;
    }
    // 3D, generic or Z-slab (same Z-pencil output on one rank)
    if (fwd) {
        const bool blk = p->blk_fwd && in != out;
        STEP(r, (blk ? p->fx_blk : p->fx).run(in, out, F, s));
        STEP(r, (blk ? p->u_xy_blk : p->u_xy).run(out, aux, F, s));
        STEP(r, p->u_yz.run(aux, out, F, s));
    } else if (p->r2c) {
        STEP(r, p->fz.run(in, aux, B, s));
        STEP(r, p->fy_zc.run(aux, aux, B, s));
        STEP(r, p->u_zx_c2r.run(aux, out, B, s));
    } else {
        const bool blk = p->blk_bwd;
        STEP(r, (blk ? p->fb_blk : p->fz).run(in, aux, B, s));
        STEP(r, (blk ? p->ub_blk : p->u_zx).run(aux, out, B, s));
        STEP(r, p->fy_cols.run(out, out, B, s));
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t exec_steps(plan_t* p, void* in, void* out, bool fwd, void* aux, steps& r) {
    if (p->transpose_only) return exec_transpose_only(p, in, out, fwd, aux, r);
    if (p->r2r) return exec_r2r(p, in, out, fwd, aux, r);
    return p->fused ? exec_fused(p, in, out, fwd, aux, r) : exec_plain(p, in, out, fwd, aux, r);
}

}  // namespace dtfft_hip

using namespace dtfft_hip;

extern "C" {

int32_t dtfft_hip_create_config(dtfft_hip_config_t* c) {
    if (!c) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *c = dtfft_hip_config_t{0, 1, 0, 2, 5, nullptr, 1, 1, 1};
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_set_config(const dtfft_hip_config_t* c) {
    if (!c) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (c->measure_warmup_iters < 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (c->measure_iters <= 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (c->lds_dma < 0 || c->lds_dma > 3) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (c->twl < 0 || c->twl > 2) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (c->blocked < 0 || c->blocked > 1) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    g_config = *c;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_create_plan_c2c(int8_t ndims, const int32_t* dims, int32_t precision, int32_t effort,
                                  int32_t executor, dtfft_hip_plan_t* plan) {
    return create_plan(ndims, dims, precision, effort, executor, false, plan);
}

int32_t dtfft_hip_create_plan_r2c(int8_t ndims, const int32_t* dims, int32_t precision, int32_t effort,
                                  int32_t executor, dtfft_hip_plan_t* plan) {
    return create_plan(ndims, dims, precision, effort, executor, true, plan);
}

int32_t dtfft_hip_create_plan_r2r(int8_t ndims, const int32_t* dims, const int32_t* kinds, int32_t precision,
                                  int32_t effort, int32_t executor, dtfft_hip_plan_t* plan) {
    return create_plan(ndims, dims, precision, effort, executor, false, plan, true, kinds);
}

int32_t dtfft_hip_execute(dtfft_hip_plan_t plan, void* in, void* out, int32_t execute_type, void* aux) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (execute_type != DTFFT_HIP_EXECUTE_FORWARD && execute_type != DTFFT_HIP_EXECUTE_BACKWARD)
        return DTFFT_HIP_ERROR_INVALID_EXECUTE_TYPE;
    if (!in || !out) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (aux && (aux == in || aux == out)) return DTFFT_HIP_ERROR_INVALID_AUX;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    const bool fwd = execute_type == DTFFT_HIP_EXECUTE_FORWARD;
    if (p->transpose_only && in == out) return DTFFT_HIP_ERROR_INPLACE_TRANSPOSE;
    steps r;
    return exec_steps(p, in, out, fwd, aux, r);
}

int32_t dtfft_hip_execute_no_alloc(dtfft_hip_plan_t plan, void* in, void* out, int32_t execute_type) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (execute_type != DTFFT_HIP_EXECUTE_FORWARD && execute_type != DTFFT_HIP_EXECUTE_BACKWARD)
        return DTFFT_HIP_ERROR_INVALID_EXECUTE_TYPE;
    if (!in || !out) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    if (p->transpose_only && in == out) return DTFFT_HIP_ERROR_INPLACE_TRANSPOSE;
    const size_t need = p->alloc_elems * p->elem_size;
    if (need > kNoAllocAuxBytes) {
        set_last_error("plan needs " + std::to_string(need) + " aux bytes, the no-alloc scratch holds " +
                       std::to_string(kNoAllocAuxBytes));
        return DTFFT_HIP_ERROR_SCRATCH_TOO_SMALL;
    }
    void* aux = nullptr;
    TRY(no_alloc_scratch(&aux));
    steps r;
    return exec_steps(p, in, out, execute_type == DTFFT_HIP_EXECUTE_FORWARD, aux, r);
}

int32_t dtfft_hip_execute_cycle_graph(dtfft_hip_plan_t plan, void* in, void* out, void* aux, void* in2) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!in || !out || !in2) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (out == in || out == in2) return DTFFT_HIP_ERROR_INVALID_USAGE;
    if (aux && (aux == in || aux == out || aux == in2)) return DTFFT_HIP_ERROR_INVALID_AUX;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    // the library's aux is allocated before the capture, never inside it
    TRY(ensure_aux(p, aux));
    void* ptrs[4] = {in, out, aux, in2};
    if (!p->gexec || !std::equal(ptrs, ptrs + 4, p->gptr)) {
        if (!p->gstream) {
            HIP_TRY(hipStreamCreateWithFlags(&p->gstream, hipStreamNonBlocking));
        }
        if (p->gexec) { hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
        if (p->graph) { hipGraphDestroy(p->graph); p->graph = nullptr; }
        // the execute sequences launch on the plan's stream: point it at the capture stream
        HIP_TRY(hipStreamBeginCapture(p->gstream, hipStreamCaptureModeThreadLocal));
        const hipStream_t user = p->stream;
        p->stream = p->gstream;
        steps r;
        int32_t rc = exec_steps(p, in, out, true, aux, r);
        if (rc == DTFFT_HIP_SUCCESS) rc = exec_steps(p, out, in2, false, aux, r);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(p->gstream, &g);
        p->stream = user;
        if (rc != DTFFT_HIP_SUCCESS) { if (g) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) { set_last_error("hipStreamEndCapture failed"); return DTFFT_HIP_ERROR_HIP_RUNTIME; }
        p->graph = g;
        HIP_TRY(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
        std::copy(ptrs, ptrs + 4, p->gptr);
    }
    // replayed on the plan's stream: after the work queued there, before what follows it
    HIP_TRY(hipGraphLaunch(p->gexec, p->stream));
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_plan_execute_stage(dtfft_hip_plan_t plan, void* in, void* out, int32_t execute_type, void* aux,
                                     int32_t stage, int32_t* n_stages) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (execute_type != DTFFT_HIP_EXECUTE_FORWARD && execute_type != DTFFT_HIP_EXECUTE_BACKWARD)
        return DTFFT_HIP_ERROR_INVALID_EXECUTE_TYPE;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    steps r;
    r.only = stage < 0 ? INT32_MAX : stage;  // stage < 0: count only
    TRY(exec_steps(p, in, out, execute_type == DTFFT_HIP_EXECUTE_FORWARD, aux, r));
    if (n_stages) *n_stages = r.count;
    if (stage >= r.count) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_transpose(dtfft_hip_plan_t plan, void* in, void* out, int32_t tt, void* aux) {
    (void)aux;
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (in == out) return DTFFT_HIP_ERROR_INPLACE_TRANSPOSE;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    hipStream_t s = p->stream;
    if (p->ndims == 2) {
        if (tt == DTFFT_HIP_TRANSPOSE_X_TO_Y) return p->t_xy.run(in, out, s);
        if (tt == DTFFT_HIP_TRANSPOSE_Y_TO_X) return p->t_yx.run(in, out, s);
        return DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE;
    }
    switch (tt) {
        case DTFFT_HIP_TRANSPOSE_X_TO_Y: return p->z_slab ? DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE : p->t_xy.run(in, out, s);
        case DTFFT_HIP_TRANSPOSE_Y_TO_X: return p->z_slab ? DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE : p->t_yx.run(in, out, s);
        case DTFFT_HIP_TRANSPOSE_Y_TO_Z: return p->z_slab ? DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE : p->t_yz.run(in, out, s);
        case DTFFT_HIP_TRANSPOSE_Z_TO_Y: return p->z_slab ? DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE : p->t_zy.run(in, out, s);
        case DTFFT_HIP_TRANSPOSE_X_TO_Z: return p->z_slab ? p->t_xz.run(in, out, s) : DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE;
        case DTFFT_HIP_TRANSPOSE_Z_TO_X: return p->z_slab ? p->t_zx.run(in, out, s) : DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE;
        default: return DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE;
    }
}

int32_t dtfft_hip_destroy(dtfft_hip_plan_t* plan) {
    if (!plan) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (!*plan) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    delete (plan_t*)*plan;
    *plan = nullptr;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_get_local_sizes(dtfft_hip_plan_t plan, int32_t* in_starts, int32_t* in_counts, int32_t* out_starts,
                                  int32_t* out_counts, size_t* alloc_size) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!in_starts && !in_counts && !out_starts && !out_counts && !alloc_size) return DTFFT_HIP_ERROR_INVALID_USAGE;
    const int nd = p->ndims;
    // Fortran-order counts; the C API reports them reversed (include/dtfft.h:657-667)
    int64_t fin[3] = {p->nx, p->ny, p->nz};
    int64_t fout[3];
    if (nd == 2) { fout[0] = p->ny; fout[1] = p->nc; fout[2] = 1; }
    else { fout[0] = p->nz; fout[1] = p->nc; fout[2] = p->ny; }
    for (int i = 0; i < nd; i++) {
        if (in_starts) in_starts[i] = 0;
        if (out_starts) out_starts[i] = 0;
        if (in_counts) in_counts[i] = (int32_t)fin[nd - 1 - i];
        if (out_counts) out_counts[i] = (int32_t)fout[nd - 1 - i];
    }
    if (alloc_size) *alloc_size = p->alloc_elems;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_get_alloc_size(dtfft_hip_plan_t plan, size_t* alloc_size) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!alloc_size) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *alloc_size = p->alloc_elems;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_get_element_size(dtfft_hip_plan_t plan, size_t* element_size) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!element_size) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *element_size = p->elem_size;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_get_alloc_bytes(dtfft_hip_plan_t plan, size_t* alloc_bytes) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!alloc_bytes) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *alloc_bytes = p->alloc_elems * p->elem_size;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_get_aux_bytes(dtfft_hip_plan_t plan, size_t* aux_bytes) {
    return dtfft_hip_get_alloc_bytes(plan, aux_bytes);
}

int32_t dtfft_hip_get_z_slab_enabled(dtfft_hip_plan_t plan, int32_t* on) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!on) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *on = p->z_slab ? 1 : 0;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_plan_mem_alloc(dtfft_hip_plan_t plan, size_t alloc_bytes, void** ptr) {
    if (!plan) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    return dtfft_hip_mem_alloc(alloc_bytes, ptr);
}

int32_t dtfft_hip_plan_mem_free(dtfft_hip_plan_t plan, void* ptr) {
    if (!plan) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    return dtfft_hip_mem_free(ptr);
}

int32_t dtfft_hip_get_stream(dtfft_hip_plan_t plan, void** stream) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!stream) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *stream = (void*)p->stream;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_report(dtfft_hip_plan_t plan) {
    auto* p = (plan_t*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    std::printf("dtfft_hip plan: %dD %s %s, dims (X fastest) %lld x %lld x %lld, %s%s%s\n", p->ndims,
                p->r2r ? "R2R" : p->r2c ? "R2C" : "C2C", p->dbl ? "double" : "single", (long long)p->nx,
                (long long)p->ny, (long long)p->nz, p->transpose_only ? "transpose-only" : "HIP executor",
                p->z_slab ? ", Z-slab" : "",
                (p->r2r ? p->r2r_fused : p->fused) ? ", fused transposes" : "");
    auto show_r2r = [](const char* name, const r2r_stage& st) {
        if (st.fused && st.fk) std::printf("  %-8s %s\n", name, st.fk->expr.c_str());
    };
    show_r2r("r2r_x", p->rx); show_r2r("r2r_zb", p->rzb); show_r2r("r2r_yb", p->ryb);
    if (p->r2r_fused) {
        show_r2r("xy+r2r", p->ryc); show_r2r("yz+r2r", p->rzc); show_r2r("zx+r2r", p->rxbc);
        show_r2r("r2r_ycol", p->rybc);
        if (p->r2r_blk_fwd) { show_r2r("r2r_frow", p->rx_blk); show_r2r("r2r_fcol", p->ryc_blk); }
        if (p->r2r_blk_bwd) { show_r2r("r2r_brow", p->rb_blk); show_r2r("r2r_bcol", p->rbc_blk); }
    }
    auto show = [](const char* name, const fft_stage& st) {
        if (!st.k) return;
        std::printf("  %-8s %s\n", name, st.cfg.expr().c_str());
    };
    show("fft_x", p->fx); show("fft_y", p->fy); show("fft_z", p->fz); show("fft_ycol", p->fy_cols);
    show("xy+fft", p->u_xy); show("yz+fft", p->u_yz); show("zx+fft", p->u_zx); show("yx+fft", p->u_yx);
    show("fft_yzc", p->fy_zc); show("zx+c2r", p->u_zx_c2r); show("yx+c2r", p->u_yx_c2r);
    show("fwd_rows", p->fx_blk); show("fwd_cols", p->u_xy_blk); show("bwd_rows", p->fb_blk);
    show("bwd_cols", p->ub_blk);
    std::fflush(stdout);
    return DTFFT_HIP_SUCCESS;
}

}  // extern "C"
