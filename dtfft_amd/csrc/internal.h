// Internal declarations shared by the host-side sources of libdtfft_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dtfft_hip.h"

namespace dtfft_hip {

// Error propagation: functions return int32 codes; HIP_TRY converts a failed
// HIP call into DTFFT_HIP_ERROR_HIP_RUNTIME after recording the message.
void set_last_error(const std::string& msg);
const std::string& last_error();

#define HIP_TRY(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::dtfft_hip::set_last_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return DTFFT_HIP_ERROR_HIP_RUNTIME;                                            \
        }                                                                                  \
    } while (0)

#define TRY(call)                          \
    do {                                   \
        int32_t rc_ = (call);              \
        if (rc_ != DTFFT_HIP_SUCCESS) return rc_; \
    } while (0)

// ---- device / runtime -------------------------------------------------------------
struct device_info {
    bool present = false;      // a gfx950 device is visible and usable
    int device = -1;
    std::string arch;          // e.g. "gfx950:sramecc+:xnack-"
    int num_cus = 0;
    size_t lds_per_block = 0;
};
const device_info& device();
bool compile_only();           // no launches / allocations (no device or forced)
// The configuration set by dtfft_hip_set_config (process-wide, as dtfft_set_config).
const dtfft_hip_config_t& global_config();

// Times autotuner candidates, each a callable that enqueues one launch on `s` (0 on success):
// measure_warmup_iters untimed launches each, then rounds in which every candidate in turn
// runs measure_iters timed launches (src/dtfft_kernel_device.F90:348-349 takes the same two
// settings).  Returns each candidate's median time per launch over the rounds, in ms (1e30
// where a launch failed).  Interleaving the candidates keeps the GPU's clock drift under
// sustained load (durations moving by 20 % over a few launches) out of the comparison.
std::vector<double> time_candidates(const std::vector<std::function<int32_t()>>& runs, hipStream_t s,
                                    int rounds = 3);
// Fills an autotuner scratch buffer with finite pseudo-random values (FP32 words, or FP64
// values when dbl) on stream s.
int32_t fill_tuning_scratch(void* p, size_t bytes, bool dbl, hipStream_t s);

// ---- hipRTC program cache ----------------------------------------------------------
// A compiled kernel: module + function.  `source` is one of the embedded device sources
// ("fft", "transpose"), `expr` a name expression such as "box_copy<8, 2>".
struct rtc_kernel {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    std::string expr;
};
// Compiles (or loads from cache) and returns the kernel; nullptr with an error code on failure.
int32_t rtc_get(const std::string& source, const std::string& expr, const rtc_kernel** out);
std::string cache_dir();

// Launch helper: kernel args are passed as one packed struct-free pointer array.
int32_t launch(const rtc_kernel* k, unsigned grid, unsigned block, hipStream_t stream, void** args);

// ---- tuning records (persisted next to the code objects) ---------------------------
bool tune_lookup(const std::string& key, std::string& value);
void tune_store(const std::string& key, const std::string& value);

// ---- FFT stage (used by the executor plugin, plans and the batched cycle) ----------
struct lines_desc {  // mirrors dtfft::lines_desc in device/fft.hip
    int64_t nl, nb;
    int64_t in_ls, in_bs, in_ns;
    int64_t out_ls, out_bs, out_ns;
    int64_t ntl, ntiles;
    int64_t in_seg_stride;  // with fft_config::seg > 0: input chunks of `seg` elements this far apart
    int64_t in_ts;           // column-in tile stride of a chunk-blocked input (0: lines side by side)
    int64_t out_seg_stride;  // with fft_config::oseg > 0: output chunks of `oseg` elements this far apart
};

// Launch configuration of one batched 1D FFT kernel.
struct fft_config {
    int n = 0;
    std::vector<int> radices;
    int tpl = 0;   // threads per line
    int lpb = 0;   // lines per block
    int pad = 0;   // LDS line padding (complex elements)
    bool in_cols = false, out_cols = false;
    bool dbl = false;
    int kind = 0;  // 0: C2C, 1: R2C (n = 2 * this->n real inputs), 2: C2R
    int seg = 0;   // > 0: FFT axis read in segments of `seg` elements (fused unpack)
    int dma = 0;  // column-in passes through LDS filled by direct loads: 1 double buffer
                  // (fft_cols_dma_kernel), 2 single buffer, two blocks per CU (fft_cols_dma1_kernel)
    int twl = 0;  // C2C (kind 0, dma 0): persistent blocks with the twiddle table in LDS
    int oseg = 0;  // C2C and R2C rows (kind 0 / 1, dma 0): FFT axis written in chunks of `oseg` elements
    // fft_any_kernel (device/anylen.hip): any == 1 runs a line of logical length nlog through
    // the n-point schedule in LDS, directly (chirp false, n == nlog) or by Bluestein's chirp
    // z-transform (chirp true, n >= 2 nlog - 1); kind 0 / 1 / 2 = C2C / R2C / C2R of nlog
    int any = 0;
    int nlog = 0;
    bool chirp = false;
    std::string expr() const;
};

// Twiddle table of one length on the device (generated in FP64 on the GPU).
struct twiddles {
    void* ptr = nullptr;
    int64_t bytes = 0;
    ~twiddles();
};
int32_t get_twiddles(int n, const std::vector<int>& radices, bool dbl, std::shared_ptr<twiddles>* out);
// Bluestein tables of logical length n with work length m (device/anylen.hip).
int32_t get_bluestein(int n, int m, bool dbl, std::shared_ptr<twiddles>* chirp, std::shared_ptr<twiddles>* bh);
// Configuration of fft_any_kernel for kind 0 / 1 / 2 of logical length n (Bluestein unless
// `direct`, which needs n to factorize); false when the work line does not fit in LDS.
bool any_config(int kind, int n, bool dbl, bool direct, fft_config& c);

bool factorize(int n, std::vector<int>& radices);
// LDS bytes of the kernel a configuration names (tile, plus the LDS twiddle tables with twl).
size_t fft_lds_bytes(const fft_config& c);
// Sets c.twl by the rule of `mode` (0 never, 1 default rule, 2 wherever the kernel has the
// form) for a stage of these lines, and only where the LDS budget allows it.
void apply_twl(fft_config& c, const lines_desc& d, int mode);
// Chooses tpl / lpb for a length and access mode (heuristic default of the autotuner).
fft_config default_fft_config(int n, bool dbl, bool in_cols, bool out_cols, bool inplace);
std::vector<fft_config> fft_candidates(int n, bool dbl, bool in_cols, bool out_cols, bool inplace);
// Radix schedules of n for the autotuner (the default first), at most max_count.
std::vector<std::vector<int>> radix_schedules(int n, int max_count);
// Parses a tuning record's "r0,r1,..." schedule of n.
bool parse_radices(const char* text, int n, std::vector<int>& out);

// One batched C2C stage: y = FFT_sign(x) over the lines of `d`.
struct fft_stage {
    fft_config cfg;
    lines_desc d{};
    const rtc_kernel* k = nullptr;
    const rtc_kernel* k_alt = nullptr;  // cfg.dma: the plain column kernel, for misaligned buffers
    std::shared_ptr<twiddles> tw;
    std::shared_ptr<twiddles> rtw;  // R2C / C2R untangling twiddles exp(-2 pi i k / 2M)
    std::shared_ptr<twiddles> chirp, bh;  // cfg.chirp: exp(-i pi j^2 / nlog), DFT_n(b) / n
    unsigned grid = 0, block = 0, grid_alt = 0;
    int32_t init(const fft_config& c, const lines_desc& desc);
    int32_t run(const void* in, void* out, int sign, hipStream_t s) const;
};
// Whether a column-in / row-out stage can use fft_cols_dma_kernel (shape, strides, LDS).
bool dma_ok(const fft_config& c, const lines_desc& d);
lines_desc make_lines(int64_t nl, int64_t nb, int64_t in_ls, int64_t in_bs, int64_t in_ns,
                      int64_t out_ls, int64_t out_bs, int64_t out_ns, int lpb);
// Picks the configuration for a stage (heuristic, or timed when effort >= MEASURE).
int32_t plan_fft_stage(int n, bool dbl, int64_t nl, int64_t nb, int64_t in_ls, int64_t in_bs,
                       int64_t in_ns, int64_t out_ls, int64_t out_bs, int64_t out_ns, bool inplace,
                       int effort, hipStream_t stream, fft_stage* st, int lds_dma = 1);
// Real-to-complex (kind 1) or complex-to-real (kind 2) stage over contiguous lines of n
// reals (n even); idist / odist as abstract_executor passes them (reals for the real side).
int32_t plan_real_stage(int kind, int n, bool dbl, int64_t how_many, int64_t idist, int64_t odist,
                        int effort, hipStream_t stream, fft_stage* st);
// R2R line geometry, mirrors r2r_desc in device/r2r.hip (strides in real elements).
struct r2r_desc_h {
    int64_t nl, nb;
    int64_t in_ls, in_bs, in_es;
    int64_t out_ls, out_bs, out_es;
    int64_t n, m;
    int64_t wdist;
    int64_t in_ts;    // chunk-blocked column input: tile stride (0: natural)
    int64_t out_seg;  // chunk-blocked row output: chunk stride (kernel built with OSEG)
};
// One R2R transform (DCT/DST 1-4, FFTW kinds 3..10) over nl x nb real lines of length n:
// pre-process -> complex FFT of length m -> post-process.  Lines of unit element stride
// with the same layout in and out take one fused launch (fft_r2r_kernel); otherwise three
// launches through a caller-owned work buffer of work_bytes().
struct r2r_stage {
    int kind = 0;
    int64_t n = 0, m = 0;
    bool dbl = false;
    int sign = -1;
    r2r_desc_h d{};
    fft_stage fft;
    const rtc_kernel* pre = nullptr;
    const rtc_kernel* post = nullptr;
    bool fused = false;
    const rtc_kernel* fk = nullptr;
    std::shared_ptr<twiddles> ftw;  // twiddles of the length-m FFT
    std::shared_ptr<twiddles> t4;   // e^{-i pi j / 4n}, j < 2n
    std::shared_ptr<twiddles> rtw;  // half-length form: e^{-2 pi i k / n}, k <= n / 2
    bool half = false;              // fused kernel runs types 2 / 3 through an n/2-point FFT
    unsigned fgrid = 0, fblock = 0;
    int32_t init(int kind, int64_t n, bool dbl, int64_t nl, int64_t nb, int64_t in_ls, int64_t in_bs,
                 int64_t in_es, int64_t out_ls, int64_t out_bs, int64_t out_es, int effort, hipStream_t stream);
    // The fused kernel reading its lines across (nl adjacent lines per batch, element stride
    // in_es) and writing rows (out_es 1) or, with out_cols, columns (out_ls 1) -- the transpose
    // before the stage fused into its load.  Types 2 / 3 of even length through the half-length
    // form; *ok is false (nothing set up) where the stage does not have that form or nl is not a
    // multiple of the 128-byte tile row.
    int32_t init_cols(int kind, int64_t n, bool dbl, int64_t nl, int64_t nb, int64_t in_bs, int64_t in_es,
                      int64_t out_ls, int64_t out_bs, int64_t out_es, bool out_cols, bool* ok);
    // Chunk-blocked pair (fused kernels, half-length form): `rows` writing its lines (nl
    // lines of nb batches, rows of n) in chunks of B elements -- element k of line (l, b) at
    // b * n * nl + l * B + (k / B) * nl * B + k % B -- so that the column stage `cols` (B lines
    // per tile, FFT length nl, lines n per batch) reads each tile as one contiguous block.
    // *ok is false (nothing set up) where the stages have no such form.
    static int32_t blocked_pair(const r2r_stage& rows, const r2r_stage& cols, r2r_stage* rows_blk,
                                r2r_stage* cols_blk, bool* ok);
    std::string expr;  // the fused kernel's name expression
    int lpb = 0;       // lines per tile of the fused kernel
    size_t work_bytes() const;
    int32_t run(const void* in, void* out, void* work, hipStream_t s) const;
};
bool r2r_kind_valid(int kind);
int r2r_inverse_kind(int kind);
int32_t get_r2r_table(int64_t n, bool dbl, std::shared_ptr<twiddles>* out);

// C2R along lines read across (fused transpose): nl lines adjacent in the input, batches
// in_bs apart, complex element stride in_ns; real output rows of n values viewed as n/2
// complex, line stride out_ls and batch stride out_bs in those complex units.
// lds_dma != 0: the LDS-DMA kernel (fft_c2r_dma_kernel) where the stage allows it.
int32_t plan_c2r_cols_stage(int n, bool dbl, int64_t nl, int64_t nb, int64_t in_bs, int64_t in_ns, int64_t out_ls,
                            int64_t out_bs, fft_stage* st, int lds_dma = 1);
int32_t get_rtwiddles(int n, bool dbl, std::shared_ptr<twiddles>* out);

// ---- allocation-free execute (noalloc.hip) ----------------------------------------
// Bytes of the per-device scratch array that dtfft_hip_execute_no_alloc uses as aux.
constexpr size_t kNoAllocAuxBytes = DTFFT_HIP_NO_ALLOC_AUX_BYTES;
constexpr int kNoAllocMaxDevices = 64;
// Device address of the calling thread's current device's scratch array.
int32_t no_alloc_scratch(void** out);

// ---- exact data movement (kernel plugin) ----------------------------------------
struct box_desc {  // mirrors box_desc in device/transpose.hip
    int64_t e0, e1, e2;
    int64_t is1, is2;
    int64_t os0, os1, os2;
    int64_t din, dout;
    int64_t t0, t1;
    int64_t nblocks;
};
struct box_move {
    box_desc d{};
    bool transpose = false;
    int es = 0;
};
// Launches one box move with the tuned kernel for its element size and shape class.
struct mover {
    int es = 0;
    bool transpose = false;
    int tile = 0, rows = 0, pad = 0, vec = 1;
    const rtc_kernel* k = nullptr;
    const rtc_kernel* k_scalar = nullptr;  // box_copy without vector access (unaligned boxes)
    int32_t init(int es, bool transpose, int tile, int rows, int pad, int vec);
    int32_t run(const box_desc& d, const void* in, void* out, hipStream_t s) const;
};
int32_t tune_mover(const box_move& m, int effort, bool force_effort, hipStream_t stream, mover* out);

}  // namespace dtfft_hip
