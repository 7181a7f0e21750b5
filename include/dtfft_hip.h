/*
 * dtfft_hip.h -- C-ABI of dtfft_amd, the MI355X (gfx950) implementation of
 * dtFFT's single-GPU hot path.
 *
 * Three groups of entry points, each a drop-in for one reference interface:
 *
 *  1. Kernel plugin (dtfft_hip_kernel_*) replaces `kernel_device`, the CUDA
 *     implementation of `abstract_kernel` (src/dtfft_abstract_kernel.F90:106-163,
 *     src/dtfft_kernel_device.F90:61-178): the local permute / pack / unpack
 *     kernels of a transpose.
 *  2. FFT executor plugin (dtfft_hip_fft_*) replaces the cuFFT / VkFFT
 *     implementations of `abstract_executor` (src/dtfft_abstract_executor.F90:47-112,
 *     src/interfaces/fft/cufft/dtfft_executor_cufft_m.F90:52-125): batched 1D / 2D
 *     C2C and R2C/C2R transforms.
 *  3. Single-rank plan (dtfft_hip_create_plan_*, dtfft_hip_execute, ...) mirrors the
 *     public C API of dtFFT (include/dtfft.h:397-917) for one process and one GPU,
 *     with the transposes optionally fused into the following FFT pass, and the
 *     batched cycle (dtfft_hip_cycle_*), which has no counterpart in the reference.
 *
 * Conventions (as in the reference):
 *  - every call returns an int32 error code, DTFFT_HIP_SUCCESS (0) on success;
 *    codes that exist in dtFFT keep dtFFT's value (include/dtfft_config.h.in:82-151);
 *  - arrays are Fortran-ordered: dims[0] of a kernel is the fastest dimension;
 *    the plan API takes global dims in reverse (C) order like dtfft_create_plan_*;
 *  - FFTs are unnormalised, forward = exp(-2*pi*i*jk/n) (sign -1), backward sign +1;
 *  - all device work is asynchronous on the stream given at creation (NULL = the
 *    default stream); nothing blocks the host unless `sync` is requested;
 *  - element indices are int32 as in the reference; byte offsets are 64-bit.
 *
 * Kernels are compiled at plan creation with hipRTC for gfx950 and cached
 * (in-process and on disk, see dtfft_hip_set_cache_dir).
 */
#ifndef DTFFT_HIP_H
#define DTFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (dtFFT values where dtFFT has the code) ------------------------ */
#define DTFFT_HIP_SUCCESS                        0
#define DTFFT_HIP_ERROR_PLAN_NOT_CREATED         1
#define DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE   2
#define DTFFT_HIP_ERROR_INVALID_N_DIMENSIONS     3
#define DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE   4
#define DTFFT_HIP_ERROR_INVALID_PRECISION        6
#define DTFFT_HIP_ERROR_INVALID_EFFORT_FLAG      7
#define DTFFT_HIP_ERROR_INVALID_EXECUTOR_TYPE    8
#define DTFFT_HIP_ERROR_MISSING_R2R_KINDS        11
#define DTFFT_HIP_ERROR_INVALID_R2R_KINDS        12
#define DTFFT_HIP_ERROR_R2C_TRANSPOSE_PLAN       13
#define DTFFT_HIP_ERROR_INPLACE_TRANSPOSE        14
#define DTFFT_HIP_ERROR_INVALID_AUX              15
#define DTFFT_HIP_ERROR_INVALID_USAGE            17
#define DTFFT_HIP_ERROR_ALLOC_FAILED             19
#define DTFFT_HIP_ERROR_FREE_FAILED              20
#define DTFFT_HIP_ERROR_INVALID_ALLOC_BYTES      21
#define DTFFT_HIP_ERROR_INVALID_EXECUTE_TYPE     43
#define DTFFT_HIP_ERROR_R2R_FFT_NOT_SUPPORTED    101
#define DTFFT_HIP_ERROR_GPU_INVALID_STREAM       201
#define DTFFT_HIP_ERROR_NOT_DEVICE_PTR           300
/* codes of the HIP layer (no counterpart in dtFFT) */
#define DTFFT_HIP_ERROR_HIP_RUNTIME              600  /* a HIP runtime call failed   */
#define DTFFT_HIP_ERROR_RTC_COMPILE              601  /* hipRTC compilation failed   */
#define DTFFT_HIP_ERROR_UNSUPPORTED_SIZE         602  /* FFT length not supported   */
#define DTFFT_HIP_ERROR_INVALID_KERNEL_TYPE      603
#define DTFFT_HIP_ERROR_INVALID_NEIGHBOR         604
#define DTFFT_HIP_ERROR_INDEX_OVERFLOW           605  /* > 2^31-1 elements per buffer */
#define DTFFT_HIP_ERROR_NO_DEVICE                606  /* no gfx950 device visible     */
#define DTFFT_HIP_ERROR_INVALID_ARGUMENT         607
#define DTFFT_HIP_ERROR_SCRATCH_TOO_SMALL        608  /* aux > DTFFT_HIP_NO_ALLOC_AUX_BYTES */

/* ---- enums (dtFFT values) ---------------------------------------------------------- */
/* kernel_type_t, src/dtfft_abstract_kernel.F90:59-98 */
#define DTFFT_HIP_KERNEL_DUMMY                       (-1)
#define DTFFT_HIP_KERNEL_PACK                        1
#define DTFFT_HIP_KERNEL_COPY_PIPELINED              2
#define DTFFT_HIP_KERNEL_UNPACK                      3
#define DTFFT_HIP_KERNEL_COPY                        4
#define DTFFT_HIP_KERNEL_UNPACK_PIPELINED            5
#define DTFFT_HIP_KERNEL_PACK_PIPELINED              6
#define DTFFT_HIP_KERNEL_PERMUTE_FORWARD             7
#define DTFFT_HIP_KERNEL_PERMUTE_BACKWARD            8
#define DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_START      9
#define DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END        10
#define DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END_PIPELINED 11
#define DTFFT_HIP_KERNEL_PACK_FORWARD                12
#define DTFFT_HIP_KERNEL_PACK_BACKWARD               13
#define DTFFT_HIP_KERNEL_UNPACK_FORWARD              15
#define DTFFT_HIP_KERNEL_UNPACK_FORWARD_PIPELINED    16
#define DTFFT_HIP_KERNEL_UNPACK_BACKWARD             17
#define DTFFT_HIP_KERNEL_UNPACK_BACKWARD_PIPELINED   18

/* effort, include/dtfft_config.h.in:57-60 */
#define DTFFT_HIP_ESTIMATE    0
#define DTFFT_HIP_MEASURE     1
#define DTFFT_HIP_PATIENT     2
#define DTFFT_HIP_EXHAUSTIVE  3
/* precision */
#define DTFFT_HIP_SINGLE      0
#define DTFFT_HIP_DOUBLE      1
/* executor: NONE = transpose-only plan; HIP = this library's FFT stages */
#define DTFFT_HIP_EXECUTOR_NONE 0
#define DTFFT_HIP_EXECUTOR_HIP  5
/* abstract_executor fft_type / fft_rank, src/dtfft_abstract_executor.F90:35-45 */
#define DTFFT_HIP_FFT_C2C 0
#define DTFFT_HIP_FFT_R2C 1
#define DTFFT_HIP_FFT_R2R 2
/* R2R kinds = FFTW REDFT/RODFT codes, include/dtfft_config.h.in:66-73 */
#define DTFFT_HIP_DCT_1 3
#define DTFFT_HIP_DCT_2 5
#define DTFFT_HIP_DCT_3 4
#define DTFFT_HIP_DCT_4 6
#define DTFFT_HIP_DST_1 7
#define DTFFT_HIP_DST_2 9
#define DTFFT_HIP_DST_3 8
#define DTFFT_HIP_DST_4 10
/* sign, src/include/_dtfft_private.h:27-28 */
#define DTFFT_HIP_FFT_FORWARD  (-1)
#define DTFFT_HIP_FFT_BACKWARD (+1)
/* execute / transpose types, include/dtfft_config.h.in:35-43 */
#define DTFFT_HIP_EXECUTE_FORWARD   11
#define DTFFT_HIP_EXECUTE_BACKWARD  12
#define DTFFT_HIP_TRANSPOSE_X_TO_Y  1
#define DTFFT_HIP_TRANSPOSE_Y_TO_X  (-1)
#define DTFFT_HIP_TRANSPOSE_Y_TO_Z  2
#define DTFFT_HIP_TRANSPOSE_Z_TO_Y  (-2)
#define DTFFT_HIP_TRANSPOSE_X_TO_Z  3
#define DTFFT_HIP_TRANSPOSE_Z_TO_X  (-3)

/* ---- library -------------------------------------------------------------------- */
int32_t dtfft_hip_get_version(void);
const char* dtfft_hip_get_error_string(int32_t error_code);
/* Detailed text of the last failure on the calling thread (hipRTC log, HIP error). */
const char* dtfft_hip_get_last_error_message(void);
/* Directory for compiled code objects and tuning records (default: $DTFFT_HIP_CACHE_DIR,
   else <library dir>/kcache).  NULL or "" disables the on-disk cache. */
int32_t dtfft_hip_set_cache_dir(const char* path);
/* Compile-only mode: when set (or with no gfx950 device), create calls compile and
   cache their kernels for gfx950 but allocate and launch nothing. */
int32_t dtfft_hip_set_compile_only(int32_t on);

/* ---- 1. kernel plugin ----------------------------------------------------------
 * Replaces kernel_device%create/execute/destroy.  `dims` are the local dims
 * (fastest first), `neighbor_data` is the 5 x n_peers column-major array of
 * (nxx, nyy, nzz, din, dout) of src/dtfft_abstract_kernel.F90:86, may be NULL for
 * permute kernels.  base_storage is the element size in bytes (4, 8 or 16).
 * Follows abstract_kernel%create (src/dtfft_abstract_kernel.F90:219-288): zero dims
 * make a dummy kernel, 2D BACKWARD kernels are remapped to FORWARD ones.
 * effort >= MEASURE times launch configurations on the device (autotune); a stored
 * tuning record is used otherwise.  force_effort != 0 with ESTIMATE takes the default
 * launch shape and ignores tuning records (src/dtfft_kernel_device.F90:326-335). */
int32_t dtfft_hip_kernel_create(int32_t kernel_type, int32_t ndims, const int32_t* dims,
                                int64_t base_storage, const int32_t* neighbor_data, int32_t n_peers,
                                int32_t effort, int32_t force_effort, void** handle);
/* Replaces kernel_device%execute (src/dtfft_kernel_device.F90:104-178).  neighbor is
 * 1-based; 0 runs every peer (PACK / UNPACK / PERMUTE_BACKWARD_END).  `stream` is a
 * hipStream_t (NULL = default stream); sync != 0 waits for completion. */
int32_t dtfft_hip_kernel_execute(void* handle, const void* in, void* out, void* stream,
                                 int32_t neighbor, int32_t sync);
int32_t dtfft_hip_kernel_destroy(void* handle);

/* ---- 2. FFT executor plugin ------------------------------------------------------
 * Replaces abstract_executor%create_private (src/dtfft_abstract_executor.F90:67-83)
 * with the same arguments: fft_sizes/inembed/onembed in FFTW (row-major) order,
 * unit element stride, distances idist/odist between consecutive transforms.
 * One plan serves both directions (is_inverse_copied in the reference).
 * r2r_kinds is required for FFT_R2R (dtFFT kind ids, include/dtfft_config.h.in:66-73), one per
 * dimension in fft_sizes order; FFTW semantics, unnormalised.  The forward sign applies the
 * kinds, the backward sign their inverses (src/dtfft_utils.F90:447-471).  An R2R stage runs
 * a complex FFT of length n, 2n (type 4), 2(n-1) (DCT-1) or 2(n+1) (DST-1).
 * Any length is accepted, as cuFFT and FFTW accept it: lengths whose primes are <= 13 run
 * the mixed-radix kernels, others Bluestein's chirp z-transform, odd R2C/C2R lengths a
 * direct odd-length kernel.  A length whose work line (n, or the Bluestein length >= 2n-1)
 * exceeds the 160 KiB of LDS of one workgroup returns DTFFT_HIP_ERROR_UNSUPPORTED_SIZE. */
int32_t dtfft_hip_fft_create(int8_t fft_rank, int8_t fft_type, int32_t precision, int32_t idist,
                             int32_t odist, int32_t how_many, const int32_t* fft_sizes,
                             const int32_t* inembed, const int32_t* onembed,
                             const int32_t* r2r_kinds, int32_t effort, void* stream, void** plan);
/* Replaces abstract_executor%execute_private: sign = FFT_FORWARD (-1) or FFT_BACKWARD (+1).
 * For R2C, forward is real->complex and backward complex->real (in != out required).
 * R2R may run in place. */
int32_t dtfft_hip_fft_execute(void* plan, const void* in, void* out, int8_t sign);
int32_t dtfft_hip_fft_destroy(void* plan);

/* Device memory (abstract_executor%mem_alloc/mem_free, dtfft_mem_alloc/free). */
int32_t dtfft_hip_mem_alloc(size_t bytes, void** ptr);
int32_t dtfft_hip_mem_free(void* ptr);

/* ---- 3. single-rank plans (include/dtfft.h) ---------------------------------------- */
typedef void* dtfft_hip_plan_t;

typedef struct {
    int32_t enable_z_slab;        /* 1: Z-slab path for 3D (2D FFT over X,Y, X->Z, FFT Z); default 0 */
    int32_t enable_fused;         /* 1: fuse each transpose into the following FFT pass; default 1 */
    int32_t enable_kernel_autotune; /* 1: time kernel launch configurations even at ESTIMATE */
    int32_t measure_warmup_iters; /* autotune warm-up launches (default 2) */
    int32_t measure_iters;        /* autotune timed launches (default 5) */
    void* stream;                 /* hipStream_t, NULL = default stream */
    int32_t lds_dma;              /* column-in FFT passes through LDS tiles filled by direct global->LDS
                                     loads: 0 never, 1 where the autotuner (or the default rule) picks
                                     them (default), 2 the double-buffer kernel wherever the stage shape
                                     allows, 3 the single-buffer kernel (one tile per block, two
                                     blocks per CU) likewise */
    int32_t twl;                  /* FFT kernels that keep the twiddle table in LDS (persistent blocks):
                                     0 never, 1 where the default rule picks them (default: FP32 rows
                                     of n >= 2048, FP32 column-in passes without LDS-DMA, the fused
                                     C2R, on stages of >= 2048 tiles), 2 every stage whose kernel has
                                     the form and whose LDS budget allows it */
    int32_t blocked;              /* 1 (default): a row FFT pass whose output the next pass reads as
                                     column tiles writes it chunk-blocked (each tile's lines as one
                                     contiguous block), so that the column pass streams it; the
                                     intermediate layout is private to the plan.  0: natural layout */
} dtfft_hip_config_t;

int32_t dtfft_hip_create_config(dtfft_hip_config_t* config);
int32_t dtfft_hip_set_config(const dtfft_hip_config_t* config);

/* dims are the global dims in reverse order (slowest first), as dtfft_create_plan_c2c. */
int32_t dtfft_hip_create_plan_c2c(int8_t ndims, const int32_t* dims, int32_t precision, int32_t effort,
                                  int32_t executor, dtfft_hip_plan_t* plan);
int32_t dtfft_hip_create_plan_r2c(int8_t ndims, const int32_t* dims, int32_t precision, int32_t effort,
                                  int32_t executor, dtfft_hip_plan_t* plan);
/* dtfft_create_plan_r2r (include/dtfft.h): kinds[ndims] in the same reversed order as dims;
 * real elements; the backward applies the inverse kinds.  executor NONE makes a
 * transpose-only plan of real elements (kinds may then be NULL). */
int32_t dtfft_hip_create_plan_r2r(int8_t ndims, const int32_t* dims, const int32_t* kinds, int32_t precision,
                                  int32_t effort, int32_t executor, dtfft_hip_plan_t* plan);
int32_t dtfft_hip_execute(dtfft_hip_plan_t plan, void* in, void* out, int32_t execute_type, void* aux);
int32_t dtfft_hip_transpose(dtfft_hip_plan_t plan, void* in, void* out, int32_t transpose_type, void* aux);
int32_t dtfft_hip_destroy(dtfft_hip_plan_t* plan);
/* Forward then backward of one plan, captured once as a hipGraph and replayed (no counterpart
   in the reference, which launches every kernel of execute FORWARD / BACKWARD itself):
   in -> out (forward), out -> in2 (backward, unnormalised), aux as in dtfft_hip_execute (NULL:
   the plan's own, allocated before the capture).  The graph is captured again when the
   pointers change.  Ordered on the plan's stream like two execute calls; the backward may
   overwrite out, as execute BACKWARD may. */
int32_t dtfft_hip_execute_cycle_graph(dtfft_hip_plan_t plan, void* in, void* out, void* aux, void* in2);
/* No counterpart in the reference: dtfft_hip_execute without an aux argument and without any
   allocation, for callers that cannot allocate (code under stream capture, tight loops,
   harnesses that forbid hipMalloc).  The aux buffer of the sequence is a fixed per-device
   array of DTFFT_HIP_NO_ALLOC_AUX_BYTES inside the library (each GPU the library runs on holds
   it until the process ends); a plan whose aux (dtfft_hip_get_aux_bytes) is larger returns
   DTFFT_HIP_ERROR_SCRATCH_TOO_SMALL before any launch.  All calls on one GPU share that
   array: calls that overlap on one GPU (other streams, other threads) race. */
#define DTFFT_HIP_NO_ALLOC_AUX_BYTES ((size_t)128 << 20)
int32_t dtfft_hip_execute_no_alloc(dtfft_hip_plan_t plan, void* in, void* out, int32_t execute_type);
/* starts/counts in reverse order, as dtfft_get_local_sizes; alloc_size in elements.  A fused
   3D R2C plan whose forward keeps its complex X-pencil padded to whole 16-line chunks
   asks for up to 1/16 more than the complex pencil (2 * ceil((nx/2+1)/16)*16 * ny * nz reals). */
int32_t dtfft_hip_get_local_sizes(dtfft_hip_plan_t plan, int32_t* in_starts, int32_t* in_counts,
                                  int32_t* out_starts, int32_t* out_counts, size_t* alloc_size);
int32_t dtfft_hip_get_alloc_size(dtfft_hip_plan_t plan, size_t* alloc_size);
int32_t dtfft_hip_get_element_size(dtfft_hip_plan_t plan, size_t* element_size);
int32_t dtfft_hip_get_alloc_bytes(dtfft_hip_plan_t plan, size_t* alloc_bytes);
int32_t dtfft_hip_get_aux_bytes(dtfft_hip_plan_t plan, size_t* aux_bytes);
int32_t dtfft_hip_get_z_slab_enabled(dtfft_hip_plan_t plan, int32_t* is_z_slab_enabled);
int32_t dtfft_hip_plan_mem_alloc(dtfft_hip_plan_t plan, size_t alloc_bytes, void** ptr);
int32_t dtfft_hip_plan_mem_free(dtfft_hip_plan_t plan, void* ptr);
int32_t dtfft_hip_get_stream(dtfft_hip_plan_t plan, void** stream);
int32_t dtfft_hip_report(dtfft_hip_plan_t plan);
/* Profiling entry point (no counterpart in the reference): runs only launch `stage`
   (0-based) of the execute sequence, with the buffers execute would use; stage < 0 runs
   nothing.  n_stages (may be NULL) receives the number of launches in the sequence. */
int32_t dtfft_hip_plan_execute_stage(dtfft_hip_plan_t plan, void* in, void* out, int32_t execute_type,
                                     void* aux, int32_t stage, int32_t* n_stages);

/* ---- receive side of a P > 1 transpose fused with the next FFT (no counterpart) -------
 * Replaces KERNEL_UNPACK over all peers (src/dtfft_reshape_handle_generic.F90:606-636)
 * followed by the FFT along the aligned axis, when the peers' chunks tile that axis in
 * equal segments (neighbor_data rows (seg, ny, nz, din0 + i*chunk, i*seg)).  dims are the
 * receive pencil's local dims (aligned axis first).  Returns INVALID_USAGE otherwise. */
int32_t dtfft_hip_unpack_fft_create(int32_t ndims, const int32_t* dims, const int32_t* neighbor_data,
                                    int32_t n_peers, int32_t precision, int32_t effort, void* stream,
                                    void** handle);
int32_t dtfft_hip_unpack_fft_execute(void* handle, const void* recv, void* out, int8_t sign);
int32_t dtfft_hip_unpack_fft_destroy(void* handle);

/* ---- batched cycle (no counterpart in the reference) ---------------------------------
 * n_grids independent 3D C2C grids (dims[3*g..3*g+2], reverse order) packed one after
 * another in one buffer; every stage of all grids runs in one persistent launch with
 * the transpose fused into the following FFT pass.  Offsets are in elements. */
typedef void* dtfft_hip_cycle_t;
int32_t dtfft_hip_cycle_create(int32_t n_grids, const int32_t* dims, int32_t precision, int32_t effort,
                               void* stream, dtfft_hip_cycle_t* cycle);
int32_t dtfft_hip_cycle_get_layout(dtfft_hip_cycle_t cycle, int64_t* offsets, size_t* alloc_size);
int32_t dtfft_hip_cycle_execute(dtfft_hip_cycle_t cycle, void* in, void* out, void* aux, int32_t execute_type);
/* Forward then backward, captured once as a hipGraph and replayed: in -> out (spectrum)
   -> in2 (unnormalised round trip). */
int32_t dtfft_hip_cycle_execute_graph(dtfft_hip_cycle_t cycle, void* in, void* out, void* aux, void* in2);
int32_t dtfft_hip_cycle_destroy(dtfft_hip_cycle_t* cycle);

/* ---- plumbing used by tests and benchmarks ----------------------------------------- */
int32_t dtfft_hip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int32_t dtfft_hip_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Asynchronous device-to-device copy on `stream` (the loopback that stands in for the
   exchange on one GPU). */
int32_t dtfft_hip_memcpy_d2d_async(void* dst, const void* src, size_t bytes, void* stream);
int32_t dtfft_hip_device_synchronize(void);
int32_t dtfft_hip_stream_create(void** stream);
int32_t dtfft_hip_stream_destroy(void* stream);
int32_t dtfft_hip_stream_synchronize(void* stream);
int32_t dtfft_hip_event_create(void** event);
int32_t dtfft_hip_event_destroy(void* event);
int32_t dtfft_hip_event_record(void* event, void* stream);
int32_t dtfft_hip_event_elapsed_ms(void* start, void* stop, float* ms);
/* Name of the gfx target of device 0 ("" if none). */
int32_t dtfft_hip_get_device_arch(char* buf, int32_t buflen);

#ifdef __cplusplus
}
#endif
#endif
