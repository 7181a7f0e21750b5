// Receive side of a P > 1 transpose fused with the FFT that follows it.
//
// In the reference the receive side is KERNEL_UNPACK (or _PIPELINED) over all peers
// (src/dtfft_reshape_handle_generic.F90:606-636, kernel _dtfft_kernel_host_routines.inc:
// 575-636) followed by a separate FFT along the now-aligned axis
// (src/dtfft_plan.F90:1088-1100).  When the peers' chunks tile the aligned axis in equal
// segments -- chunk i holds elements [i*seg, (i+1)*seg) of every line, stored as
// [seg, ny, nz] at din_i = i * chunk -- both are one pass here: the FFT kernel reads
// each line segment by segment straight from the receive buffer and writes the aligned
// pencil (no counterpart entry point in the reference).
#include "internal.h"

namespace dtfft_hip {

struct unpack_fft_t {
    fft_stage st;
    hipStream_t stream = nullptr;
    int64_t din0 = 0;
};

}  // namespace dtfft_hip

using namespace dtfft_hip;

extern "C" {

int32_t dtfft_hip_unpack_fft_create(int32_t ndims, const int32_t* dims, const int32_t* neighbor_data, int32_t n_peers,
                                    int32_t precision, int32_t effort, void* stream, void** handle) {
    if (!handle || !dims || !neighbor_data || n_peers <= 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *handle = nullptr;
    if (ndims != 2 && ndims != 3) return DTFFT_HIP_ERROR_INVALID_N_DIMENSIONS;
    if (precision != DTFFT_HIP_SINGLE && precision != DTFFT_HIP_DOUBLE) return DTFFT_HIP_ERROR_INVALID_PRECISION;
    const int64_t nx = dims[0], ny = dims[1], nz = ndims == 3 ? dims[2] : 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE;
    if (nx * ny * nz > INT32_MAX) return DTFFT_HIP_ERROR_INDEX_OVERFLOW;
    // the chunks must tile the aligned axis in equal segments at a constant spacing
    const int64_t seg = neighbor_data[0];
    const int64_t chunk = n_peers > 1 ? (int64_t)neighbor_data[5 + 3] - neighbor_data[3] : seg * ny * nz;
    for (int i = 0; i < n_peers; i++) {
        const int32_t* r = neighbor_data + 5 * i;
        const int64_t rz = ndims == 3 ? r[2] : 1;
        if (r[0] != seg || r[1] != ny || rz != nz || r[4] != i * seg || r[3] != neighbor_data[3] + i * chunk) {
            dtfft_hip::set_last_error("unpack+FFT needs equal chunks tiling the aligned axis");
            return DTFFT_HIP_ERROR_INVALID_USAGE;
        }
    }
    if (seg * n_peers != nx || chunk < seg * ny * nz) {
        dtfft_hip::set_last_error("unpack+FFT: chunks do not cover the aligned axis");
        return DTFFT_HIP_ERROR_INVALID_USAGE;
    }
    auto h = std::make_unique<unpack_fft_t>();
    h->stream = (hipStream_t)stream;
    const bool dbl = precision == DTFFT_HIP_DOUBLE;
    std::vector<int> rad;
    if (!factorize((int)nx, rad)) return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
    fft_config c = default_fft_config((int)nx, dbl, false, false, false);
    c.seg = (int)seg;
    // line (y, z) = l: input base l * seg (inside chunk 0) + din_0, output l * nx
    lines_desc d = make_lines(ny * nz, 1, seg, 0, 1, nx, 0, 1, c.lpb);
    d.in_seg_stride = chunk;
    // long FP32 lines keep the twiddle table in LDS (persistent blocks), as plain rows do
    apply_twl(c, d, global_config().twl);
    TRY(h->st.init(c, d));
    h->din0 = neighbor_data[3];  // din of peer 0, applied as a pointer offset
    *handle = h.release();
    (void)effort;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_unpack_fft_execute(void* handle, const void* recv, void* out, int8_t sign) {
    auto* h = (unpack_fft_t*)handle;
    if (!h) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (sign != DTFFT_HIP_FFT_FORWARD && sign != DTFFT_HIP_FFT_BACKWARD) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (recv == out) return DTFFT_HIP_ERROR_INPLACE_TRANSPOSE;
    const int64_t es = h->st.cfg.dbl ? 16 : 8;
    return h->st.run((const char*)recv + h->din0 * es, out, sign, h->stream);
}

int32_t dtfft_hip_unpack_fft_destroy(void* handle) {
    delete (unpack_fft_t*)handle;
    return DTFFT_HIP_SUCCESS;
}

}  // extern "C"
