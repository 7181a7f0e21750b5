// R2R stages (DCT / DST types 1-4, FFTW REDFT / RODFT semantics): the real-to-real
// transforms dtFFT hands to FFTW (src/interfaces/fft/fftw/dtfft_executor_fftw_m.F90:
// 149-223) or VkFFT (src/interfaces/fft/vkfft/dtfft_executor_vkfft_m.F90:97-123).  A stage
// is three launches: r2r_pre (real line -> complex line of length M), the Stockham C2C
// stage over the work lines, r2r_post (complex -> n real outputs); device/r2r.hip lists
// the per-kind maps.  The work buffer (lines x M complex) belongs to the caller.
#include "internal.h"

namespace dtfft_hip {

bool r2r_kind_valid(int kind) { return kind >= 3 && kind <= 10; }

int r2r_inverse_kind(int kind) {  // src/dtfft_utils.F90:447-471
    switch (kind) {
        case 5: return 4;
        case 4: return 5;
        case 9: return 8;
        case 8: return 9;
        default: return kind;  // DCT-1, DCT-4, DST-1, DST-4 are their own inverses
    }
}

static int64_t work_length(int kind, int64_t n) {
    if (kind == 3) return 2 * (n - 1);
    if (kind == 7) return 2 * (n + 1);
    if (kind == 6 || kind == 10) return 2 * n;
    return n;
}

int32_t r2r_stage::init(int kind_, int64_t n_, bool dbl_, int64_t nl, int64_t nb, int64_t in_ls, int64_t in_bs,
                        int64_t in_es, int64_t out_ls, int64_t out_bs, int64_t out_es, int effort,
                        hipStream_t stream) {
    if (!r2r_kind_valid(kind_)) {
        set_last_error("R2R kind " + std::to_string(kind_) + " is not one of DCT/DST 1-4 (3..10)");
        return DTFFT_HIP_ERROR_INVALID_R2R_KINDS;
    }
    if (n_ < 1 || (kind_ == 3 && n_ < 2)) {
        set_last_error("R2R length " + std::to_string(n_) + " too short for kind " + std::to_string(kind_));
        return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
    }
    kind = kind_;
    n = n_;
    dbl = dbl_;
    m = work_length(kind, n);
    d = r2r_desc_h{nl, nb, in_ls, in_bs, in_es, out_ls, out_bs, out_es, n, m, m};
    const std::string t = dbl ? "double" : "float";
    TRY(rtc_get("r2r", "r2r_pre<" + t + ", " + std::to_string(kind) + ">", &pre));
    TRY(rtc_get("r2r", "r2r_post<" + t + ", " + std::to_string(kind) + ">", &post));
    // types 3 (DCT-3, DST-3) run the backward transform
    sign = (kind == 4 || kind == 8) ? DTFFT_HIP_FFT_BACKWARD : DTFFT_HIP_FFT_FORWARD;
    return plan_fft_stage((int)m, dbl, nl * nb, 1, m, 0, 1, m, 0, 1, true, effort, stream, &fft);
}

size_t r2r_stage::work_bytes() const { return (size_t)(d.nl * d.nb * d.wdist) * (dbl ? 16 : 8); }

static unsigned grid_for(int64_t items) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, 1 << 20));
}

int32_t r2r_stage::run(const void* in, void* out, void* work, hipStream_t s) const {
    if (d.nl * d.nb == 0) return DTFFT_HIP_SUCCESS;
    if (!work) {
        set_last_error("R2R stage without a work buffer");
        return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    }
    r2r_desc_h dd = d;
    {
        void* args[] = {(void*)&in, &work, &dd};
        TRY(launch(pre, grid_for(d.nl * d.nb * d.m), 256, s, args));
    }
    TRY(fft.run(work, work, sign, s));
    void* args[] = {&work, &out, &dd};
    return launch(post, grid_for(d.nl * d.nb * d.n), 256, s, args);
}

}  // namespace dtfft_hip
