// R2R stages (DCT / DST types 1-4, FFTW REDFT / RODFT semantics): the real-to-real
// transforms dtFFT hands to FFTW (src/interfaces/fft/fftw/dtfft_executor_fftw_m.F90:
// 149-223) or VkFFT (src/interfaces/fft/vkfft/dtfft_executor_vkfft_m.F90:97-123).  A stage
// is r2r_pre (real line -> complex line of length M), the Stockham C2C transform of the
// work line, and r2r_post (complex -> n real outputs); device/r2r.hip lists the per-kind
// maps.  Contiguous lines run all three in one kernel through LDS; other layouts run three
// launches through a work buffer (lines x M complex) that belongs to the caller.
#include <map>
#include <mutex>

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

// Radix schedule of the half-length form: the power-of-two part in radix-4 passes (one radix 2
// if the exponent is odd), the odd primes of factorize after it.  With M / 4 threads per line
// every thread holds one radix-4 butterfly per pass and 8 real elements of the prefetch;
// factorize's radix-16 passes would leave 16 threads per line holding 32 (293 VGPRs at
// M = 256 against 117).
static bool half_schedule(int m, std::vector<int>& out) {
    std::vector<int> f;
    if (!factorize(m, f)) return false;
    int p2 = 0, rest = m;
    while (rest % 2 == 0) { rest /= 2; p2++; }
    out.assign((size_t)(p2 / 2), 4);
    if (p2 % 2) out.push_back(2);
    for (int r : f) if (r % 2 != 0) out.push_back(r);
    if (out.empty()) out.push_back(1);
    return out.size() <= 10;
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
    // types 3 (DCT-3, DST-3) run the backward transform
    sign = (kind == 4 || kind == 8) ? DTFFT_HIP_FFT_BACKWARD : DTFFT_HIP_FFT_FORWARD;
    // A work length with a prime factor above 13 runs the three launches, whose C2C stage
    // is then the Bluestein form of plan_fft_stage.
    std::vector<int> rad;
    const bool factors = factorize((int)m, rad);
    // One kernel for unit-stride, evenly spaced lines whose output overwrites nothing but the
    // line itself: a block reads its lines whole before writing them, so in place is safe.
    // Batches that continue the line spacing fold into one range of lines.
    const bool even = nb == 1 || (in_bs == nl * in_ls && out_bs == nl * out_ls);
    // Types 2 and 3 of even length run the fused kernel's half-length form: the n-point
    // transform of a real sequence (or of a Hermitian spectrum) as an n/2-point one.
    std::vector<int> hrad;
    const bool half_ok = (kind == 4 || kind == 5 || kind == 8 || kind == 9) && n % 2 == 0 && n >= 8 &&
                         half_schedule((int)(n / 2), hrad);
    const int64_t mf = half_ok ? n / 2 : m;  // length of the fused kernel's FFT
    if ((factors || half_ok) && in_es == 1 && out_es == 1 && in_ls == out_ls && in_bs == out_bs && even && m > 1) {
        fft_config c = default_fft_config((int)mf, dbl, false, false, false);
        if (half_ok) {
            // M / 4 threads per line (one butterfly of radix <= 4 per pass), 512-thread blocks.
            // Measured on MI355X at 512^3 FP32 (medians of interleaved rounds), DCT-2 / DCT-3
            // stage: 269 / 239 us with 8 lines of 64 threads, 297 / 254 us with 4 lines,
            // 306 / 232 us with radix-8 passes at 32 threads per line, 328 / 314 us for the
            // full-length form.
            c.radices = hrad;
            const int rmax = *std::max_element(hrad.begin(), hrad.end());
            c.tpl = (int)std::max<int64_t>(1, mf / std::max(4, rmax));
            c.lpb = std::max(1, 512 / c.tpl);
        } else {
            // one butterfly of the largest radix per thread and pass
            const int rmax = *std::max_element(c.radices.begin(), c.radices.end());
            if ((mf / rmax) * c.lpb <= 1024) c.tpl = (int)(mf / rmax);
        }
        const size_t ec = dbl ? 16 : 8, er = dbl ? 8 : 4;
        // work lines, real lines, FFT twiddles (mf - 1 of 16 bytes), table (2n complex), and
        // the half-length form's untangling table (n/2 + 1 of 16 bytes)
        const size_t lds = (size_t)c.lpb * (size_t)(c.n + c.pad) * ec + (size_t)c.lpb * n * er +
                           (size_t)std::max<int64_t>(1, mf - 1) * 16 + (size_t)(2 * n) * ec +
                           (half_ok ? (size_t)(n / 2 + 1) * 16 : 16);
        if (lds <= 160 * 1024) {
            d = r2r_desc_h{nl * nb, 1, in_ls, nl * nb * in_ls, 1, out_ls, nl * nb * out_ls, 1, n, m, m};
            std::string e = "fft_r2r_kernel<" + t + ", dtfft::sched<" + std::to_string(mf);
            for (int r : c.radices) e += ", " + std::to_string(r);
            e += ">, " + std::to_string(c.tpl) + ", " + std::to_string(c.lpb) + ", " + std::to_string(c.pad) + ", " +
                 std::to_string(kind) + ((nl * nb) % c.lpb == 0 ? ", true" : ", false") +
                 (half_ok ? ", true>" : ">");
            TRY(rtc_get("r2r", e, &fk));
            expr = e;
            lpb = c.lpb;
            TRY(get_twiddles((int)mf, c.radices, dbl, &ftw));
            TRY(get_r2r_table(n, dbl, &t4));
            if (half_ok) TRY(get_rtwiddles((int)n, dbl, &rtw));
            half = half_ok;
            const int64_t tiles = (nl * nb + c.lpb - 1) / c.lpb;
            fblock = (unsigned)(c.tpl * c.lpb);
            // persistent: as many blocks as are resident
            int per_cu = 1;
            if (!compile_only() && fk && fk->fn &&
                hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fk->fn, (int)fblock, 0) != hipSuccess)
                per_cu = 1;
            per_cu = std::max(1, per_cu);
            fgrid = (unsigned)std::max<int64_t>(
                1, std::min<int64_t>(tiles, (int64_t)per_cu * std::max(1, device().num_cus)));
            fused = true;
            return DTFFT_HIP_SUCCESS;
        }
    }
    TRY(rtc_get("r2r", "r2r_pre<" + t + ", " + std::to_string(kind) + ">", &pre));
    TRY(rtc_get("r2r", "r2r_post<" + t + ", " + std::to_string(kind) + ">", &post));
    return plan_fft_stage((int)m, dbl, nl * nb, 1, m, 0, 1, m, 0, 1, true, effort, stream, &fft);
}

int32_t r2r_stage::blocked_pair(const r2r_stage& rows, const r2r_stage& cols, r2r_stage* rows_blk,
                                r2r_stage* cols_blk, bool* ok) {
    *ok = false;
    if (!global_config().blocked) return DTFFT_HIP_SUCCESS;
    // rows: the fused half-length row form over evenly spaced contiguous rows (batches folded)
    if (!rows.fused || !rows.half || rows.d.nb != 1 || rows.d.in_es != 1 || rows.d.out_es != 1 ||
        rows.d.in_ls != rows.n || rows.d.out_ls != rows.n || rows.expr.empty() || rows.expr.back() != '>')
        return DTFFT_HIP_SUCCESS;
    // cols: the fused column form reading the rows' output across lines: n lines side by side,
    // its FFT axis (nlc rows) n elements apart, batches of whole planes
    const int64_t n = rows.n, nlc = cols.n, B = cols.lpb;
    const r2r_desc_h& cd = cols.d;
    if (!cols.fused || !cols.half || cols.dbl != rows.dbl || B <= 0 || n % B != 0 || cd.nl != n || cd.in_ls != 1 ||
        cd.in_es != n || cd.in_ts != 0 || (cd.nb > 1 && cd.in_bs != n * nlc) || rows.d.nl != nlc * cd.nb)
        return DTFFT_HIP_SUCCESS;
    if (rows.lpb <= 0 || nlc % rows.lpb != 0) return DTFFT_HIP_SUCCESS;  // row tiles within one batch
    r2r_stage r = rows;
    r.expr = rows.expr.substr(0, rows.expr.size() - 1) + ", false, false, " + std::to_string(B) + ">";
    TRY(rtc_get("r2r", r.expr, &r.fk));
    r.d.nl = nlc;
    r.d.nb = cd.nb;
    r.d.in_bs = n * nlc;  // rows stay evenly spaced: line (l, b) at (l + nlc b) n
    r.d.out_ls = B;
    r.d.out_bs = n * nlc;
    r.d.out_seg = nlc * B;
    int per_cu = 1;
    if (!compile_only() && r.fk && r.fk->fn &&
        hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, r.fk->fn, (int)r.fblock, 0) != hipSuccess)
        per_cu = 1;
    per_cu = std::max(1, per_cu);
    const int64_t tiles = (nlc * cd.nb + r.lpb - 1) / r.lpb;
    r.fgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)per_cu * std::max(1, device().num_cus)));
    r2r_stage c = cols;
    c.d.in_es = B;
    c.d.in_ts = nlc * B;
    c.d.in_bs = n * nlc;
    *rows_blk = r;
    *cols_blk = c;
    *ok = true;
    return DTFFT_HIP_SUCCESS;
}

size_t r2r_stage::work_bytes() const { return fused ? 0 : (size_t)(d.nl * d.nb * d.wdist) * (dbl ? 16 : 8); }

int32_t r2r_stage::init_cols(int kind_, int64_t n_, bool dbl_, int64_t nl, int64_t nb, int64_t in_bs, int64_t in_es,
                             int64_t out_ls, int64_t out_bs, int64_t out_es, bool out_cols, bool* ok) {
    *ok = false;
    if (!r2r_kind_valid(kind_) || n_ < 8 || n_ % 2 != 0) return DTFFT_HIP_SUCCESS;
    if (kind_ != 4 && kind_ != 5 && kind_ != 8 && kind_ != 9) return DTFFT_HIP_SUCCESS;
    std::vector<int> hrad;
    if (!half_schedule((int)(n_ / 2), hrad)) return DTFFT_HIP_SUCCESS;
    const size_t ec = dbl_ ? 16 : 8, er = dbl_ ? 8 : 4;
    const int lpb = (int)(128 / er);  // a tile row is one 128-byte piece of LPB reals
    if (nl % lpb != 0) return DTFFT_HIP_SUCCESS;
    const int64_t mf = n_ / 2;
    fft_config c = default_fft_config((int)mf, dbl_, false, false, false);
    c.radices = hrad;
    c.pad = 1;
    c.lpb = lpb;
    const int rmax = *std::max_element(hrad.begin(), hrad.end());
    c.tpl = (int)std::max<int64_t>(1, std::min<int64_t>(mf / std::max(4, rmax), 1024 / lpb));
    const size_t lds = (size_t)lpb * (size_t)((mf + c.pad)) * ec + (size_t)lpb * n_ * er +
                       (size_t)std::max<int64_t>(1, mf - 1) * 16 + (size_t)(2 * n_) * ec + (size_t)(mf + 1) * 16;
    if (lds > 160 * 1024) return DTFFT_HIP_SUCCESS;
    kind = kind_;
    n = n_;
    dbl = dbl_;
    m = work_length(kind, n);
    sign = (kind == 4 || kind == 8) ? DTFFT_HIP_FFT_BACKWARD : DTFFT_HIP_FFT_FORWARD;
    d = r2r_desc_h{nl, nb, 1, in_bs, in_es, out_ls, out_bs, out_es, n, m, m};
    std::string e = std::string("fft_r2r_kernel<") + (dbl ? "double" : "float") + ", dtfft::sched<" +
                    std::to_string(mf);
    for (int r : c.radices) e += ", " + std::to_string(r);
    e += ">, " + std::to_string(c.tpl) + ", " + std::to_string(c.lpb) + ", " + std::to_string(c.pad) + ", " +
         std::to_string(kind) + ", true, true, true" + (out_cols ? ", true>" : ", false>");
    TRY(rtc_get("r2r", e, &fk));
    expr = e;
    this->lpb = c.lpb;
    TRY(get_twiddles((int)mf, c.radices, dbl, &ftw));
    TRY(get_r2r_table(n, dbl, &t4));
    TRY(get_rtwiddles((int)n, dbl, &rtw));
    half = true;
    const int64_t tiles = nl * nb / lpb;
    fblock = (unsigned)(c.tpl * c.lpb);
    int per_cu = 1;
    if (!compile_only() && fk && fk->fn &&
        hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fk->fn, (int)fblock, 0) != hipSuccess)
        per_cu = 1;
    per_cu = std::max(1, per_cu);
    fgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)per_cu * std::max(1, device().num_cus)));
    fused = true;
    *ok = true;
    return DTFFT_HIP_SUCCESS;
}

// e^{-i pi j / 4n} for j < 2n, generated on the device in FP64; one table per (n, precision)
// for the life of the process.
static std::mutex g_t4_mu;
static std::map<std::pair<int64_t, bool>, std::shared_ptr<twiddles>> g_t4;

int32_t get_r2r_table(int64_t n, bool dbl, std::shared_ptr<twiddles>* out) {
    std::lock_guard<std::mutex> lk(g_t4_mu);
    auto it = g_t4.find({n, dbl});
    if (it != g_t4.end()) { *out = it->second; return DTFFT_HIP_SUCCESS; }
    auto t = std::make_shared<twiddles>();
    const rtc_kernel* k = nullptr;
    TRY(rtc_get("r2r", std::string("r2r_table<") + (dbl ? "double" : "float") + ">", &k));
    t->bytes = 2 * n * (dbl ? 16 : 8);
    if (!compile_only()) {
        HIP_TRY(hipMalloc(&t->ptr, (size_t)t->bytes));
        void* ptr = t->ptr;
        int64_t nn = n;
        void* args[] = {&ptr, &nn};
        TRY(launch(k, (unsigned)((2 * n + 255) / 256), 256, nullptr, args));
        HIP_TRY(hipDeviceSynchronize());
    }
    g_t4[{n, dbl}] = t;
    *out = t;
    return DTFFT_HIP_SUCCESS;
}

static unsigned grid_for(int64_t items) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, 1 << 20));
}

int32_t r2r_stage::run(const void* in, void* out, void* work, hipStream_t s) const {
    if (d.nl * d.nb == 0) return DTFFT_HIP_SUCCESS;
    if (fused) {
        r2r_desc_h dd = d;
        void* twp = ftw ? ftw->ptr : nullptr;
        void* tp = t4 ? t4->ptr : nullptr;
        void* rp = rtw ? rtw->ptr : nullptr;
        void* args[] = {(void*)&in, &out, &twp, &tp, &rp, &dd};
        return launch(fk, fgrid, fblock, s, args);
    }
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
