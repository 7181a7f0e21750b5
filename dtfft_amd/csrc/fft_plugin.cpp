// FFT executor plugin: batched 1D / 2D C2C (and R2C / C2R) stages on gfx950.
//
// Drop-in for the cuFFT / VkFFT executors behind abstract_executor
// (src/dtfft_abstract_executor.F90:47-112; cuFFT: src/interfaces/fft/cufft/
// dtfft_executor_cufft_m.F90:52-125).  create receives exactly what
// abstract_executor%create computes from the pencils (:115-216): fft_rank, the sizes
// in FFTW row-major order, inembed/onembed, idist/odist and how_many, unit stride.
//
// Each 1D stage is one launch of the Stockham kernel in device/fft.hip, compiled by
// hipRTC for its length.  Radix schedule, threads per line, lines per block and LDS
// padding come from a heuristic or, at effort >= MEASURE, from timing the candidates
// on the device (the project's replacement for dtfft_nvrtc_block_optimizer).
#include <algorithm>
#include <array>
#include <cmath>
#include <functional>

#include "internal.h"

namespace dtfft_hip {

// ---- schedules ----------------------------------------------------------------------
bool factorize(int n, std::vector<int>& radices) {
    radices.clear();
    if (n < 1) return false;
    if (n == 1) { radices.push_back(1); return true; }
    int a = 0, m = n;
    while (m % 2 == 0) { m /= 2; a++; }
    // power-of-two part: fewest passes with radices 16/8/4/2
    if (a > 0) {
        int passes = (a + 3) / 4;
        std::vector<int> e(passes, a / passes);
        for (int i = 0; i < a % passes; i++) e[i]++;
        for (int x : e) radices.push_back(1 << x);
    }
    const int primes[] = {3, 5, 7, 11, 13};
    for (int p : primes)
        while (m % p == 0) { radices.push_back(p); m /= p; }
    if (m != 1) return false;
    if (radices.size() > 10) return false;  // dtfft::sched holds up to 10 passes
    return true;
}

std::string fft_config::expr() const {
    if (any) {
        std::string s = std::string("fft_any_kernel<") + (dbl ? "double" : "float") + ", dtfft::sched<" +
                        std::to_string(n);
        for (int r : radices) s += ", " + std::to_string(r);
        return s + ">, " + std::to_string(tpl) + ", " + std::to_string(lpb) + ", " + std::to_string(pad) + ", " +
               std::to_string(kind) + (chirp ? ", true>" : ", false>");
    }
    if (dma && kind == 2) {
        std::string s = std::string(dma == 1 ? "fft_c2r_dma2_kernel<" : "fft_c2r_dma_kernel<") +
                        (dbl ? "double" : "float") + ", dtfft::sched<" + std::to_string(n);
        for (int r : radices) s += ", " + std::to_string(r);
        s += ">, " + std::to_string(tpl) + ", " + std::to_string(lpb) + ", " + std::to_string(pad);
        if (dma == 1) return s + ">";
        return s + (twl ? ", true>" : ", false>");
    }
    if (dma) {
        std::string s = std::string(dma == 2 ? "fft_cols_dma1_kernel<" : "fft_cols_dma_kernel<") +
                        (dbl ? "double" : "float") + ", dtfft::sched<" + std::to_string(n);
        for (int r : radices) s += ", " + std::to_string(r);
        return s + ">, " + std::to_string(tpl) + ", " + std::to_string(lpb) + ", " + std::to_string(pad) +
               (out_cols ? ", true>" : ">");
    }
    std::string s = kind == 0 ? "fft_c2c_kernel<" : (kind == 1 ? "fft_r2c_kernel<" : "fft_c2r_kernel<");
    s += dbl ? "double" : "float";
    s += ", dtfft::sched<" + std::to_string(n);
    for (int r : radices) s += ", " + std::to_string(r);
    s += ">, " + std::to_string(tpl) + ", " + std::to_string(lpb) + ", ";
    if (kind == 0) s += std::string(in_cols ? "true" : "false") + ", " + (out_cols ? "true" : "false") + ", ";
    s += std::to_string(pad);
    if (kind == 0 && (seg > 0 || twl || oseg > 0)) s += ", " + std::to_string(seg);
    if (kind == 0 && (twl || oseg > 0)) s += twl ? ", true" : ", false";
    if (kind == 0 && oseg > 0) s += ", " + std::to_string(oseg);
    if (kind == 1 && oseg > 0) s += ", " + std::to_string(oseg);
    if (kind == 2 && (in_cols || twl)) s += in_cols ? ", true" : ", false";
    if (kind == 2 && twl) s += ", true";
    s += ">";
    return s;
}

static int elem_bytes(bool dbl) { return dbl ? 16 : 8; }
static const size_t kMaxLds = 160 * 1024;

// LDS bytes of fft_cols_dma_kernel: two tile buffers (load layout of 1152-byte groups of 8
// rows, or the line-major staging layout, whichever is larger) and the twiddle table.
static size_t lds_dma(const fft_config& c) {
    const size_t es = (size_t)elem_bytes(c.dbl);
    if (c.kind == 2) {
        // fft_c2r_dma(2)_kernel: buffers of the n + 1 input rows (the last group unpadded) or the
        // staging layout; dma 1 (double buffer): both twiddle tables in LDS (the FFT table and
        // the untangling half of n / 2 + 1 entries); dma 2: the FFT table when twl is set
        const size_t groups = (size_t)((c.n + 1 + 7) / 8);
        const size_t load = (groups - 1) * 1152 + ((size_t)(c.n + 1) - 8 * (groups - 1)) * 128;
        const size_t stage = (size_t)c.lpb * (size_t)((c.n + c.pad) | 1) * es;
        size_t tw = 0, L = 1;
        for (int r : c.radices) { tw += L * (size_t)(r - 1); L *= (size_t)r; }
        const size_t twb = std::max<size_t>(1, tw) * 16 + 16;
        if (c.dma == 1) return 2 * std::max(load, stage) + twb + (size_t)(c.n / 2 + 1) * 16 + 16;
        return std::max(load, stage) + (c.twl ? twb : 0);
    }
    const size_t load = (size_t)(c.n / 8) * 1152;
    const size_t stage = (size_t)c.lpb * (size_t)((c.n + c.pad) | 1) * es;
    size_t tw = 0, L = 1;
    for (int r : c.radices) { tw += L * (size_t)(r - 1); L *= (size_t)r; }
    return (c.dma == 2 ? 1 : 2) * std::max(load, stage) + std::max<size_t>(1, tw) * 16;
}

// Whether a fused column-in C2R stage can use fft_c2r_dma_kernel.
static bool c2r_dma_ok(const fft_config& c, const lines_desc& d) {
    const int64_t es = elem_bytes(c.dbl);
    return c.kind == 2 && (c.dma == 1 || c.dma == 2) && c.n > 1 && c.n % 8 == 0 && c.lpb * es == 128 && (c.tpl * c.lpb) % 64 == 0 &&
           c.tpl * c.lpb <= 1024 && d.in_ls == 1 && d.out_ns == 1 && d.nl % c.lpb == 0 && (d.in_ns * es) % 16 == 0 &&
           (d.in_bs * es) % 16 == 0 && lds_dma(c) <= kMaxLds;
}

bool dma_ok(const fft_config& c, const lines_desc& d) {
    if (c.kind == 2) return c2r_dma_ok(c, d);
    const int64_t es = elem_bytes(c.dbl);
    const bool out_ok = c.out_cols ? (d.out_ls == 1 && (d.out_ns * es) % 16 == 0 && (d.out_bs * es) % 16 == 0)
                                   : d.out_ns == 1;
    return c.kind == 0 && c.in_cols && c.seg == 0 && c.n > 1 && c.n % 8 == 0 && c.lpb * es == 128 &&
           (c.tpl * c.lpb) % 64 == 0 && c.tpl * c.lpb <= 1024 && d.in_ls == 1 && out_ok && d.nl % c.lpb == 0 &&
           (d.in_ns * es) % 16 == 0 && (d.in_bs * es) % 16 == 0 && lds_dma(c) <= kMaxLds;
}

// LDS line stride of fft_tile (device/fft.hip): n + pad, or with pad 2 / 3 one spare slot per
// 16 elements plus (pad & 1), made odd.
static size_t lds_line(int n, int pad) {
    return pad >= 2 ? (size_t)((n + (n >> 4) + (pad & 1)) | 1) : (size_t)(n + pad);
}

// Entries of the schedule's twiddle table: sum over passes of L_p * (R_p - 1).
static size_t tw_entries(const fft_config& c) {
    size_t tw = 0, L = 1;
    for (int r : c.radices) { tw += L * (size_t)(r - 1); L *= (size_t)r; }
    return tw;
}

size_t fft_lds_bytes(const fft_config& c) {
    if (c.dma) return lds_dma(c);
    if (c.any) return (size_t)c.lpb * lds_line(c.n, c.pad) * elem_bytes(c.dbl);
    size_t b = (size_t)c.lpb * lds_line(c.n, c.pad) * elem_bytes(c.dbl);
    if (c.twl) {
        // the tables are separate __shared__ arrays of 16-byte entries (FP32 hi+lo pairs or
        // FP64 complex); 16 bytes of slack each for their alignment
        b += std::max<size_t>(1, tw_entries(c)) * 16 + 16;
        if (c.kind == 2) b += (size_t)(c.n + 1) * 16 + 16;  // C2R untangling table, M + 1 entries
    }
    return b;
}

static bool valid(const fft_config& c) {
    const int threads = c.tpl * c.lpb;
    if (threads < 1 || threads > 1024) return false;
    if (fft_lds_bytes(c) > kMaxLds) return false;
    return true;
}

void apply_twl(fft_config& c, const lines_desc& d, int mode) {
    c.twl = 0;
    // the LDS-DMA column kernels keep their twiddles in LDS anyway; R2C rows have no such form
    if (mode == 0 || c.dma || c.any || c.kind == 1 || c.n < 2) return;
    const int64_t ntiles = ((d.nl + c.lpb - 1) / c.lpb) * d.nb;
    bool want = mode >= 2;
    if (mode == 1) {
        // Default rule, measured on MI355X.  Long FP32 rows (n >= 2048, plain or fused
        // unpack+FFT): the table (n - 1 entries of 16 bytes) otherwise shares the L1 cache with
        // the tile; the 2048-point rows of a 2048 x 512 x 512 pencil went from 2028 to 1702 us.
        // At n = 512 the one-tile-per-block launch stays faster (384 against 412 us).  FP32
        // column-in passes without LDS-DMA (e.g. 513 lines per row, the complex X pencil of an
        // R2C plan) run one 1024-thread block per CU anyway, so persistence costs no occupancy:
        // 1024 x 512 x 256 R2C's X->Y pass went from 325 to 276 us; the fused Z->X C2R of the
        // same plan from 453 to 386 us.  FP64 measured no gain.
        const bool rows = c.kind == 0 && !c.in_cols && !c.out_cols;
        const bool cols = (c.kind == 0 && c.in_cols && c.seg == 0) || (c.kind == 2 && c.in_cols);
        want = !c.dbl && ntiles >= 2048 && ((rows && c.n >= 2048) || cols);
    }
    if (!want) return;
    c.twl = 1;
    if (!valid(c)) c.twl = 0;  // tile + tables over the LDS budget: global twiddles
}

fft_config default_fft_config(int n, bool dbl, bool in_cols, bool out_cols, bool inplace) {
    (void)inplace;
    fft_config c;
    c.n = n; c.dbl = dbl; c.in_cols = in_cols; c.out_cols = out_cols;
    factorize(n, c.radices);
    int rmax = *std::max_element(c.radices.begin(), c.radices.end());
    c.tpl = std::max(1, n / rmax);
    if (in_cols || out_cols) {
        // Columns: 16 lines per tile (128-byte rows in FP32).  FP64 at n >= 256 takes 16
        // lines too (256-byte rows) with two butterflies per thread, which measured faster
        // on MI355X (512^3: 825 us per pass against 845 with 8 lines).
        c.pad = 1;
        c.lpb = dbl ? 8 : 16;
        if (dbl && n >= 256) {
            c.lpb = 16;
            c.tpl = std::max(1, c.tpl / 2);
        }
        while (c.tpl * c.lpb > 1024 && c.tpl > 1) c.tpl = (c.tpl + 1) / 2;
        while (!valid(c) && c.lpb > 1) c.lpb /= 2;
    } else {
        // Rows: small blocks (about 128 threads in FP32, 64 in FP64, 256 once a line takes 64
        // threads or more), several resident per CU so that one block's loads overlap
        // another's passes; at n = 512..1024 (FP32) and 512 (FP64) two butterflies per thread
        // and pass measured faster on MI355X (512 FP32: 5.6 TB/s against 5.2 with one; 512
        // FP64: 731 us per pass with 2 lines per block against 763 with 4).
        c.pad = 0;
        if ((!dbl && n >= 512 && n <= 1024) || (dbl && n == 512)) c.tpl = std::max(1, c.tpl / 2);
        const int target = c.tpl >= 64 ? 256 : (dbl ? 64 : 128);
        c.lpb = 1;
        while (c.lpb * c.tpl < target) c.lpb *= 2;
        while ((size_t)c.lpb * lds_line(n, c.pad) * elem_bytes(dbl) > 64 * 1024 && c.lpb > 1) c.lpb /= 2;
        while (c.tpl * c.lpb > 1024 && c.tpl > 1) c.tpl = (c.tpl + 1) / 2;
    }
    return c;
}

// Radix schedules of n for the autotuner, the default (factorize) first: the power-of-two
// part split into the fewest passes of radix <= 16 or one more, in descending and ascending
// order, with the odd primes after (default) or before it.  At most `max_count`.
std::vector<std::vector<int>> radix_schedules(int n, int max_count) {
    std::vector<std::vector<int>> out;
    std::vector<int> def;
    if (!factorize(n, def)) return out;
    out.push_back(def);
    int a = 0, m = n;
    while (m % 2 == 0) { m /= 2; a++; }
    std::vector<int> odd;
    for (int r : def)
        if (r % 2 != 0 && r != 1) odd.push_back(r);
    std::vector<std::vector<int>> pow2;  // exponent lists, each in [1, 4]
    if (a == 0) pow2.push_back({});
    const int kmin = (a + 3) / 4;
    for (int k = kmin; a > 0 && k <= std::min(a, kmin + 1); k++) {
        // multisets of k exponents in [1, 4] summing to a, descending
        std::vector<int> e(k, 1);
        std::function<void(int, int, int)> rec = [&](int i, int left, int hi) {
            if (i == k) { if (left == 0) pow2.push_back(e); return; }
            for (int x = std::min(hi, 4); x >= 1; x--) {
                if (x > left - (k - i - 1)) continue;
                e[i] = x;
                rec(i + 1, left - x, x);
            }
        };
        rec(0, a, 4);
    }
    auto push = [&](const std::vector<int>& r) {
        if ((int)out.size() >= max_count || r.empty() || r.size() > 10) return;
        for (auto& o : out) if (o == r) return;
        out.push_back(r);
    };
    for (auto& ex : pow2) {
        std::vector<int> p2;
        for (int x : ex) p2.push_back(1 << x);
        std::vector<int> asc(p2.rbegin(), p2.rend());
        for (const auto* seq : {&p2, &asc}) {
            std::vector<int> r = *seq;
            r.insert(r.end(), odd.begin(), odd.end());
            push(r);
            std::vector<int> f = odd;
            f.insert(f.end(), seq->begin(), seq->end());
            push(f);
        }
    }
    return out;
}

std::vector<fft_config> fft_candidates(int n, bool dbl, bool in_cols, bool out_cols, bool inplace) {
    std::vector<fft_config> out;
    fft_config base = default_fft_config(n, dbl, in_cols, out_cols, inplace);
    out.push_back(base);
    std::vector<int> lpbs;
    if (in_cols || out_cols) lpbs = {4, 8, 16, 32};
    else lpbs = {1, 2, 4, 8, 16, 32, 64};
    auto add = [&](const fft_config& c) {
        if (!valid(c) || c.tpl * c.lpb < 64) return;
        for (auto& o : out)
            if (o.tpl == c.tpl && o.lpb == c.lpb && o.pad == c.pad && o.radices == c.radices) return;
        out.push_back(c);
    };
    // other radix schedules first (MEASURE times the first 12 candidates), each with the
    // threads per line of its own largest radix (one and two butterflies per thread and pass)
    // and the default's lines per block and padding
    const auto scheds = radix_schedules(n, 4);
    for (size_t i = 1; i < scheds.size(); i++) {
        const int rmax = *std::max_element(scheds[i].begin(), scheds[i].end());
        for (int t : {n / rmax, n / rmax / 2}) {
            if (t < 1) continue;
            fft_config c = base;
            c.radices = scheds[i];
            c.tpl = t;
            while (c.tpl * c.lpb > 1024 && c.lpb > 1) c.lpb /= 2;
            add(c);
        }
    }
    // launch shapes of the default schedule
    std::vector<int> tpls;
    for (int t = base.tpl; t >= 1 && (int)tpls.size() < 3; t /= 2) tpls.push_back(t);
    for (int t : tpls)
        for (int l : lpbs)
            for (int p : {0, 1, 2, 3}) {
                fft_config c = base;
                c.tpl = t; c.lpb = l; c.pad = p;
                add(c);
            }
    return out;
}

// Any-length kernel: the work line of m points sits whole in LDS.  Bluestein takes the
// smallest m >= 2n - 1 whose primes are <= 7.  One butterfly of the
// largest radix per thread and pass; lines per block grow to about 256 threads while the
// tile stays within 64 KiB of LDS, so that several blocks share a CU.
bool any_config(int kind, int n, bool dbl, bool direct, fft_config& c) {
    c = fft_config{};
    c.any = 1;
    c.kind = kind;
    c.nlog = n;
    c.dbl = dbl;
    c.chirp = !direct;
    const int64_t es = elem_bytes(dbl);
    const int64_t mmax = (int64_t)(kMaxLds / (size_t)es);
    int64_t m = direct ? n : 2 * (int64_t)n - 1;
    for (;; m++) {
        if (m > mmax) return false;
        if (!factorize((int)m, c.radices)) {
            if (direct) return false;
            continue;
        }
        if (direct) break;
        bool small = true;
        for (int r : c.radices) small &= (r == 1 || r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 16);
        if (small) break;
    }
    c.n = (int)m;
    const int rmax = *std::max_element(c.radices.begin(), c.radices.end());
    c.tpl = std::max<int>(1, (int)(m / rmax));
    while (c.tpl > 1024) c.tpl = (c.tpl + 1) / 2;
    c.lpb = 1;
    c.pad = 0;
    while (c.lpb * c.tpl < 256 && c.lpb * 2 * c.tpl <= 1024 &&
           (size_t)(c.lpb * 2) * lds_line(c.n, 1) * (size_t)es <= 64 * 1024)
        c.lpb *= 2;
    if (c.lpb > 1) c.pad = 1;
    return valid(c);
}

// ---- twiddles -------------------------------------------------------------------------
twiddles::~twiddles() {
    // Tables live in a process-wide cache and are released at exit; freeing device memory
    // from a static destructor after the runtime shut down is not safe, so leave it.
}

static std::mutex g_tw_mu;
static std::map<std::string, std::shared_ptr<twiddles>> g_tw;

int32_t get_twiddles(int n, const std::vector<int>& radices, bool dbl, std::shared_ptr<twiddles>* out) {
    std::string key = std::to_string(n) + (dbl ? "d" : "f");
    for (int r : radices) key += "," + std::to_string(r);
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) { *out = it->second; return DTFFT_HIP_SUCCESS; }
    auto t = std::make_shared<twiddles>();
    int64_t total = 0, L = 1;
    std::vector<std::array<int64_t, 3>> passes;  // (off, L, R)
    for (int r : radices) {
        passes.push_back({total, L, r});
        total += L * (r - 1);
        L *= r;
    }
    const rtc_kernel* k = nullptr;
    TRY(rtc_get("fft", dbl ? "gen_twiddles_f64" : "gen_twiddles_f32", &k));
    t->bytes = std::max<int64_t>(total, 1) * 16;
    if (!compile_only()) {
        HIP_TRY(hipMalloc(&t->ptr, (size_t)t->bytes));
        for (auto& p : passes) {
            int off = (int)p[0], Lp = (int)p[1], R = (int)p[2];
            int cnt = Lp * (R - 1);
            if (cnt <= 0) continue;
            void* ptr = t->ptr;
            void* args[] = {&ptr, &off, &Lp, &R};
            TRY(launch(k, (unsigned)((cnt + 255) / 256), 256, nullptr, args));
        }
        HIP_TRY(hipDeviceSynchronize());
    }
    g_tw[key] = t;
    *out = t;
    return DTFFT_HIP_SUCCESS;
}

int32_t get_rtwiddles(int n, bool dbl, std::shared_ptr<twiddles>* out) {
    std::string key = "r" + std::to_string(n) + (dbl ? "d" : "f");
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) { *out = it->second; return DTFFT_HIP_SUCCESS; }
    auto t = std::make_shared<twiddles>();
    int cnt = n / 2 + 1;
    const rtc_kernel* k = nullptr;
    TRY(rtc_get("fft", dbl ? "gen_rtwiddles_f64" : "gen_rtwiddles_f32", &k));
    t->bytes = (int64_t)cnt * 16;
    if (!compile_only()) {
        HIP_TRY(hipMalloc(&t->ptr, (size_t)t->bytes));
        void* ptr = t->ptr;
        int off = 0, nn = n;
        void* args[] = {&ptr, &off, &nn, &cnt};
        TRY(launch(k, (unsigned)((cnt + 255) / 256), 256, nullptr, args));
        HIP_TRY(hipDeviceSynchronize());
    }
    g_tw[key] = t;
    *out = t;
    return DTFFT_HIP_SUCCESS;
}

int32_t get_bluestein(int n, int m, bool dbl, std::shared_ptr<twiddles>* chirp, std::shared_ptr<twiddles>* bh) {
    const std::string kc = "c" + std::to_string(n) + (dbl ? "d" : "f");
    const std::string kb = "b" + std::to_string(n) + "," + std::to_string(m) + (dbl ? "d" : "f");
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto ic = g_tw.find(kc), ib = g_tw.find(kb);
    if (ic != g_tw.end() && ib != g_tw.end()) { *chirp = ic->second; *bh = ib->second; return DTFFT_HIP_SUCCESS; }
    const std::string t = dbl ? "double" : "float";
    const rtc_kernel *kch = nullptr, *kbh = nullptr;
    TRY(rtc_get("anylen", "gen_chirp<" + t + ">", &kch));
    TRY(rtc_get("anylen", "gen_bluestein_bh<" + t + ">", &kbh));
    auto c = std::make_shared<twiddles>(), b = std::make_shared<twiddles>();
    c->bytes = (int64_t)n * 16;
    b->bytes = (int64_t)m * 16;
    if (!compile_only()) {
        HIP_TRY(hipMalloc(&c->ptr, (size_t)c->bytes));
        HIP_TRY(hipMalloc(&b->ptr, (size_t)b->bytes));
        int nn = n, mm = m;
        void* ac[] = {&c->ptr, &nn};
        TRY(launch(kch, (unsigned)((n + 255) / 256), 256, nullptr, ac));
        void* ab[] = {&b->ptr, &nn, &mm};
        TRY(launch(kbh, (unsigned)m, 256, nullptr, ab));
        HIP_TRY(hipDeviceSynchronize());
    }
    g_tw[kc] = c;
    g_tw[kb] = b;
    *chirp = c;
    *bh = b;
    return DTFFT_HIP_SUCCESS;
}

// ---- stages ---------------------------------------------------------------------------
lines_desc make_lines(int64_t nl, int64_t nb, int64_t in_ls, int64_t in_bs, int64_t in_ns, int64_t out_ls,
                      int64_t out_bs, int64_t out_ns, int lpb) {
    lines_desc d{};
    d.nl = nl; d.nb = nb;
    d.in_ls = in_ls; d.in_bs = in_bs; d.in_ns = in_ns;
    d.out_ls = out_ls; d.out_bs = out_bs; d.out_ns = out_ns;
    d.ntl = (nl + lpb - 1) / lpb;
    d.ntiles = d.ntl * nb;
    return d;
}

int32_t fft_stage::init(const fft_config& c, const lines_desc& desc) {
    cfg = c;
    d = desc;
    d.ntl = (d.nl + c.lpb - 1) / c.lpb;
    d.ntiles = d.ntl * d.nb;
    if (c.any) {
        TRY(rtc_get("anylen", c.expr(), &k));
        TRY(get_twiddles(c.n, c.radices, c.dbl, &tw));
        if (c.chirp) TRY(get_bluestein(c.nlog, c.n, c.dbl, &chirp, &bh));
        block = (unsigned)(c.tpl * c.lpb);
        grid = (unsigned)std::min<int64_t>(d.ntiles, 1 << 22);
        return DTFFT_HIP_SUCCESS;
    }
    if (c.n > 1 || c.kind != 0) {
        TRY(rtc_get("fft", c.expr(), &k));
        TRY(get_twiddles(c.n, c.radices, c.dbl, &tw));
    }
    if (c.kind != 0) TRY(get_rtwiddles(2 * c.n, c.dbl, &rtw));
    block = (unsigned)(c.tpl * c.lpb);
    grid = (unsigned)std::min<int64_t>(d.ntiles, 1 << 22);
    if (c.twl && (c.kind == 0 || c.kind == 2) && !c.dma) {
        // persistent: as many blocks as are resident, each copying the twiddles to LDS once
        int per_cu = 1;
        if (!compile_only() && k && k->fn &&
            hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k->fn, (int)block, 0) != hipSuccess)
            per_cu = 1;
        per_cu = std::max(1, per_cu);
        grid = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>(d.ntiles, (int64_t)per_cu * std::max(1, device().num_cus)));
    }
    if (c.dma) {
        if (!dma_ok(c, d)) {
            set_last_error("LDS-DMA column kernel requested for an unsuitable stage");
            return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
        }
        fft_config alt = c;
        alt.dma = 0;
        if (c.kind == 2) {
            // the fallback for misaligned buffers: the plain fused C2R of the same block shape
            alt.twl = 0;
            apply_twl(alt, d, global_config().twl);
        }
        TRY(rtc_get("fft", alt.expr(), &k_alt));
        grid_alt = grid;
        if (c.dma == 1) {
            // persistent: as many blocks as are resident (one per CU with two tile buffers)
            int per_cu = 1;
            if (!compile_only() && k && k->fn &&
                hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k->fn, (int)block, 0) != hipSuccess)
                per_cu = 1;
            per_cu = std::max(1, per_cu);
            grid = (unsigned)std::max<int64_t>(
                1, std::min<int64_t>(d.ntiles, (int64_t)per_cu * std::max(1, device().num_cus)));
        }
        // dma == 2: one tile per block (the kernel strides by the grid beyond 2^22 tiles).
        // Blocks start in tile order, so the blocks resident together gather neighbouring
        // 128-byte pieces of the same rows; on MI355X this ran the 512^3 column-in / row-out
        // pass in 430 us against 446 us for the persistent double buffer, whose blocks drift
        // apart over their 64 tiles each.
    }
    return DTFFT_HIP_SUCCESS;
}

int32_t fft_stage::run(const void* in, void* out, int sign, hipStream_t s) const {
    if (d.ntiles == 0) return DTFFT_HIP_SUCCESS;
    if (cfg.any) {
        void* twp = tw ? tw->ptr : nullptr;
        void* cp = chirp ? chirp->ptr : nullptr;
        void* bp = bh ? bh->ptr : nullptr;
        lines_desc dd = d;
        int inv = sign > 0 ? 1 : 0, nl = cfg.nlog;
        void* args[] = {(void*)&in, &out, &twp, &cp, &bp, &dd, &inv, &nl};
        return launch(k, grid, block, s, args);
    }
    if (cfg.kind != 0) {
        void* twp = tw ? tw->ptr : nullptr;
        void* rtwp = rtw ? rtw->ptr : nullptr;
        lines_desc dd = d;
        void* args[] = {(void*)&in, &out, &twp, &rtwp, &dd};
        if (cfg.dma && ((uintptr_t)in % 16) != 0) return launch(k_alt, grid_alt, block, s, args);
        return launch(k, grid, block, s, args);
    }
    if (cfg.n == 1) {
        // length-1 transform is the identity (copy the lines)
        if (in == out) return DTFFT_HIP_SUCCESS;
        if (d.in_ls == d.out_ls && d.in_bs == d.out_bs && (d.nb == 1 || d.in_bs == d.nl * d.in_ls) && d.in_ls == 1) {
            HIP_TRY(hipMemcpyAsync(out, in, (size_t)(d.nl * d.nb) * elem_bytes(cfg.dbl), hipMemcpyDeviceToDevice, s));
            return DTFFT_HIP_SUCCESS;
        }
        set_last_error("length-1 FFT stage with a strided layout");
        return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
    }
    int inv = sign > 0 ? 1 : 0;
    void* twp = tw ? tw->ptr : nullptr;
    lines_desc dd = d;
    void* args[] = {(void*)&in, &out, &twp, &dd, &inv};
    if (cfg.dma && (((uintptr_t)in % 16) != 0 || ((uintptr_t)out % 16) != 0))
        return launch(k_alt, grid_alt, block, s, args);
    return launch(k, grid, block, s, args);
}

// "r0,r1,..." -> radices whose product is n, each a radix the kernels have
bool parse_radices(const char* text, int n, std::vector<int>& out) {
    out.clear();
    long prod = 1;
    const char* p = text;
    while (*p) {
        char* end = nullptr;
        const long r = std::strtol(p, &end, 10);
        if (end == p) return false;
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 11 && r != 13 && r != 16) return false;
        out.push_back((int)r);
        prod *= r;
        p = end;
        if (*p == ',') p++;
        else if (*p) return false;
    }
    return !out.empty() && out.size() <= 10 && prod == n;
}

static int64_t span(const lines_desc& d, bool in, int n) {
    int64_t ls = in ? d.in_ls : d.out_ls, bs = in ? d.in_bs : d.out_bs, ns = in ? d.in_ns : d.out_ns;
    return (d.nl - 1) * ls + (d.nb - 1) * bs + (int64_t)(n - 1) * ns + 1;
}

int32_t plan_fft_stage(int n, bool dbl, int64_t nl, int64_t nb, int64_t in_ls, int64_t in_bs, int64_t in_ns,
                       int64_t out_ls, int64_t out_bs, int64_t out_ns, bool inplace, int effort,
                       hipStream_t stream, fft_stage* st, int lds_dma) {
    std::vector<int> rad;
    if (!factorize(n, rad)) {
        // a prime factor above 13, or more passes than a schedule holds: Bluestein
        fft_config c;
        if (!any_config(0, n, dbl, false, c)) {
            set_last_error("FFT length " + std::to_string(n) + ": the Bluestein work line does not fit in LDS");
            return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
        }
        return st->init(c, make_lines(nl, nb, in_ls, in_bs, in_ns, out_ls, out_bs, out_ns, c.lpb));
    }
    const bool in_cols = (in_ns != 1 && in_ls == 1);
    const bool out_cols = (out_ns != 1 && out_ls == 1);
    fft_config cfg = default_fft_config(n, dbl, in_cols, out_cols, inplace);
    lines_desc d = make_lines(nl, nb, in_ls, in_bs, in_ns, out_ls, out_bs, out_ns, cfg.lpb);
    char kb[256];
    std::snprintf(kb, sizeof kb, "fft|%d|%c|%d%d%d|%lld|%lld|%lld,%lld,%lld|%lld,%lld,%lld", n, dbl ? 'd' : 'f',
                  (int)in_cols, (int)out_cols, (int)inplace, (long long)nl, (long long)nb, (long long)in_ls,
                  (long long)in_bs, (long long)in_ns, (long long)out_ls, (long long)out_bs, (long long)out_ns);
    const std::string key(kb);
    const int twl_mode = global_config().twl;
    std::string rec;
    bool tuned = false;
    int rec_twl = -1;  // twl of a tuning record (-1: the record predates the field)
    if (n > 1 && tune_lookup(key, rec)) {
        tuned = true;
        int t, l, p, dm = 0, tw = -1, used = 0;
        char rads[128] = {0};
        const int got = std::sscanf(rec.c_str(), "%d %d %d %d %d %n%127s", &t, &l, &p, &dm, &tw, &used, rads);
        if (got >= 3) {
            fft_config c = cfg;
            c.tpl = t; c.lpb = l; c.pad = p;
            c.dma = dm;
            // the radix schedule, "r0,r1,..." (records without one keep the default)
            std::vector<int> r;
            if (got >= 6 && parse_radices(rads, n, r)) c.radices = r;
            if (valid(c) && (!c.dma || dma_ok(c, d))) { cfg = c; rec_twl = tw; }
        }
    } else if (n > 1 && effort >= DTFFT_HIP_MEASURE && !compile_only()) {
        const int64_t si = span(d, true, n), so = span(d, false, n);
        void *bi = nullptr, *bo = nullptr;
        const size_t es = (size_t)elem_bytes(dbl);
        HIP_TRY(hipMalloc(&bi, (size_t)si * es));
        if (!inplace && hipMalloc(&bo, (size_t)so * es) != hipSuccess) { hipFree(bi); return DTFFT_HIP_ERROR_ALLOC_FAILED; }
        if (inplace) bo = bi;
        TRY(fill_tuning_scratch(bi, (size_t)si * es, dbl, stream));
        double best = 1e30;
        auto cands = fft_candidates(n, dbl, in_cols, out_cols, inplace);
        if (effort < DTFFT_HIP_EXHAUSTIVE && cands.size() > 12) cands.resize(12);
        for (size_t i = 0, nc = cands.size(); lds_dma != 0 && i < nc; i++) {  // LDS-DMA twins
            for (int v : {1, 2}) {
                fft_config c = cands[i];
                c.dma = v;
                if (dma_ok(c, d)) cands.push_back(c);
            }
        }
        // each candidate is timed in the form it would run: with the LDS twiddle table where
        // the rule of this configuration gives it one, and at EXHAUSTIVE also in the other form
        for (size_t i = 0, nc = cands.size(); i < nc; i++) {
            apply_twl(cands[i], d, twl_mode);
            if (effort >= DTFFT_HIP_EXHAUSTIVE && twl_mode != 0 && !cands[i].dma) {
                fft_config c = cands[i];
                apply_twl(c, d, c.twl ? 0 : 2);
                if (c.twl != cands[i].twl) cands.push_back(c);
            }
        }
        std::vector<fft_stage> built;
        std::vector<fft_config> kept;
        for (auto& c : cands) {
            fft_stage cand;
            if (cand.init(c, d)) continue;
            built.push_back(cand);
            kept.push_back(c);
        }
        std::vector<std::function<int32_t()>> runs;
        for (auto& st : built) runs.push_back([&st, bi, bo, stream]() { return st.run(bi, bo, -1, stream); });
        const std::vector<double> ms = time_candidates(runs, stream);
        for (size_t i = 0; i < kept.size(); i++)
            if (ms[i] < best) { best = ms[i]; cfg = kept[i]; }
        hipStreamSynchronize(stream);
        hipFree(bi);
        if (!inplace) hipFree(bo);
        std::string r;
        for (size_t i = 0; i < cfg.radices.size(); i++) r += (i ? "," : "") + std::to_string(cfg.radices[i]);
        char buf[160];
        std::snprintf(buf, sizeof buf, "%d %d %d %d %d %s", cfg.tpl, cfg.lpb, cfg.pad, cfg.dma, cfg.twl, r.c_str());
        tune_store(key, buf);
        tuned = true;
        rec_twl = cfg.twl;
    }
    if (lds_dma == 0) {
        cfg.dma = 0;
    } else if (!cfg.dma && (lds_dma >= 2 || (!tuned && !dbl))) {
        // Default for FP32 column-in passes: the LDS-DMA kernels measured 10-14 % faster than
        // the plain column kernel on MI355X.  512^3 column-in / row-out pass: 431 us with one
        // tile per block, 444 us persistent with two buffers, 498 us plain.  The in-place
        // column pass runs best persistent with two buffers and half the threads per line:
        // 444 us, against 475 us with one tile per block and 495 us plain.  FP64 keeps the
        // plain kernel, which the autotuner preferred there.
        fft_config c = cfg;
        c.dma = lds_dma == 2 ? 1 : lds_dma == 3 ? 2 : (c.out_cols ? 1 : 2);
        c.lpb = 128 / elem_bytes(dbl);  // a tile row is one 128-byte piece
        if (c.out_cols && lds_dma < 2) c.tpl = std::max(1, c.tpl / 2);
        bool wide = false;
        if (lds_dma == 1 && !dbl && inplace && cfg.out_cols && n >= 512) {
            // In-place FP32 column passes of n >= 512 where a 32-line tile (256-byte pieces of
            // every row) fits with the LDS twiddle table: the plain kernel, persistent, four
            // butterflies per thread and pass.  EXHAUSTIVE picked it at 512^3 in three runs,
            // 416 us against 427 us for the LDS-DMA double buffer (the 32-line tile fits only
            // up to n = 512; at 256^3 and 384^3 the tuner picked other kernels, which keep the
            // rule below).
            fft_config w = cfg;
            const int rmax = *std::max_element(w.radices.begin(), w.radices.end());
            w.dma = 0;
            w.lpb = 32;
            w.pad = 1;
            w.tpl = std::max(1, n / rmax / 4);
            apply_twl(w, d, 1);  // the measured form: persistent with LDS twiddles (>= 2048 tiles)
            if (w.twl && valid(w)) { cfg = w; rec_twl = -1; wide = true; }
        }
        if (!wide && dma_ok(c, d)) { cfg = c; rec_twl = -1; }
    }
    // persistent blocks with the twiddle table in LDS: as the tuning record says (the form it
    // was timed in), else by the rule of the configuration
    if (rec_twl >= 0 && twl_mode == 1) {
        apply_twl(cfg, d, rec_twl ? 2 : 0);
    } else {
        apply_twl(cfg, d, twl_mode);
    }
    return st->init(cfg, d);
}

int32_t plan_real_stage(int kind, int n, bool dbl, int64_t how_many, int64_t idist, int64_t odist, int effort,
                        hipStream_t stream, fft_stage* st) {
    (void)effort; (void)stream;
    if (n < 1) {
        set_last_error("R2C/C2R length must be positive, got " + std::to_string(n));
        return DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE;
    }
    const int m = n / 2;
    fft_config c;
    if (n % 2 != 0 || !factorize(m, c.radices)) {
        // odd lengths (directly when n factorizes) and halves with a prime factor above 13
        // (Bluestein) take the any-length kernel: real side in real units
        std::vector<int> rad;
        const bool direct = n % 2 != 0 && factorize(n, rad);
        if (!any_config(kind, n, dbl, direct, c)) {
            set_last_error("R2C/C2R length " + std::to_string(n) + ": the work line does not fit in LDS");
            return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
        }
        return st->init(c, make_lines(how_many, 1, idist, 0, 1, odist, 0, 1, c.lpb));
    }
    c = default_fft_config(m, dbl, false, false, false);
    c.kind = kind;
    if (m == 1) {
        c.radices = {1}; c.tpl = 1; c.lpb = 64;
    } else {
        // one butterfly of the largest radix per thread and pass: the untangling adds
        // registers to the row FFT, and fewer per thread keep more waves resident
        const int rmax = *std::max_element(c.radices.begin(), c.radices.end());
        if ((m / rmax) * c.lpb <= 1024) c.tpl = m / rmax;
    }
    lines_desc d;
    if (kind == 1) {  // real lines (idist reals) viewed as complex; complex out (odist)
        if (idist % 2) { set_last_error("R2C idist must be even"); return DTFFT_HIP_ERROR_INVALID_ARGUMENT; }
        d = make_lines(how_many, 1, idist / 2, 0, 1, odist, 0, 1, c.lpb);
    } else {          // complex in (idist), real out viewed as complex (odist / 2)
        if (odist % 2) { set_last_error("C2R odist must be even"); return DTFFT_HIP_ERROR_INVALID_ARGUMENT; }
        d = make_lines(how_many, 1, idist, 0, 1, odist / 2, 0, 1, c.lpb);
    }
    apply_twl(c, d, global_config().twl);  // C2R rows: only when forced (no gain measured)
    return st->init(c, d);
}

int32_t plan_c2r_cols_stage(int n, bool dbl, int64_t nl, int64_t nb, int64_t in_bs, int64_t in_ns, int64_t out_ls,
                            int64_t out_bs, fft_stage* st, int lds_dma) {
    if (n < 2 || n % 2 != 0) {
        set_last_error("C2R stages need an even length, got " + std::to_string(n));
        return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
    }
    const int m = n / 2;
    fft_config c;
    if (!factorize(m, c.radices) || m < 2) {
        // Bluestein (or the direct form for n = 2) on the any-length kernel, which writes
        // real units: the strides of the complex view double
        std::vector<int> rad;
        if (!any_config(2, n, dbl, factorize(n, rad) && m < 2, c)) {
            set_last_error("fused C2R length " + std::to_string(n) + ": the work line does not fit in LDS");
            return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
        }
        return st->init(c, make_lines(nl, nb, 1, in_bs, in_ns, 2 * out_ls, 2 * out_bs, 1, c.lpb));
    }
    c = default_fft_config(m, dbl, true, false, false);
    c.kind = 2;
    lines_desc d = make_lines(nl, nb, 1, in_bs, in_ns, out_ls, out_bs, 1, c.lpb);
    // persistent blocks with both twiddle tables in LDS by the configuration's rule (FP32 by
    // default: one 1024-thread block per CU either way)
    apply_twl(c, d, global_config().twl);
    if (lds_dma != 0) {
        // LDS-DMA kernels: 128-byte input row pieces loaded straight into LDS, each read once;
        // two butterflies per thread and pass.  lds_dma 2: the persistent double buffer
        // (fft_c2r_dma2_kernel, one block per CU); 3: one tile per block (fft_c2r_dma_kernel,
        // two blocks per CU, the FFT twiddles in LDS only when forced with twl = 2, which leaves
        // room for one block per CU); 1: the double buffer where it fits, else one tile per block
        fft_config x = c;
        x.lpb = 128 / elem_bytes(dbl);
        x.pad = 1;
        const int rmax = *std::max_element(x.radices.begin(), x.radices.end());
        x.tpl = std::max(1, (m / rmax) / 2);
        const lines_desc dx = make_lines(nl, nb, 1, in_bs, in_ns, out_ls, out_bs, 1, x.lpb);
        for (int v : {1, 2}) {
            if ((lds_dma == 2 && v != 1) || (lds_dma == 3 && v != 2)) continue;
            x.dma = v;
            x.twl = v == 2 && global_config().twl == 2 ? 1 : 0;
            if (dma_ok(x, dx)) return st->init(x, dx);
        }
    }
    return st->init(c, d);
}

// ---- executor handle -------------------------------------------------------------------
struct fft_plan {
    int rank = 1, type = DTFFT_HIP_FFT_C2C;
    bool dbl = false;
    hipStream_t stream = nullptr;
    // C2C: rank 1 one stage; rank 2 fast axis (in -> out) then slow axis (in place on out).
    std::vector<fft_stage> stages;
    // R2C: forward R2C along the fast axis then C2C along the slow one (in place on out);
    // backward C2C along the slow axis in place on the input (FFTW_DESTROY_INPUT, as the
    // reference's FFTW flags allow) then C2R along the fast axis.
    std::vector<fft_stage> bwd;
    // R2R: forward kinds along the fast axis (in -> out) then the slow one (in place on
    // out); backward the same order with the inverse kinds.  One work buffer serves all.
    std::vector<r2r_stage> rf, rb;
    void* work = nullptr;
    fft_plan() = default;
    fft_plan(const fft_plan&) = delete;
    ~fft_plan() { if (work) hipFree(work); }
};

}  // namespace dtfft_hip

using namespace dtfft_hip;

extern "C" {

int32_t dtfft_hip_fft_create(int8_t fft_rank, int8_t fft_type, int32_t precision, int32_t idist, int32_t odist,
                             int32_t how_many, const int32_t* fft_sizes, const int32_t* inembed,
                             const int32_t* onembed, const int32_t* r2r_kinds, int32_t effort, void* stream,
                             void** plan) {
    (void)inembed;
    if (!plan || !fft_sizes) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *plan = nullptr;
    if (fft_rank != 1 && fft_rank != 2) return DTFFT_HIP_ERROR_INVALID_N_DIMENSIONS;
    if (precision != DTFFT_HIP_SINGLE && precision != DTFFT_HIP_DOUBLE) return DTFFT_HIP_ERROR_INVALID_PRECISION;
    if (fft_type != DTFFT_HIP_FFT_C2C && fft_type != DTFFT_HIP_FFT_R2C && fft_type != DTFFT_HIP_FFT_R2R)
        return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (fft_type == DTFFT_HIP_FFT_R2R) {
        if (!r2r_kinds) return DTFFT_HIP_ERROR_MISSING_R2R_KINDS;
        for (int i = 0; i < fft_rank; i++)
            if (!r2r_kind_valid(r2r_kinds[i])) return DTFFT_HIP_ERROR_INVALID_R2R_KINDS;
    }
    if (how_many < 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (effort < DTFFT_HIP_ESTIMATE || effort > DTFFT_HIP_EXHAUSTIVE) return DTFFT_HIP_ERROR_INVALID_EFFORT_FLAG;
    for (int i = 0; i < fft_rank; i++) if (fft_sizes[i] <= 0) return DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE;
    auto p = std::make_unique<fft_plan>();
    p->rank = fft_rank;
    p->type = fft_type;
    p->dbl = precision == DTFFT_HIP_DOUBLE;
    p->stream = (hipStream_t)stream;
    if (fft_type == DTFFT_HIP_FFT_R2R) {
        for (int dir = 0; dir < 2; dir++) {
            auto& v = dir == 0 ? p->rf : p->rb;
            auto kind_of = [&](int i) { return dir == 0 ? r2r_kinds[i] : r2r_inverse_kind(r2r_kinds[i]); };
            r2r_stage a;
            if (fft_rank == 1) {
                TRY(a.init(kind_of(0), fft_sizes[0], p->dbl, how_many, 1, dir == 0 ? idist : odist, 0, 1,
                           dir == 0 ? odist : idist, 0, 1, effort, p->stream));
                v.push_back(a);
            } else {
                const int64_t n0 = fft_sizes[0], n1 = fft_sizes[1];  // n0 slow, n1 fast (FFTW order)
                const int64_t id = dir == 0 ? idist : odist, od = dir == 0 ? odist : idist;
                r2r_stage b;
                TRY(a.init(kind_of(1), n1, p->dbl, n0, how_many, n1, id, 1, n1, od, 1, effort, p->stream));
                TRY(b.init(kind_of(0), n0, p->dbl, n1, how_many, 1, od, n1, 1, od, n1, effort, p->stream));
                v.push_back(a);
                v.push_back(b);
            }
        }
        size_t wb = 0;
        for (auto* v : {&p->rf, &p->rb})
            for (auto& st : *v) wb = std::max(wb, st.work_bytes());
        if (!compile_only()) HIP_TRY(hipMalloc(&p->work, std::max<size_t>(wb, 16)));
    } else if (fft_type == DTFFT_HIP_FFT_C2C) {
        if (fft_rank == 1) {
            fft_stage st;
            TRY(plan_fft_stage(fft_sizes[0], p->dbl, how_many, 1, idist, 0, 1, odist, 0, 1, false, effort,
                               p->stream, &st));
            p->stages.push_back(st);
        } else {
            const int n0 = fft_sizes[0], n1 = fft_sizes[1];  // n0 slow, n1 fast (FFTW order)
            fft_stage a, b;
            TRY(plan_fft_stage(n1, p->dbl, (int64_t)n0 * how_many, 1, n1, 0, 1, n1, 0, 1, false, effort,
                               p->stream, &a));
            TRY(plan_fft_stage(n0, p->dbl, n1, how_many, 1, odist, n1, 1, odist, n1, true, effort, p->stream, &b));
            p->stages.push_back(a);
            p->stages.push_back(b);
        }
    } else {
        if (fft_rank == 1) {
            fft_stage f, b;
            TRY(plan_real_stage(1, fft_sizes[0], p->dbl, how_many, idist, odist, effort, p->stream, &f));
            TRY(plan_real_stage(2, fft_sizes[0], p->dbl, how_many, odist, idist, effort, p->stream, &b));
            p->stages.push_back(f);
            p->bwd.push_back(b);
        } else {
            const int n0 = fft_sizes[0], n1 = fft_sizes[1];
            const int c1 = onembed ? onembed[1] : n1 / 2 + 1;  // complex fast extent
            fft_stage f, fs, bs, b;
            TRY(plan_real_stage(1, n1, p->dbl, (int64_t)n0 * how_many, n1, c1, effort, p->stream, &f));
            TRY(plan_fft_stage(n0, p->dbl, c1, how_many, 1, odist, c1, 1, odist, c1, true, effort, p->stream, &fs));
            TRY(plan_fft_stage(n0, p->dbl, c1, how_many, 1, odist, c1, 1, odist, c1, true, effort, p->stream, &bs));
            TRY(plan_real_stage(2, n1, p->dbl, (int64_t)n0 * how_many, c1, n1, effort, p->stream, &b));
            p->stages.push_back(f);
            p->stages.push_back(fs);
            p->bwd.push_back(bs);
            p->bwd.push_back(b);
        }
    }
    *plan = p.release();
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_fft_execute(void* plan, const void* in, void* out, int8_t sign) {
    auto* p = (fft_plan*)plan;
    if (!p) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (sign != DTFFT_HIP_FFT_FORWARD && sign != DTFFT_HIP_FFT_BACKWARD) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (p->type == DTFFT_HIP_FFT_R2R) {
        const auto& v = sign == DTFFT_HIP_FFT_FORWARD ? p->rf : p->rb;
        TRY(v[0].run(in, out, p->work, p->stream));
        for (size_t i = 1; i < v.size(); i++) TRY(v[i].run(out, out, p->work, p->stream));
        return DTFFT_HIP_SUCCESS;
    }
    if (p->type == DTFFT_HIP_FFT_R2C) {
        if (in == out) return DTFFT_HIP_ERROR_INVALID_USAGE;  // src/dtfft_abstract_executor.F90:226-228
        if (sign == DTFFT_HIP_FFT_FORWARD) {
            TRY(p->stages[0].run(in, out, sign, p->stream));
            for (size_t i = 1; i < p->stages.size(); i++) TRY(p->stages[i].run(out, out, sign, p->stream));
        } else {
            if (p->bwd.size() == 2) {
                TRY(p->bwd[0].run(in, (void*)in, sign, p->stream));
                TRY(p->bwd[1].run(in, out, sign, p->stream));
            } else {
                TRY(p->bwd[0].run(in, out, sign, p->stream));
            }
        }
        return DTFFT_HIP_SUCCESS;
    }
    TRY(p->stages[0].run(in, out, sign, p->stream));
    for (size_t i = 1; i < p->stages.size(); i++) TRY(p->stages[i].run(out, out, sign, p->stream));
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_fft_destroy(void* plan) {
    delete (fft_plan*)plan;
    return DTFFT_HIP_SUCCESS;
}

}  // extern "C"
