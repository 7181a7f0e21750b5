// Runtime services of libdtfft_hip: device discovery, hipRTC compilation with an
// in-process and on-disk code-object cache, tuning records, error strings and the
// small memory / stream / event helpers of the C-ABI.
//
// This is the MI355X counterpart of the reference's NVRTC layer
// (src/dtfft_nvrtc_module.F90:260-432 compile + name expressions,
// src/dtfft_nvrtc_module_cache.F90:96-183 process-global module cache): kernels are
// C++ templates instantiated through hipRTC name expressions for gfx950.
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include <algorithm>

#include "internal.h"

#include "device_sources.inc"  // generated at build: k_src_common, k_src_fft, k_src_transpose

namespace dtfft_hip {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) {
    g_last_error = msg;
    if (std::getenv("DTFFT_HIP_DEBUG")) std::fprintf(stderr, "[dtfft_hip] %s\n", msg.c_str());
}
const std::string& last_error() { return g_last_error; }

// ---- device -------------------------------------------------------------------------
static std::once_flag g_dev_once;
static device_info g_dev;
static int g_compile_only = -1;  // -1: not forced

const device_info& device() {
    std::call_once(g_dev_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return;
        g_dev.device = dev;
        g_dev.arch = p.gcnArchName;
        g_dev.num_cus = p.multiProcessorCount;
        g_dev.lds_per_block = p.sharedMemPerBlock;
        g_dev.present = g_dev.arch.rfind("gfx950", 0) == 0;
    });
    return g_dev;
}

bool compile_only() {
    if (g_compile_only >= 0) return g_compile_only != 0;
    const char* e = std::getenv("DTFFT_HIP_COMPILE_ONLY");
    if (e && *e && *e != '0') return true;
    return !device().present;
}

// ---- on-disk cache ----------------------------------------------------------------
static std::mutex g_mu;
static std::string g_cache_dir;
static bool g_cache_dir_set = false;

static std::string lib_dir() {
    Dl_info info;
    if (dladdr((void*)&lib_dir, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        auto pos = p.rfind('/');
        if (pos != std::string::npos) return p.substr(0, pos);
    }
    return ".";
}

std::string cache_dir() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_cache_dir_set) {
        const char* e = std::getenv("DTFFT_HIP_CACHE_DIR");
        g_cache_dir = e ? std::string(e) : lib_dir() + "/kcache";
        g_cache_dir_set = true;
    }
    return g_cache_dir;
}

static void mkdirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

static uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
    return h;
}

static bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

static void write_file_atomic(const std::string& path, const std::string& data) {
    std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), (std::streamsize)data.size());
    }
    std::rename(tmp.c_str(), path.c_str());
}

// ---- hipRTC ---------------------------------------------------------------------------
static const char* source_text(const std::string& name) {
    if (name == "fft") return k_src_fft;
    if (name == "transpose") return k_src_transpose;
    if (name == "r2r") return k_src_r2r;
    if (name == "anylen") return k_src_anylen;
    return nullptr;
}

static const std::vector<std::string>& rtc_options() {
    static const std::vector<std::string> opts = {
        "--offload-arch=gfx950", "-O3", "-std=c++17",
        // keep hipcc from forming packed FP32 instructions
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    };
    return opts;
}

static std::map<std::string, std::unique_ptr<rtc_kernel>> g_kernels;

static int32_t compile(const std::string& src_name, const std::string& expr, std::string& code,
                       std::string& lowered) {
    const char* src = source_text(src_name);
    if (!src) {
        set_last_error("unknown device source " + src_name);
        return DTFFT_HIP_ERROR_RTC_COMPILE;
    }
    // r2r.hip and anylen.hip include fft.hip for the passes of fft_tile
    const char* hdrs[] = {k_src_common, k_src_fft};
    const char* hdr_names[] = {"common.hip", "fft.hip"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src, (src_name + ".hip").c_str(), 2, hdrs, hdr_names) != HIPRTC_SUCCESS) {
        set_last_error("hiprtcCreateProgram failed");
        return DTFFT_HIP_ERROR_RTC_COMPILE;
    }
    hiprtcAddNameExpression(prog, expr.c_str());
    std::vector<const char*> opts;
    for (auto& o : rtc_options()) opts.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        set_last_error("hipRTC compile of " + expr + " failed:\n" + log);
        hiprtcDestroyProgram(&prog);
        return DTFFT_HIP_ERROR_RTC_COMPILE;
    }
    const char* ln = nullptr;
    if (hiprtcGetLoweredName(prog, expr.c_str(), &ln) != HIPRTC_SUCCESS || !ln) {
        set_last_error("hiprtcGetLoweredName failed for " + expr);
        hiprtcDestroyProgram(&prog);
        return DTFFT_HIP_ERROR_RTC_COMPILE;
    }
    lowered = ln;
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, &code[0]);
    hiprtcDestroyProgram(&prog);
    return DTFFT_HIP_SUCCESS;
}

int32_t rtc_get(const std::string& source, const std::string& expr, const rtc_kernel** out) {
    const std::string key = source + "|" + expr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_kernels.find(key);
        if (it != g_kernels.end()) { *out = it->second.get(); return DTFFT_HIP_SUCCESS; }
    }
    const char* src = source_text(source);
    if (!src) { set_last_error("unknown source " + source); return DTFFT_HIP_ERROR_RTC_COMPILE; }
    std::string h = source + "|" + expr + "|" + std::string(src) + "|" + std::string(k_src_common);
    if (source == "r2r" || source == "anylen") h += "|" + std::string(k_src_fft);
    for (auto& o : rtc_options()) h += "|" + o;
    char hex[32];
    std::snprintf(hex, sizeof hex, "%016llx", (unsigned long long)fnv1a(h));
    const std::string dir = cache_dir();
    const std::string base = dir.empty() ? std::string() : dir + "/" + hex;
    std::string code, lowered;
    bool cached = !base.empty() && read_file(base + ".co", code) && read_file(base + ".name", lowered) &&
                  !code.empty() && !lowered.empty();
    if (!cached) {
        TRY(compile(source, expr, code, lowered));
        if (!base.empty()) {
            mkdirs(dir);
            write_file_atomic(base + ".co", code);
            write_file_atomic(base + ".name", lowered);
        }
    }
    auto k = std::make_unique<rtc_kernel>();
    k->expr = expr;
    if (!compile_only()) {
        HIP_TRY(hipModuleLoadData(&k->module, code.data()));
        HIP_TRY(hipModuleGetFunction(&k->fn, k->module, lowered.c_str()));
    }
    std::lock_guard<std::mutex> lk(g_mu);
    auto& slot = g_kernels[key];
    if (!slot) slot = std::move(k);
    *out = slot.get();
    return DTFFT_HIP_SUCCESS;
}

int32_t launch(const rtc_kernel* k, unsigned grid, unsigned block, hipStream_t stream, void** args) {
    if (compile_only()) return DTFFT_HIP_SUCCESS;
    if (!k || !k->fn) { set_last_error("kernel not loaded"); return DTFFT_HIP_ERROR_HIP_RUNTIME; }
    if (grid == 0) return DTFFT_HIP_SUCCESS;
    HIP_TRY(hipModuleLaunchKernel(k->fn, grid, 1, 1, block, 1, 1, 0, stream, args, nullptr));
    return DTFFT_HIP_SUCCESS;
}

// ---- candidate timing -----------------------------------------------------------
std::vector<double> time_candidates(const std::vector<std::function<int32_t()>>& runs, hipStream_t s, int rounds) {
    const dtfft_hip_config_t& c = global_config();
    const int warm = std::max(0, c.measure_warmup_iters), iters = std::max(1, c.measure_iters);
    const size_t n = runs.size();
    std::vector<double> out(n, 1e30);
    std::vector<std::vector<float>> t(n);
    std::vector<char> ok(n, 1);
    for (size_t i = 0; i < n; i++)
        for (int w = 0; w < warm && ok[i]; w++) ok[i] = runs[i]() == 0;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return out;
    if (hipEventCreate(&b) != hipSuccess) { hipEventDestroy(a); return out; }
    for (int r = 0; r < rounds; r++)
        for (size_t i = 0; i < n; i++) {
            if (!ok[i]) continue;
            hipEventRecord(a, s);
            for (int k = 0; k < iters && ok[i]; k++) ok[i] = runs[i]() == 0;
            hipEventRecord(b, s);
            hipEventSynchronize(b);
            float ms = 0;
            hipEventElapsedTime(&ms, a, b);
            if (ok[i]) t[i].push_back(ms / iters);
        }
    hipEventDestroy(a);
    hipEventDestroy(b);
    for (size_t i = 0; i < n; i++) {
        if (!ok[i] || t[i].empty()) continue;
        std::vector<float> v = t[i];
        std::sort(v.begin(), v.end());
        out[i] = v[v.size() / 2];
    }
    return out;
}

int32_t fill_tuning_scratch(void* p, size_t bytes, bool dbl, hipStream_t s) {
    const rtc_kernel* k = nullptr;
    TRY(rtc_get("transpose", "fill_tuning_data", &k));
    int64_t nwords = (int64_t)(bytes / 4);
    int d = dbl ? 1 : 0;
    if (nwords == 0) return DTFFT_HIP_SUCCESS;
    void* args[] = {&p, &nwords, &d};
    return launch(k, (unsigned)std::min<int64_t>((nwords + 255) / 256, 1 << 16), 256, s, args);
}

// ---- tuning records ------------------------------------------------------------
static std::map<std::string, std::string> g_tune;
static bool g_tune_loaded = false;

static std::string tune_path() {
    std::string d = cache_dir();
    return d.empty() ? std::string() : d + "/tuning.txt";
}

// Caller holds g_mu and passes the path (tune_path() takes g_mu itself).
static void tune_load_locked(const std::string& path) {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    std::string data;
    if (path.empty() || !read_file(path, data)) return;
    std::stringstream ss(data);
    std::string line;
    while (std::getline(ss, line)) {
        auto p = line.find('\t');
        if (p != std::string::npos) g_tune[line.substr(0, p)] = line.substr(p + 1);
    }
}

bool tune_lookup(const std::string& key, std::string& value) {
    const std::string path = tune_path();
    std::lock_guard<std::mutex> lk(g_mu);
    tune_load_locked(path);
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    value = it->second;
    return true;
}

void tune_store(const std::string& key, const std::string& value) {
    std::string path = tune_path();
    std::lock_guard<std::mutex> lk(g_mu);
    tune_load_locked(path);
    g_tune[key] = value;
    if (path.empty()) return;
    mkdirs(path.substr(0, path.rfind('/')));
    std::string data;
    for (auto& kv : g_tune) data += kv.first + "\t" + kv.second + "\n";
    write_file_atomic(path, data);
}

}  // namespace dtfft_hip

using namespace dtfft_hip;

extern "C" {

int32_t dtfft_hip_get_version(void) { return 100; }  // 0.1.0

const char* dtfft_hip_get_error_string(int32_t code) {
    switch (code) {
        case DTFFT_HIP_SUCCESS: return "DTFFT_SUCCESS";
        case DTFFT_HIP_ERROR_PLAN_NOT_CREATED: return "Plan not created";
        case DTFFT_HIP_ERROR_INVALID_TRANSPOSE_TYPE: return "Invalid transpose_type provided";
        case DTFFT_HIP_ERROR_INVALID_N_DIMENSIONS: return "Invalid number of dimensions provided";
        case DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE: return "One or more provided dimension sizes <= 0";
        case DTFFT_HIP_ERROR_INVALID_PRECISION: return "Invalid precision parameter provided";
        case DTFFT_HIP_ERROR_INVALID_EFFORT_FLAG: return "Invalid effort parameter provided";
        case DTFFT_HIP_ERROR_INVALID_EXECUTOR_TYPE: return "Invalid executor parameter provided";
        case DTFFT_HIP_ERROR_MISSING_R2R_KINDS: return "Missing R2R kinds";
        case DTFFT_HIP_ERROR_INVALID_R2R_KINDS: return "Invalid R2R kinds";
        case DTFFT_HIP_ERROR_R2C_TRANSPOSE_PLAN: return "Transpose plan is not supported for R2C";
        case DTFFT_HIP_ERROR_INPLACE_TRANSPOSE: return "Inplace transpose is not supported";
        case DTFFT_HIP_ERROR_INVALID_AUX: return "Invalid aux buffer provided";
        case DTFFT_HIP_ERROR_INVALID_USAGE: return "Invalid usage";
        case DTFFT_HIP_ERROR_ALLOC_FAILED: return "Failed to allocate memory";
        case DTFFT_HIP_ERROR_FREE_FAILED: return "Failed to free memory";
        case DTFFT_HIP_ERROR_INVALID_ALLOC_BYTES: return "Invalid alloc_bytes provided";
        case DTFFT_HIP_ERROR_INVALID_EXECUTE_TYPE: return "Invalid execute_type provided";
        case DTFFT_HIP_ERROR_R2R_FFT_NOT_SUPPORTED: return "R2R FFT is not supported by this executor";
        case DTFFT_HIP_ERROR_GPU_INVALID_STREAM: return "Invalid stream provided";
        case DTFFT_HIP_ERROR_NOT_DEVICE_PTR: return "Pointer is not a device pointer";
        case DTFFT_HIP_ERROR_HIP_RUNTIME: return "HIP runtime call failed";
        case DTFFT_HIP_ERROR_RTC_COMPILE: return "hipRTC compilation failed";
        case DTFFT_HIP_ERROR_UNSUPPORTED_SIZE: return "FFT length not supported by the HIP executor";
        case DTFFT_HIP_ERROR_INVALID_KERNEL_TYPE: return "Invalid kernel type";
        case DTFFT_HIP_ERROR_INVALID_NEIGHBOR: return "Neighbor index out of bounds";
        case DTFFT_HIP_ERROR_INDEX_OVERFLOW: return "Local size exceeds 2^31-1 elements";
        case DTFFT_HIP_ERROR_NO_DEVICE: return "No gfx950 device available";
        case DTFFT_HIP_ERROR_INVALID_ARGUMENT: return "Invalid argument";
        case DTFFT_HIP_ERROR_SCRATCH_TOO_SMALL: return "Aux buffer larger than the no-alloc scratch";
        default: return "Unknown error";
    }
}

// Detailed message of the last failure on this thread (hipRTC log, HIP error text).
const char* dtfft_hip_get_last_error_message(void) { return last_error().c_str(); }

int32_t dtfft_hip_set_cache_dir(const char* path) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache_dir = path ? std::string(path) : std::string();
    g_cache_dir_set = true;
    g_tune.clear();
    g_tune_loaded = false;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_set_compile_only(int32_t on) {
    g_compile_only = on ? 1 : 0;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_mem_alloc(size_t bytes, void** ptr) {
    if (!ptr) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *ptr = nullptr;
    if (bytes == 0) return DTFFT_HIP_ERROR_INVALID_ALLOC_BYTES;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        set_last_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        *ptr = nullptr;
        return DTFFT_HIP_ERROR_ALLOC_FAILED;
    }
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_mem_free(void* ptr) {
    if (!ptr) return DTFFT_HIP_SUCCESS;
    if (hipFree(ptr) != hipSuccess) return DTFFT_HIP_ERROR_FREE_FAILED;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_memcpy_d2d_async(void* dst, const void* src, size_t bytes, void* stream) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_stream_create(void** stream) {
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_stream_destroy(void* stream) {
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_stream_synchronize(void* stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_event_create(void** event) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *event = (void*)e;
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_event_destroy(void* event) {
    HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_event_record(void* event, void* stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_event_elapsed_ms(void* start, void* stop, float* ms) {
    HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return DTFFT_HIP_SUCCESS;
}
int32_t dtfft_hip_get_device_arch(char* buf, int32_t buflen) {
    if (!buf || buflen <= 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    std::snprintf(buf, (size_t)buflen, "%s", device().arch.c_str());
    return DTFFT_HIP_SUCCESS;
}

}  // extern "C"
