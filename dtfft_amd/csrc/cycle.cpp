// Batched cycle over many independent 3D C2C grids (no counterpart in the reference,
// where every plan launches its own kernels).
//
// The grids are packed one after another in each buffer.  A direction is three stages,
// the fused sequence of a single-rank plan:
//   forward : FFT_X (in -> out), X->Y + FFT_Y (out -> aux), Y->Z + FFT_Z (aux -> out)
//   backward: FFT_Z (in -> aux), Z->X + FFT_X (aux -> out), FFT_Y in place (out)
// Within a stage the grids are grouped by the length of the stage's axis.  Each group is
// one persistent launch of a kernel compiled for that length alone (so its register budget
// is that of a single-length kernel), walking a work list of (grid, tile) items.  The groups
// of a stage are independent: they are spread over side streams, forked from and joined
// back to the caller's stream with events, so that small groups overlap.  The
// forward+backward cycle is captured once as a hipGraph (the forks become parallel
// branches) and replayed.
#include <algorithm>
#include <array>
#include <cmath>
#include <atomic>
#include <memory>
#include <numeric>
#include <thread>

#include "internal.h"

namespace dtfft_hip {

struct grid_stage_h {  // mirrors dtfft::grid_stage
    lines_desc d;
    int64_t in_off, out_off;
    const void* tw;
    int32_t cls;
    int32_t pad_;
};

struct stage_spec {  // per-grid geometry of one stage
    int64_t n, nl, nb, in_ls, in_bs, in_ns, out_ls, out_bs, out_ns;
};

// One launch: every grid of a stage whose axis has this length.
struct length_group {
    fft_config cfg;
    std::shared_ptr<twiddles> tw;
    const rtc_kernel* k = nullptr;
    void* d_gs = nullptr;
    void* d_items = nullptr;
    int nitems = 0;
    unsigned grid = 0, block = 0;
    int64_t work = 0;  // elements moved, for ordering the launches
    ~length_group() {
        if (d_gs) hipFree(d_gs);
        if (d_items) hipFree(d_items);
    }
};

static constexpr int kLanes = 4;  // the caller's stream + 3 side streams

struct cycle_stage {
    std::vector<std::unique_ptr<length_group>> groups;  // largest first
    hipEvent_t fork = nullptr;
    hipEvent_t join[kLanes] = {};
    ~cycle_stage() {
        if (fork) hipEventDestroy(fork);
        for (auto& e : join) if (e) hipEventDestroy(e);
    }
};

struct cycle_t {
    int ng = 0;
    std::vector<std::array<int64_t, 3>> dims;  // (nx, ny, nz) per grid
    std::vector<int64_t> off;
    int64_t total = 0;
    bool dbl = false;
    int effort = 0;
    hipStream_t stream = nullptr;
    hipStream_t side[kLanes - 1] = {};
    cycle_stage st[2][3];  // [fwd/bwd][stage]
    // graph of fwd+bwd for one set of pointers
    hipStream_t gstream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    void* gptr[4] = {nullptr, nullptr, nullptr, nullptr};
    ~cycle_t() {
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        if (gstream) hipStreamDestroy(gstream);
        for (auto& x : side) if (x) hipStreamDestroy(x);
        if (ev_a) hipEventDestroy(ev_a);
        if (ev_b) hipEventDestroy(ev_b);
    }
};

static stage_spec spec_of(int dir, int s, const std::array<int64_t, 3>& g) {
    const int64_t nx = g[0], ny = g[1], nz = g[2];
    if (dir == 0) {
        if (s == 0) return {nx, ny * nz, 1, nx, 0, 1, nx, 0, 1};                 // FFT_X
        if (s == 1) return {ny, nx, nz, 1, nx * ny, nx, ny * nz, ny, 1};          // X->Y + FFT_Y
        return {nz, ny, nx, 1, ny * nz, ny, nz * nx, nz, 1};                      // Y->Z + FFT_Z
    }
    // backward in the order whose strided reads stay one row apart (same result on one grid)
    if (s == 0) return {nz, nx * ny, 1, nz, 0, 1, nz, 0, 1};                      // FFT_Z
    if (s == 1) return {nx, nz, ny, 1, nz * nx, nz, nx * ny, nx, 1};              // Z->X + FFT_X
    return {ny, nx, nz, 1, nx * ny, nx, 1, nx * ny, nx};                          // FFT_Y in place
}

static std::string tile_type(const fft_config& c) {
    std::string s = "dtfft::fft_tile<";
    s += c.dbl ? "double" : "float";
    s += ", dtfft::sched<" + std::to_string(c.n);
    for (int r : c.radices) s += ", " + std::to_string(r);
    s += ">, " + std::to_string(c.tpl) + ", " + std::to_string(c.lpb) + ", " + (c.in_cols ? "true" : "false") +
         ", " + (c.out_cols ? "true" : "false") + ", " + std::to_string(c.pad) + ">";
    return s;
}

static int32_t launch_group(const length_group& grp, int inv, const void* in, void* out, hipStream_t st);

static std::string group_expr(const fft_config& cf, bool dbl) {
    return std::string("batched_fft_kernel<") + (dbl ? "double" : "float") + ", " + std::to_string(cf.tpl * cf.lpb) +
           ", " + tile_type(cf) + ">";
}

// One launch of a stage for the grids `members`, all of one axis length, with tile shape cf.
static int32_t make_group(const cycle_t* cy, const std::vector<stage_spec>& specs, const std::vector<int>& members,
                          const fft_config& cf, std::unique_ptr<length_group>& out) {
    auto grp = std::make_unique<length_group>();
    grp->cfg = cf;
    grp->block = (unsigned)(cf.tpl * cf.lpb);
    TRY(rtc_get("fft", group_expr(cf, cy->dbl), &grp->k));
    TRY(get_twiddles(cf.n, cf.radices, cy->dbl, &grp->tw));
    std::vector<grid_stage_h> gs;
    std::vector<int> flat;
    for (int g : members) {
        const stage_spec& q = specs[g];
        grid_stage_h h{};
        h.d = make_lines(q.nl, q.nb, q.in_ls, q.in_bs, q.in_ns, q.out_ls, q.out_bs, q.out_ns, cf.lpb);
        h.in_off = cy->off[g];
        h.out_off = cy->off[g];
        h.tw = grp->tw->ptr;
        h.cls = 0;
        const int idx = (int)gs.size();
        for (int64_t t = 0; t < h.d.ntiles; t++) { flat.push_back(idx); flat.push_back((int)t); }
        grp->work += q.n * q.nl * q.nb;
        gs.push_back(h);
    }
    grp->nitems = (int)(flat.size() / 2);
    if (!compile_only()) {
        HIP_TRY(hipMalloc(&grp->d_gs, sizeof(grid_stage_h) * gs.size()));
        HIP_TRY(hipMemcpy(grp->d_gs, gs.data(), sizeof(grid_stage_h) * gs.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&grp->d_items, sizeof(int) * flat.size()));
        HIP_TRY(hipMemcpy(grp->d_items, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice));
        int per_cu = 1;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grp->k->fn, (int)grp->block, 0) !=
                hipSuccess || per_cu < 1)
            per_cu = 1;
        const int64_t resident = (int64_t)per_cu * std::max(1, device().num_cus);
        grp->grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grp->nitems, resident));
    }
    out = std::move(grp);
    return DTFFT_HIP_SUCCESS;
}

// Candidate tile shapes of a group for the autotuner, the default first: radix schedules of
// the length (radix_schedules), threads per line of one or two butterflies of the largest
// radix per pass, and the default's lines per block, halved and doubled.
static std::vector<fft_config> group_candidates(int n, bool dbl, bool in_cols, bool out_cols, int effort) {
    std::vector<fft_config> out;
    const fft_config base = default_fft_config(n, dbl, in_cols, out_cols, in_cols && out_cols);
    out.push_back(base);
    const auto scheds = radix_schedules(n, effort >= DTFFT_HIP_EXHAUSTIVE ? 6 : 3);
    for (const auto& r : scheds) {
        const int rmax = *std::max_element(r.begin(), r.end());
        for (int t : {n / rmax, n / rmax / 2})
            for (int l : {base.lpb, base.lpb * 2, base.lpb / 2}) {
                if (t < 1 || l < 1 || t * l > 1024 || t * l < 64) continue;
                fft_config c = base;
                c.radices = r;
                c.tpl = t;
                c.lpb = l;
                if (fft_lds_bytes(c) > 64 * 1024) continue;  // leave room for several blocks per CU
                bool dup = false;
                for (auto& o : out) dup |= (o.radices == c.radices && o.tpl == c.tpl && o.lpb == c.lpb);
                if (!dup) out.push_back(c);
            }
    }
    if (effort < DTFFT_HIP_EXHAUSTIVE && out.size() > 8) out.resize(8);
    return out;
}

static uint64_t fnv1a_ints(const std::vector<int64_t>& v) {
    uint64_t h = 1469598103934665603ull;
    for (int64_t x : v)
        for (int b = 0; b < 8; b++) { h ^= (uint64_t)((x >> (8 * b)) & 0xff); h *= 1099511628211ull; }
    return h;
}

// At effort >= MEASURE: time the candidate shapes of one group on scratch buffers and keep the
// fastest (the project's replacement of dtfft_nvrtc_block_optimizer for the batched cycle).
// The choice is stored as a tuning record keyed by the group's length, layout and lines.
static int32_t tune_group(cycle_t* cy, const std::vector<stage_spec>& specs, const std::vector<int>& members,
                          bool in_cols, bool out_cols, void* scratch_in, void* scratch_out, fft_config& best) {
    const int n = (int)specs[members[0]].n;
    std::vector<int64_t> sig;
    for (int g : members) {
        const stage_spec& q = specs[g];
        sig.insert(sig.end(), {q.nl, q.nb, q.in_ls, q.in_bs, q.in_ns, q.out_ls, q.out_bs, q.out_ns});
    }
    char kb[160];
    std::snprintf(kb, sizeof kb, "cyc|%d|%c|%d%d|%zu|%016llx", n, cy->dbl ? 'd' : 'f', (int)in_cols, (int)out_cols,
                  members.size(), (unsigned long long)fnv1a_ints(sig));
    const std::string key(kb);
    best = default_fft_config(n, cy->dbl, in_cols, out_cols, in_cols && out_cols);
    std::string rec;
    if (tune_lookup(key, rec)) {
        int t, l, p;
        char rads[128] = {0};
        std::vector<int> r;
        if (std::sscanf(rec.c_str(), "%d %d %d %127s", &t, &l, &p, rads) == 4 && parse_radices(rads, n, r)) {
            fft_config c = best;
            c.tpl = t; c.lpb = l; c.pad = p; c.radices = r;
            if (t * l >= 1 && t * l <= 1024 && fft_lds_bytes(c) <= 160 * 1024) best = c;
        }
        return DTFFT_HIP_SUCCESS;
    }
    if (cy->effort < DTFFT_HIP_MEASURE || compile_only() || !scratch_in) return DTFFT_HIP_SUCCESS;
    auto cands = group_candidates(n, cy->dbl, in_cols, out_cols, cy->effort);
    // compile the candidates side by side (hipRTC programs are independent)
    {
        std::vector<std::thread> pool;
        std::atomic<size_t> next{0};
        const unsigned nt = std::max(1u, std::min<unsigned>(8, std::thread::hardware_concurrency()));
        for (unsigned i = 0; i < nt; i++)
            pool.emplace_back([&]() {
                for (size_t j; (j = next++) < cands.size();) {
                    const rtc_kernel* k = nullptr;
                    rtc_get("fft", group_expr(cands[j], cy->dbl), &k);
                }
            });
        for (auto& th : pool) th.join();
    }
    double best_ms = 1e30;
    std::vector<std::unique_ptr<length_group>> grps;
    std::vector<fft_config> kept;
    for (const auto& c : cands) {
        std::unique_ptr<length_group> grp;
        if (make_group(cy, specs, members, c, grp) != DTFFT_HIP_SUCCESS) continue;
        grps.push_back(std::move(grp));
        kept.push_back(c);
    }
    std::vector<std::function<int32_t()>> runs;
    for (auto& g : grps) {
        const length_group* gp = g.get();
        runs.push_back([gp, scratch_in, scratch_out, cy]() {
            return launch_group(*gp, 0, scratch_in, scratch_out, cy->stream);
        });
    }
    const std::vector<double> ms = time_candidates(runs, cy->stream);
    for (size_t i = 0; i < kept.size(); i++)
        if (ms[i] < best_ms) { best_ms = ms[i]; best = kept[i]; }
    std::string r;
    for (size_t i = 0; i < best.radices.size(); i++) r += (i ? "," : "") + std::to_string(best.radices[i]);
    char buf[160];
    std::snprintf(buf, sizeof buf, "%d %d %d %s", best.tpl, best.lpb, best.pad, r.c_str());
    tune_store(key, buf);
    return DTFFT_HIP_SUCCESS;
}

static int32_t build_stage(cycle_t* cy, int dir, int s, void* scratch_in, void* scratch_out) {
    cycle_stage& cs = cy->st[dir][s];
    std::vector<stage_spec> specs(cy->ng);
    std::vector<int> group_of(cy->ng, -1);
    std::vector<int64_t> lengths;
    for (int g = 0; g < cy->ng; g++) {
        specs[g] = spec_of(dir, s, cy->dims[g]);
        const stage_spec& sp = specs[g];
        std::vector<int> rad;
        if (!factorize((int)sp.n, rad) || sp.n < 2) {
            set_last_error("batched cycle: unsupported length " + std::to_string(sp.n));
            return DTFFT_HIP_ERROR_UNSUPPORTED_SIZE;
        }
        auto it = std::find(lengths.begin(), lengths.end(), sp.n);
        if (it == lengths.end()) { lengths.push_back(sp.n); it = lengths.end() - 1; }
        group_of[g] = (int)(it - lengths.begin());
    }
    // every grid of a stage has the same layout kind (rows or columns on each side)
    std::vector<std::unique_ptr<length_group>> groups;
    for (size_t c = 0; c < lengths.size(); c++) {
        std::vector<int> members;
        for (int g = 0; g < cy->ng; g++) if (group_of[g] == (int)c) members.push_back(g);
        const stage_spec& sp = specs[members[0]];
        const bool in_cols = sp.in_ns != 1 && sp.in_ls == 1;
        const bool out_cols = sp.out_ns != 1 && sp.out_ls == 1;
        fft_config cf;
        TRY(tune_group(cy, specs, members, in_cols, out_cols, scratch_in, scratch_out, cf));
        std::unique_ptr<length_group> grp;
        TRY(make_group(cy, specs, members, cf, grp));
        groups.push_back(std::move(grp));
    }
    std::stable_sort(groups.begin(), groups.end(),
                     [](const std::unique_ptr<length_group>& a, const std::unique_ptr<length_group>& b) {
                         return a->work > b->work;
                     });
    cs.groups = std::move(groups);
    if (!compile_only() && cs.groups.size() > 1) {
        HIP_TRY(hipEventCreateWithFlags(&cs.fork, hipEventDisableTiming));
        for (auto& e : cs.join) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t launch_group(const length_group& grp, int inv, const void* in, void* out, hipStream_t st) {
    void* gsp = grp.d_gs;
    void* itp = grp.d_items;
    int nitems = grp.nitems;
    void* args[] = {(void*)&in, &out, &gsp, &itp, &nitems, &inv};
    return launch(grp.k, grp.grid, grp.block, st, args);
}

// One stage: group i runs on lane i % kLanes (lane 0 is the caller's stream); the side
// lanes wait for the stage's fork event and the caller's stream waits for their joins.
static int32_t run_stage(const cycle_t* cy, int dir, int s, const void* in, void* out, hipStream_t st) {
    const cycle_stage& cs = cy->st[dir][s];
    const int inv = dir == 1 ? 1 : 0;
    const int ng = (int)cs.groups.size();
    if (ng == 1) return launch_group(*cs.groups[0], inv, in, out, st);
    const int lanes = std::min(ng, kLanes);
    HIP_TRY(hipEventRecord(cs.fork, st));
    for (int l = 1; l < lanes; l++) HIP_TRY(hipStreamWaitEvent(cy->side[l - 1], cs.fork, 0));
    for (int i = 0; i < ng; i++) {
        const int l = i % lanes;
        TRY(launch_group(*cs.groups[i], inv, in, out, l == 0 ? st : cy->side[l - 1]));
    }
    for (int l = 1; l < lanes; l++) {
        HIP_TRY(hipEventRecord(cs.join[l], cy->side[l - 1]));
        HIP_TRY(hipStreamWaitEvent(st, cs.join[l], 0));
    }
    return DTFFT_HIP_SUCCESS;
}

static int32_t run_dir(const cycle_t* cy, int dir, void* in, void* out, void* aux, hipStream_t st) {
    if (dir == 0) {  // in -> out -> aux -> out
        TRY(run_stage(cy, 0, 0, in, out, st));
        TRY(run_stage(cy, 0, 1, out, aux, st));
        return run_stage(cy, 0, 2, aux, out, st);
    }
    // in -> aux -> out, then in place on out
    TRY(run_stage(cy, 1, 0, in, aux, st));
    TRY(run_stage(cy, 1, 1, aux, out, st));
    return run_stage(cy, 1, 2, out, out, st);
}

}  // namespace dtfft_hip

using namespace dtfft_hip;

extern "C" {

int32_t dtfft_hip_cycle_create(int32_t n_grids, const int32_t* dims, int32_t precision, int32_t effort, void* stream,
                               dtfft_hip_cycle_t* cycle) {
    if (!cycle || !dims) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    *cycle = nullptr;
    if (n_grids <= 0) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (precision != DTFFT_HIP_SINGLE && precision != DTFFT_HIP_DOUBLE) return DTFFT_HIP_ERROR_INVALID_PRECISION;
    if (effort < DTFFT_HIP_ESTIMATE || effort > DTFFT_HIP_EXHAUSTIVE) return DTFFT_HIP_ERROR_INVALID_EFFORT_FLAG;
    auto cy = std::make_unique<cycle_t>();
    cy->ng = n_grids;
    cy->dbl = precision == DTFFT_HIP_DOUBLE;
    cy->effort = effort;
    cy->stream = (hipStream_t)stream;
    int64_t off = 0;
    for (int g = 0; g < n_grids; g++) {
        // reverse order per grid, as dtfft_create_plan_c2c
        std::array<int64_t, 3> d = {dims[3 * g + 2], dims[3 * g + 1], dims[3 * g + 0]};
        for (auto v : d) if (v <= 0) return DTFFT_HIP_ERROR_INVALID_DIMENSION_SIZE;
        const int64_t n = d[0] * d[1] * d[2];
        if (n > INT32_MAX) return DTFFT_HIP_ERROR_INDEX_OVERFLOW;
        cy->dims.push_back(d);
        cy->off.push_back(off);
        off += (n + 63) / 64 * 64;  // keep every grid 512-byte aligned
    }
    cy->total = off;
    if (!compile_only())
        for (auto& x : cy->side) HIP_TRY(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    // scratch buffers for timing candidate tile shapes (effort >= MEASURE), freed after creation
    void *si = nullptr, *so = nullptr;
    if (effort >= DTFFT_HIP_MEASURE && !compile_only()) {
        const size_t bytes = (size_t)cy->total * (cy->dbl ? 16 : 8);
        HIP_TRY(hipMalloc(&si, bytes));
        if (hipMalloc(&so, bytes) != hipSuccess) { hipFree(si); return DTFFT_HIP_ERROR_ALLOC_FAILED; }
        if (fill_tuning_scratch(si, bytes, cy->dbl, cy->stream) != DTFFT_HIP_SUCCESS) {
            hipFree(si);
            hipFree(so);
            return DTFFT_HIP_ERROR_HIP_RUNTIME;
        }
    }
    int32_t rc = DTFFT_HIP_SUCCESS;
    for (int dir = 0; dir < 2 && rc == DTFFT_HIP_SUCCESS; dir++)
        for (int s = 0; s < 3 && rc == DTFFT_HIP_SUCCESS; s++) rc = build_stage(cy.get(), dir, s, si, so);
    if (si) { hipStreamSynchronize(cy->stream); hipFree(si); hipFree(so); }
    TRY(rc);
    *cycle = cy.release();
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_cycle_get_layout(dtfft_hip_cycle_t cycle, int64_t* offsets, size_t* alloc_size) {
    auto* cy = (cycle_t*)cycle;
    if (!cy) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (offsets) for (int g = 0; g < cy->ng; g++) offsets[g] = cy->off[g];
    if (alloc_size) *alloc_size = (size_t)cy->total;
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_cycle_execute(dtfft_hip_cycle_t cycle, void* in, void* out, void* aux, int32_t execute_type) {
    auto* cy = (cycle_t*)cycle;
    if (!cy) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (execute_type != DTFFT_HIP_EXECUTE_FORWARD && execute_type != DTFFT_HIP_EXECUTE_BACKWARD)
        return DTFFT_HIP_ERROR_INVALID_EXECUTE_TYPE;
    if (!in || !out || !aux) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (aux == in || aux == out) return DTFFT_HIP_ERROR_INVALID_AUX;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    return run_dir(cy, execute_type == DTFFT_HIP_EXECUTE_FORWARD ? 0 : 1, in, out, aux, cy->stream);
}

int32_t dtfft_hip_cycle_execute_graph(dtfft_hip_cycle_t cycle, void* in, void* out, void* aux, void* in2) {
    auto* cy = (cycle_t*)cycle;
    if (!cy) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    if (!in || !out || !aux || !in2) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (aux == in || aux == out || aux == in2 || out == in || out == in2) return DTFFT_HIP_ERROR_INVALID_AUX;
    if (compile_only()) return DTFFT_HIP_ERROR_NO_DEVICE;
    void* p[4] = {in, out, aux, in2};
    if (!cy->gexec || !std::equal(p, p + 4, cy->gptr)) {
        if (!cy->gstream) {
            HIP_TRY(hipStreamCreateWithFlags(&cy->gstream, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&cy->ev_a, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&cy->ev_b, hipEventDisableTiming));
        }
        if (cy->gexec) { hipGraphExecDestroy(cy->gexec); cy->gexec = nullptr; }
        if (cy->graph) { hipGraphDestroy(cy->graph); cy->graph = nullptr; }
        HIP_TRY(hipStreamBeginCapture(cy->gstream, hipStreamCaptureModeThreadLocal));
        int32_t rc = run_dir(cy, 0, in, out, aux, cy->gstream);
        if (rc == DTFFT_HIP_SUCCESS) rc = run_dir(cy, 1, out, in2, aux, cy->gstream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cy->gstream, &g);
        if (rc != DTFFT_HIP_SUCCESS) { if (g) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) { set_last_error("hipStreamEndCapture failed"); return DTFFT_HIP_ERROR_HIP_RUNTIME; }
        cy->graph = g;
        HIP_TRY(hipGraphInstantiate(&cy->gexec, cy->graph, nullptr, nullptr, 0));
        std::copy(p, p + 4, cy->gptr);
    }
    // order the replay after the caller's stream and the caller's stream after it
    HIP_TRY(hipEventRecord(cy->ev_a, cy->stream));
    HIP_TRY(hipStreamWaitEvent(cy->gstream, cy->ev_a, 0));
    HIP_TRY(hipGraphLaunch(cy->gexec, cy->gstream));
    HIP_TRY(hipEventRecord(cy->ev_b, cy->gstream));
    HIP_TRY(hipStreamWaitEvent(cy->stream, cy->ev_b, 0));
    return DTFFT_HIP_SUCCESS;
}

int32_t dtfft_hip_cycle_destroy(dtfft_hip_cycle_t* cycle) {
    if (!cycle) return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    if (!*cycle) return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
    delete (cycle_t*)*cycle;
    *cycle = nullptr;
    return DTFFT_HIP_SUCCESS;
}

}  // extern "C"
