// C-ABI implementation of libmbe.so: graph handles, generators and the HIP
// engine that launches the persistent enumeration kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include <unistd.h>

#include "../../include/mbe.h"
#include "engine.h"
#include "graph.h"

namespace mbe {
hipError_t launch_mbe(const KParams& kp, uint32_t workers, hipStream_t stream);
hipError_t launch_dq_init(unsigned long long* dq, uint32_t n_slots, hipStream_t stream);
hipError_t set_lds_limit(uint32_t lds_words);
}  // namespace mbe

struct mbe_graph {
    mbe::Graph g;
};

namespace {

thread_local std::string g_err;
thread_local int g_hip_err = 0;

mbe_status fail(mbe_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

#define HIPC(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            g_hip_err = (int)e_;                                                           \
            return fail(MBE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
        }                                                                                  \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Device buffer that grows on demand and is reused across solves.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    template <class T>
    T* as() const { return (T*)p; }
};

// Process-wide engine state for device 0: worker resources are allocated once
// and reused by every solve (a solve is serialised by the mutex).
struct Engine {
    std::mutex mu;
    bool init = false;
    int cus = 0;
    size_t lds_per_cu = 160 * 1024;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf arena, mark, cpos, ctl, graphs, data, ovf, slots, dq;
    uint32_t lds_set = 0;
    unsigned long long* progress = nullptr;  // host-mapped, 2 per worker
    size_t progress_slots = 0;
    unsigned long long* abort_flag = nullptr;  // host-mapped watchdog flag

    mbe_status ensure_init() {
        if (init) return MBE_OK;
        int n = 0;
        HIPC(hipGetDeviceCount(&n));
        if (n < 1) return fail(MBE_EHIP, "no HIP device");
        HIPC(hipSetDevice(0));
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, 0));
        cus = prop.multiProcessorCount;
        if (prop.maxSharedMemoryPerMultiProcessor) lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
        HIPC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPC(hipEventCreate(&ev0));
        HIPC(hipEventCreate(&ev1));
        HIPC(hipHostMalloc((void**)&abort_flag, 64, hipHostMallocMapped | hipHostMallocCoherent));
        init = true;
        return MBE_OK;
    }
};

Engine& engine() {
    static Engine* e = new Engine();  // never destroyed: outlives HIP teardown
    return *e;
}

struct Tune {
    uint32_t workers, lds_words, list_cutoff, donate_depth, compact_ratio, donate_interval;
    uint64_t arena_words;
    int side;
    uint32_t flags;
    uint32_t root_begin, root_end, root_stride;
    uint64_t max_nodes;
    double time_limit_ms;
};

Tune resolve(const mbe_tuning* t, int cus, size_t lds_per_cu) {
    Tune r;
    r.lds_words = (t && t->lds_words) ? t->lds_words : 4096;  // 32 KiB: 5 workers per CU
    r.lds_words = std::max<uint32_t>(r.lds_words, mbe::kLdsHeaderWords + 64);
    r.lds_words = std::min<uint32_t>(r.lds_words, (uint32_t)(lds_per_cu / 8));
    // default occupancy: as many single-wave workers per CU as the LDS holds
    uint32_t fit = (uint32_t)std::max<size_t>(1, lds_per_cu / ((size_t)r.lds_words * 8));
    uint32_t wpc = (t && t->waves_per_cu) ? t->waves_per_cu : std::min<uint32_t>(fit, 4);
    r.workers = (t && t->workers) ? t->workers : (uint32_t)cus * wpc;
    r.arena_words = (t && t->arena_words) ? t->arena_words : (1u << 18);
    r.list_cutoff = (t && t->list_cutoff) ? t->list_cutoff : 0;
    r.donate_depth = (t && t->donate_depth) ? t->donate_depth : 0;
    r.compact_ratio = (t && t->compact_ratio) ? (t->compact_ratio == UINT32_MAX ? 0 : t->compact_ratio) : 2;
    r.side = t ? t->side : -1;
    r.flags = t ? t->flags : 0;
    r.root_begin = t ? t->root_begin : 0;
    r.root_end = t ? t->root_end : 0;
    r.root_stride = (t && t->root_stride) ? t->root_stride : 1;
    r.max_nodes = t ? t->max_nodes : 0;
    r.time_limit_ms = t ? t->time_limit_ms : 0;
    r.donate_interval = (t && t->donate_interval) ? t->donate_interval : 64;
    return r;
}

// Upload prepared graphs into one device buffer and build the descriptors.
mbe_status upload(Engine& E, const std::vector<mbe::Prepared>& ps, std::vector<mbe::DevGraph>& dg,
                  uint64_t& total_tasks) {
    // one padding rule for sizing and placement: every array gets its bytes
    // plus 8 spare, rounded up to 256
    auto slot = [](size_t n) { return (n + 8 + 255) & ~(size_t)255; };
    size_t bytes = 0;
    for (auto& p : ps) {
        bytes += slot(p.offR.size() * 4) + slot(p.adjR.size() * 4) + slot(p.offL.size() * 4) +
                 slot(p.adjL.size() * 4) + slot(p.hR.size() * 8) + slot(p.hL.size() * 8);
    }
    HIPC(E.data.ensure(bytes));
    std::vector<char> host(bytes);
    size_t o = 0;
    dg.resize(ps.size());
    total_tasks = 0;
    char* dbase = E.data.as<char>();
    auto put = [&](const void* src, size_t n) {
        size_t at = o;
        if (n) memcpy(host.data() + at, src, n);
        o += slot(n);
        return (const void*)(dbase + at);
    };
    for (size_t i = 0; i < ps.size(); i++) {
        auto& p = ps[i];
        auto& d = dg[i];
        d.offR = (const uint32_t*)put(p.offR.data(), p.offR.size() * 4);
        d.adjR = (const uint32_t*)put(p.adjR.data(), p.adjR.size() * 4);
        d.offL = (const uint32_t*)put(p.offL.data(), p.offL.size() * 4);
        d.adjL = (const uint32_t*)put(p.adjL.data(), p.adjL.size() * 4);
        d.hR = (const uint64_t*)put(p.hR.data(), p.hR.size() * 8);
        d.hL = (const uint64_t*)put(p.hL.data(), p.hL.size() * 8);
        d.nR = p.nR;
        d.nL = p.nL;
        d.n_roots = p.n_roots;
        d.r_is_u = p.r_is_u ? 1 : 0;
        d.task_base = total_tasks;
        total_tasks += p.n_roots;
    }
    if (o != bytes) return fail(MBE_EINVAL, "upload layout mismatch");
    HIPC(hipMemcpyAsync(E.data.p, host.data(), o, hipMemcpyHostToDevice, E.stream));
    HIPC(E.graphs.ensure(sizeof(mbe::DevGraph) * std::max<size_t>(1, dg.size())));
    HIPC(hipMemcpyAsync(E.graphs.p, dg.data(), sizeof(mbe::DevGraph) * dg.size(),
                        hipMemcpyHostToDevice, E.stream));
    HIPC(hipStreamSynchronize(E.stream));
    return MBE_OK;
}

// Wait for the launch; with MBE_PROGRESS=<seconds> set, print the roots and
// nodes the workers have finished (read from host-mapped slots) to stderr.
hipError_t wait_with_progress(Engine& E, uint32_t nw, uint64_t n_tasks, int pass,
                              double deadline_ms) {
    const char* env = getenv("MBE_PROGRESS");
    const double every = env ? atof(env) * 1000.0 : 0.0;
    const double t0 = now_ms();
    double last = t0;
    for (;;) {
        hipError_t q = hipEventQuery(E.ev1);
        if (q != hipErrorNotReady) return q == hipSuccess ? hipEventSynchronize(E.ev1) : q;
        usleep(2000);
        if (deadline_ms > 0 && now_ms() > deadline_ms)
            __atomic_store_n(E.abort_flag, 1ull, __ATOMIC_SEQ_CST);
        if (every > 0 && now_ms() - last >= every) {
            last = now_ms();
            unsigned long long roots = 0, nodes = 0;
            for (uint32_t w = 0; w < nw; w++) {
                roots += ((volatile unsigned long long*)E.progress)[2 * w];
                nodes += ((volatile unsigned long long*)E.progress)[2 * w + 1];
            }
            fprintf(stderr, "[mbe] pass %d: %.1f s, %llu/%llu roots, %llu nodes\n", pass,
                    (last - t0) / 1000.0, roots, (unsigned long long)n_tasks, nodes);
        }
    }
}

// Run every root task of the uploaded graphs; roots whose local problem does
// not fit a worker arena are retried with fewer workers and larger arenas.
mbe_status run_tasks(Engine& E, const std::vector<mbe::Prepared>& ps,
                     const std::vector<mbe::DevGraph>& dg, uint64_t total_tasks, int variant,
                     const Tune& tu, std::vector<uint64_t>& res, mbe_result& out) {
    uint32_t max_nr = 1;
    for (auto& p : ps) max_nr = std::max(max_nr, p.nR);
    const uint64_t mark_words = (max_nr + 63) / 64 + 1;
    const uint64_t cpos_len = max_nr + 1;
    const size_t ng = dg.size();
    // control block: cursor, overflow count, stats, results
    const size_t ctl_words = 2 + mbe::kStatCount + mbe::kSecCount + 2 * ng;
    HIPC(E.ctl.ensure(ctl_words * 8));
    HIPC(hipMemsetAsync(E.ctl.p, 0, ctl_words * 8, E.stream));
    unsigned long long* cursor = E.ctl.as<unsigned long long>();
    unsigned long long* ovf_count = cursor + 1;
    unsigned long long* stats = cursor + 2;
    unsigned long long* secs = stats + mbe::kStatCount;
    unsigned long long* results = secs + mbe::kSecCount;

    if (E.lds_set < tu.lds_words) {
        HIPC(mbe::set_lds_limit(tu.lds_words));
        E.lds_set = tu.lds_words;
    }

    uint64_t n_tasks = total_tasks;
    uint64_t arena_words = tu.arena_words;
    uint32_t workers = tu.workers;
    std::vector<uint64_t> task_list;  // empty = range mode
    if (ng == 1 && (tu.root_begin || tu.root_end || tu.root_stride > 1)) {
        const uint64_t end = tu.root_end ? std::min<uint64_t>(tu.root_end, total_tasks) : total_tasks;
        for (uint64_t r = tu.root_begin; r < end; r += tu.root_stride) task_list.push_back(r);
        n_tasks = task_list.size();
    }
    float kernel_ms = 0;
    uint32_t launches = 0;
    const uint64_t ovf_cap = std::max<uint64_t>(total_tasks, 1);
    HIPC(E.ovf.ensure(ovf_cap * 8 * 2));
    uint64_t* ovf_list = E.ovf.as<uint64_t>();
    uint64_t* task_dev = ovf_list + ovf_cap;
    const uint64_t budget_words = (uint64_t)48 << 30 >> 3;  // 48 GiB of arenas
    const double deadline = tu.time_limit_ms > 0 ? now_ms() + tu.time_limit_ms : 0;
    __atomic_store_n(E.abort_flag, 0ull, __ATOMIC_SEQ_CST);
    for (int pass = 0; n_tasks > 0; pass++) {
        if (pass > 6) return fail(MBE_ENOMEM, "root subproblem does not fit any arena");
        HIPC(E.arena.ensure(arena_words * workers * 8));
        HIPC(E.mark.ensure(mark_words * workers * 8));
        HIPC(E.cpos.ensure(cpos_len * workers * 4));
        if (pass == 0) HIPC(hipMemsetAsync(E.mark.p, 0, mark_words * workers * 8, E.stream));
        if (!task_list.empty()) {
            HIPC(hipMemcpyAsync(task_dev, task_list.data(), task_list.size() * 8,
                                hipMemcpyHostToDevice, E.stream));
        }
        HIPC(hipMemsetAsync(cursor, 0, 16, E.stream));
        const uint32_t n_slots = 1024;
        // a slot holds a compacted level of any problem that fits an arena,
        // up to 4 MiB per slot (1024 slots)
        const uint64_t slot_words = std::min<uint64_t>(std::max<uint64_t>(4096, arena_words), 1u << 19);
        HIPC(E.slots.ensure((size_t)n_slots * slot_words * 8));
        HIPC(E.dq.ensure((4 + (size_t)n_slots) * 8));
        HIPC(mbe::launch_dq_init(E.dq.as<unsigned long long>(), n_slots, E.stream));
        mbe::KParams kp{};
        kp.graphs = E.graphs.as<mbe::DevGraph>();
        kp.n_graphs = (uint32_t)ng;
        kp.variant = (uint32_t)variant;
        kp.n_tasks = n_tasks;
        kp.task_list = task_list.empty() ? nullptr : task_dev;
        kp.cursor = cursor;
        kp.results = results;
        kp.stats = stats;
        kp.arena = E.arena.as<uint64_t>();
        kp.arena_words = arena_words;
        kp.mark = E.mark.as<uint64_t>();
        kp.mark_words = mark_words;
        kp.cpos = E.cpos.as<uint32_t>();
        kp.cpos_len = cpos_len;
        kp.lds_words = tu.lds_words;
        kp.list_cutoff = tu.list_cutoff;
        kp.compact_ratio = tu.compact_ratio;
        kp.sections = (tu.flags & MBE_FLAG_SECTIONS) ? secs : nullptr;
        HIPC(hipHostGetDevicePointer((void**)&kp.abort, E.abort_flag, 0));
        kp.node_count = stats + mbe::kStatNodeBudget;
        kp.node_budget = tu.max_nodes;
        kp.overflow_count = ovf_count;
        kp.overflow_list = ovf_list;
        kp.overflow_cap = ovf_cap;
        kp.donate_depth = tu.donate_depth;
        kp.donate_interval = tu.donate_interval;
        kp.slots = E.slots.as<uint64_t>();
        kp.slot_words = slot_words;
        kp.n_slots = n_slots;
        kp.dq = E.dq.as<unsigned long long>();
        // all workers start: those without a root task steal donated subtrees
        uint32_t nw = workers;
        if (E.progress_slots < 2 * (size_t)nw) {
            if (E.progress) HIPC(hipHostFree(E.progress));
            E.progress = nullptr;
            HIPC(hipHostMalloc((void**)&E.progress, 2 * (size_t)nw * 8, hipHostMallocMapped | hipHostMallocCoherent));
            E.progress_slots = 2 * (size_t)nw;
        }
        memset(E.progress, 0, 2 * (size_t)nw * 8);
        unsigned long long* dprog = nullptr;
        HIPC(hipHostGetDevicePointer((void**)&dprog, E.progress, 0));
        kp.progress = dprog;
        HIPC(hipEventRecord(E.ev0, E.stream));
        HIPC(mbe::launch_mbe(kp, nw, E.stream));
        HIPC(hipEventRecord(E.ev1, E.stream));
        HIPC(wait_with_progress(E, nw, n_tasks, pass, deadline));
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, E.ev0, E.ev1));
        kernel_ms += ms;
        launches++;
        out.workers = std::max(out.workers, nw);
        unsigned long long novf = 0;
        HIPC(hipMemcpy(&novf, ovf_count, 8, hipMemcpyDeviceToHost));
        if (__atomic_load_n(E.abort_flag, __ATOMIC_SEQ_CST)) break;
        if (!novf) break;
        task_list.resize(novf);
        HIPC(hipMemcpy(task_list.data(), ovf_list, novf * 8, hipMemcpyDeviceToHost));
        n_tasks = novf;
        arena_words *= 16;
        workers = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(workers, budget_words / arena_words));
        HIPC(hipMemsetAsync(ovf_count, 0, 8, E.stream));
    }
    std::vector<unsigned long long> st(mbe::kStatCount), rr(2 * ng), sc(mbe::kSecCount);
    HIPC(hipMemcpy(st.data(), stats, st.size() * 8, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(sc.data(), secs, sc.size() * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < mbe::kSecCount; i++) out.sections[i] = sc[i];
    HIPC(hipMemcpy(rr.data(), results, rr.size() * 8, hipMemcpyDeviceToHost));
    res.assign(rr.begin(), rr.end());
    out.kernel_ms = kernel_ms;
    out.nodes = st[mbe::kStatNodes];
    out.donated = st[mbe::kStatDonated];
    out.launches = launches;
    out.words = st[mbe::kStatWords];
    out.build_words = st[mbe::kStatBuildWords];
    if (__atomic_load_n(E.abort_flag, __ATOMIC_SEQ_CST))
        return fail(MBE_ETIMEOUT, "watchdog: node budget or time limit reached; results partial");
    if (st[mbe::kStatDropped]) return fail(MBE_ENOMEM, "donated subtree did not fit a worker arena");
    return MBE_OK;
}

}  // namespace

extern "C" {

const char* mbe_status_str(mbe_status s) {
    switch (s) {
        case MBE_OK: return "ok";
        case MBE_EINVAL: return "invalid argument";
        case MBE_ENOMEM: return "out of memory";
        case MBE_EHIP: return "HIP error";
        case MBE_ETIMEOUT: return "timeout";
        case MBE_EOVERBUDGET: return "over admission budget";
        case MBE_EIO: return "I/O error";
        case MBE_EBUSY: return "busy";
        case MBE_ENOTREADY: return "not ready";
    }
    return "unknown status";
}

int mbe_last_hip_error(void) { return g_hip_err; }
const char* mbe_last_error(void) { return g_err.c_str(); }

mbe_status mbe_graph_load_bi(const char* path, mbe_graph** out) {
    if (!path || !out) return fail(MBE_EINVAL, "null argument");
    auto h = std::make_unique<mbe_graph>();
    std::string err;
    if (!mbe::read_bi(path, h->g, &err)) return fail(err.rfind("cannot open", 0) == 0 ? MBE_EIO : MBE_EINVAL, err);
    *out = h.release();
    return MBE_OK;
}

mbe_status mbe_graph_write_bi(const mbe_graph* g, const char* path) {
    if (!g || !path) return fail(MBE_EINVAL, "null argument");
    std::string err;
    if (!mbe::write_bi(path, g->g, &err)) return fail(MBE_EIO, err);
    return MBE_OK;
}

mbe_status mbe_graph_from_edge_list(const char* path, int skip_words, int skip_per_edge,
                                    mbe_graph** out) {
    if (!path || !out || skip_words < 0 || skip_per_edge < 0) return fail(MBE_EINVAL, "bad argument");
    auto h = std::make_unique<mbe_graph>();
    std::string err;
    if (!mbe::edge_list_to_graph(path, skip_words, skip_per_edge, h->g, &err)) return fail(MBE_EIO, err);
    *out = h.release();
    return MBE_OK;
}

mbe_status mbe_graph_from_csr(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                              const uint32_t* adj_u, mbe_graph** out) {
    if (!out || (n_u && !off_u)) return fail(MBE_EINVAL, "null argument");
    if (n_u && off_u[n_u] && !adj_u) return fail(MBE_EINVAL, "null adjacency");
    auto h = std::make_unique<mbe_graph>();
    std::string err;
    if (!mbe::graph_from_csr(n_u, n_v, off_u, adj_u, h->g, &err)) return fail(MBE_EINVAL, err);
    *out = h.release();
    return MBE_OK;
}

mbe_status mbe_graph_info(const mbe_graph* g, uint32_t* n_u, uint32_t* n_v, uint64_t* edges) {
    if (!g) return fail(MBE_EINVAL, "null graph");
    if (n_u) *n_u = g->g.n_u;
    if (n_v) *n_v = g->g.n_v;
    if (edges) *edges = g->g.edges();
    return MBE_OK;
}

mbe_status mbe_graph_csr(const mbe_graph* g, const uint32_t** off_u, const uint32_t** adj_u) {
    if (!g || !off_u || !adj_u) return fail(MBE_EINVAL, "null argument");
    *off_u = g->g.off.data();
    *adj_u = g->g.adj.data();
    return MBE_OK;
}

void mbe_graph_free(mbe_graph* g) { delete g; }

mbe_status mbe_gen_chung_lu(uint32_t n_u, uint32_t n_v, uint64_t m, double gamma_u, double gamma_v,
                            uint32_t max_deg, uint64_t seed, mbe_graph** out) {
    if (!out || gamma_u <= 1.0 || gamma_v <= 1.0) return fail(MBE_EINVAL, "bad argument");
    auto h = std::make_unique<mbe_graph>();
    h->g = mbe::gen_chung_lu(n_u, n_v, m, gamma_u, gamma_v, max_deg, seed);
    *out = h.release();
    return MBE_OK;
}

mbe_status mbe_gen_er_planted(uint32_t n_u, uint32_t n_v, double p, uint32_t n_planted, uint32_t lo,
                              uint32_t hi, double top_frac, uint64_t seed, mbe_graph** out) {
    if (!out || p < 0 || p > 1 || lo > hi || top_frac < 0 || top_frac > 1)
        return fail(MBE_EINVAL, "bad argument");
    auto h = std::make_unique<mbe_graph>();
    h->g = mbe::gen_er_planted(n_u, n_v, p, n_planted, lo, hi, top_frac, seed);
    *out = h.release();
    return MBE_OK;
}

mbe_status mbe_gen_batch_graph(uint64_t seed, mbe_graph** out) {
    if (!out) return fail(MBE_EINVAL, "null argument");
    auto h = std::make_unique<mbe_graph>();
    h->g = mbe::gen_batch_graph(seed);
    *out = h.release();
    return MBE_OK;
}

mbe_status mbe_batch_solve(const mbe_graph* const* graphs, size_t n, mbe_variant variant,
                           const mbe_tuning* tuning, mbe_result* per_graph, mbe_result* total) {
    if ((!graphs && n) || (unsigned)variant > 3u) return fail(MBE_EINVAL, "bad argument");
    for (size_t i = 0; i < n; i++)
        if (!graphs[i]) return fail(MBE_EINVAL, "null graph in batch");
    const double t0 = now_ms();
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    mbe_status s = E.ensure_init();
    if (s != MBE_OK) return s;
    Tune tu = resolve(tuning, E.cus, E.lds_per_cu);
    if (tu.side < -1 || tu.side > 1) return fail(MBE_EINVAL, "bad side");
    std::vector<mbe::Prepared> ps(n);
    for (size_t i = 0; i < n; i++) mbe::prepare(graphs[i]->g, tu.side, ps[i]);
    std::vector<mbe::DevGraph> dg;
    uint64_t total_tasks = 0;
    mbe_result agg{};
    std::vector<uint64_t> res;
    if (n) {
        s = upload(E, ps, dg, total_tasks);
        if (s != MBE_OK) return s;
        s = run_tasks(E, ps, dg, total_tasks, (int)variant, tu, res, agg);
        if (s != MBE_OK && s != MBE_ETIMEOUT) return s;  // a timeout keeps partial results
    }
    agg.total_ms = now_ms() - t0;
    for (size_t i = 0; i < n; i++) {
        agg.count += res[2 * i];
        agg.hash += res[2 * i + 1];
        if (per_graph) {
            per_graph[i] = agg;
            per_graph[i].count = res[2 * i];
            per_graph[i].hash = res[2 * i + 1];
        }
    }
    if (per_graph)
        for (size_t i = 0; i < n; i++) {
            per_graph[i].kernel_ms = agg.kernel_ms;
            per_graph[i].total_ms = agg.total_ms;
        }
    if (total) *total = agg;
    return s;
}

mbe_status mbe_solve(const mbe_graph* g, mbe_variant variant, const mbe_tuning* tuning,
                     mbe_result* out) {
    if (!g || !out) return fail(MBE_EINVAL, "null argument");
    const mbe_graph* one[1] = {g};
    mbe_result r{};
    mbe_status s = mbe_batch_solve(one, 1, variant, tuning, nullptr, &r);
    if (s == MBE_OK || s == MBE_ETIMEOUT) *out = r;
    return s;
}

}  // extern "C"
