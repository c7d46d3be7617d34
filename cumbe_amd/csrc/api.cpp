// C-ABI implementation of libmbe.so: graph handles, generators and the HIP
// engine that launches the persistent enumeration kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>
#include <unistd.h>

#include "hash.h"
#include "host.h"

namespace {
thread_local std::string g_err;
thread_local int g_hip_err = 0;
thread_local std::vector<uint64_t> g_busy;  // per-worker busy ticks, last variance solve
}  // namespace

namespace mbe {
namespace host {

mbe_status fail(mbe_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

void set_hip_error(int e) { g_hip_err = e; }


double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}



// Process-wide engine state for device 0: worker resources are allocated once
// and reused by every solve (a solve is serialised by the mutex).
struct Engine {
    std::mutex mu;
    bool init = false;
    int cus = 0;
    size_t lds_per_cu = 160 * 1024;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Bufs b;
    unsigned long long* abort_flag = nullptr;  // host-mapped watchdog flag
    // device address of the allocation-free solve's static scratch array
    // (taken once, with the engine's device current)
    unsigned long long* scratch = nullptr;

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
        HIPC(mbe::noalloc_scratch_address((void**)&scratch));
        init = true;
        return MBE_OK;
    }
};

Engine& engine() {
    static Engine* e = new Engine();  // never destroyed: outlives HIP teardown
    return *e;
}


Tune resolve(const mbe_tuning* t, int cus, size_t lds_per_cu) {
    Tune r;
    r.lds_words = (t && t->lds_words) ? t->lds_words : 3072;  // 24 KiB: 6 workers per CU
    r.lds_words = std::max<uint32_t>(r.lds_words, mbe::kLdsHeaderWords + 64);
    r.lds_words = std::min<uint32_t>(r.lds_words, (uint32_t)(lds_per_cu / 8));
    // default occupancy: as many single-wave workers per CU as the LDS holds
    uint32_t fit = (uint32_t)std::max<size_t>(1, lds_per_cu / ((size_t)r.lds_words * 8));
    uint32_t wpc = (t && t->waves_per_cu) ? t->waves_per_cu : std::min<uint32_t>(fit, 8);
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
    r.sparse_ratio = (t && t->sparse_ratio) ? (t->sparse_ratio == UINT32_MAX ? 0 : t->sparse_ratio) : 8;
    r.split_nodes = (t && t->split_nodes) ? (t->split_nodes == UINT32_MAX ? 0 : t->split_nodes) : 4096;
    return r;
}

// Upload prepared graphs into one device buffer and build the descriptors.
mbe_status upload(Bufs& B, hipStream_t stream, const std::vector<const mbe::Prepared*>& ps,
                  std::vector<mbe::DevGraph>& dg, uint64_t& total_tasks) {
    auto slot = upload_slot;
    size_t bytes = 0;
    for (const mbe::Prepared* pp : ps) bytes += packed_bytes(*pp);
    HIPC(B.data.ensure(bytes));
    std::vector<char> host(bytes);
    size_t o = 0;
    dg.resize(ps.size());
    total_tasks = 0;
    char* dbase = B.data.as<char>();
    auto put = [&](const void* src, size_t n) {
        size_t at = o;
        if (n) memcpy(host.data() + at, src, n);
        o += slot(n);
        return (const void*)(dbase + at);
    };
    for (size_t i = 0; i < ps.size(); i++) {
        const mbe::Prepared& p = *ps[i];
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
    HIPC(hipMemcpyAsync(B.data.p, host.data(), o, hipMemcpyHostToDevice, stream));
    HIPC(B.graphs.ensure(sizeof(mbe::DevGraph) * std::max<size_t>(1, dg.size())));
    HIPC(hipMemcpyAsync(B.graphs.p, dg.data(), sizeof(mbe::DevGraph) * dg.size(),
                        hipMemcpyHostToDevice, stream));
    HIPC(hipStreamSynchronize(stream));
    return MBE_OK;
}

// Wait for the launch; with MBE_PROGRESS=<seconds> set, print the roots and
// nodes the workers have finished (read from host-mapped slots) to stderr.
hipError_t wait_with_progress(Engine& E, const Bufs& B, uint32_t nw, uint64_t n_tasks, int pass,
                              double deadline_ms) {
    const char* env = getenv("MBE_PROGRESS");
    const double every = env ? atof(env) * 1000.0 : 0.0;
    const double t0 = now_ms();
    double last = t0;
    for (;;) {
        hipError_t q = hipEventQuery(E.ev1);
        if (q != hipErrorNotReady) return q == hipSuccess ? hipEventSynchronize(E.ev1) : q;
        usleep(200);
        if (deadline_ms > 0 && now_ms() > deadline_ms)
            __atomic_store_n(E.abort_flag, 1ull, __ATOMIC_SEQ_CST);
        if (every > 0 && now_ms() - last >= every) {
            last = now_ms();
            unsigned long long roots = 0, nodes = 0;
            for (uint32_t w = 0; w < nw; w++) {
                roots += ((volatile unsigned long long*)B.progress)[2 * w];
                nodes += ((volatile unsigned long long*)B.progress)[2 * w + 1];
            }
            fprintf(stderr, "[mbe] pass %d: %.1f s, %llu/%llu roots, %llu nodes\n", pass,
                    (last - t0) / 1000.0, roots, (unsigned long long)n_tasks, nodes);
        }
    }
}



// Size the buffers of one pass and fill its kernel parameters. The task list
// (explicit global task ids, or none for all tasks) is copied to the device.
mbe_status setup_pass(Bufs& B, hipStream_t stream, unsigned long long* abort_host, const Meta& m,
                      const Tune& tu, int variant, const std::vector<uint64_t>& task_list,
                      uint64_t n_tasks, uint64_t arena_words, uint32_t workers, mbe::KParams& kp) {
    HIPC(B.ctl.ensure(m.ctl_words() * 8));
    const uint64_t ovf_cap = std::max<uint64_t>(m.total_tasks, 1);
    HIPC(B.ovf.ensure(ovf_cap * 8 * 2));
    HIPC(B.arena.ensure(arena_words * workers * 8));
    HIPC(B.mark.ensure(m.mark_words() * workers * 8));
    HIPC(B.cpos.ensure(m.cpos_len() * workers * 4));
    // a slot holds a compacted level of any problem that fits an arena, up
    // to 4 MiB per slot
    const uint64_t slot_words = std::min<uint64_t>(std::max<uint64_t>(4096, arena_words), 1u << 19);
    HIPC(B.slots.ensure((size_t)kSlots * slot_words * 8));
    HIPC(B.dq.ensure((4 + (size_t)kSlots) * 8));
    HIPC(B.ensure_progress(workers));
    memset(B.progress, 0, 2 * (size_t)workers * 8);
    unsigned long long* ctl = B.ctl.as<unsigned long long>();
    uint64_t* ovf_list = B.ovf.as<uint64_t>();
    uint64_t* task_dev = ovf_list + ovf_cap;
    if (!task_list.empty())
        HIPC(hipMemcpyAsync(task_dev, task_list.data(), task_list.size() * 8,
                            hipMemcpyHostToDevice, stream));
    kp = mbe::KParams{};
    kp.graphs = B.graphs.as<mbe::DevGraph>();
    kp.n_graphs = m.ng;
    kp.variant = (uint32_t)variant;
    kp.n_tasks = n_tasks;
    kp.task_list = task_list.empty() ? nullptr : task_dev;
    kp.cursor = ctl;
    kp.overflow_count = ctl + 1;
    kp.stats = ctl + 2;
    const bool variance = tu.flags & MBE_FLAG_VARIANCE;
    kp.sections = (tu.flags & MBE_FLAG_SECTIONS) || variance ? ctl + 2 + mbe::kStatCount : nullptr;
    if (variance) {
        HIPC(B.busy.ensure((size_t)workers * 8));
        kp.worker_busy = B.busy.as<unsigned long long>();
    }
    kp.results = ctl + 2 + mbe::kStatCount + mbe::kSecCount;
    kp.arena = B.arena.as<uint64_t>();
    kp.arena_words = arena_words;
    kp.mark = B.mark.as<uint64_t>();
    kp.mark_words = m.mark_words();
    kp.cpos = B.cpos.as<uint32_t>();
    kp.cpos_len = m.cpos_len();
    kp.lds_words = tu.lds_words;
    kp.list_cutoff = tu.list_cutoff;
    kp.compact_ratio = tu.compact_ratio;
    kp.sparse_ratio = tu.sparse_ratio;
    kp.split_nodes = tu.split_nodes;
    kp.twins = (tu.flags & MBE_FLAG_REFERENCE_TREE) ? 0 : 1;
    HIPC(hipHostGetDevicePointer((void**)&kp.abort, abort_host, 0));
    kp.node_count = kp.stats + mbe::kStatNodeBudget;
    kp.node_budget = tu.max_nodes;
    kp.overflow_list = ovf_list;
    kp.overflow_cap = ovf_cap;
    kp.donate_depth = tu.donate_depth;
    kp.donate_interval = tu.donate_interval;
    kp.slots = B.slots.as<uint64_t>();
    kp.slot_words = slot_words;
    kp.n_slots = kSlots;
    kp.dq = B.dq.as<unsigned long long>();
    HIPC(hipHostGetDevicePointer((void**)&kp.progress, B.progress, 0));
    return MBE_OK;
}

// Enqueue one pass with kernels only (so that it can be captured into a
// graph): clear the control words (all of them on the first pass; cursor and
// overflow count later, which keeps the results), clear the root marks,
// reset the deque, launch the persistent workers.
hipError_t enqueue_pass(const mbe::KParams& kp, uint32_t workers, size_t ctl_clear,
                        hipStream_t s) {
    hipError_t e = mbe::launch_fill(kp.cursor, ctl_clear, 0ull, s);
    if (e == hipSuccess)
        e = mbe::launch_fill((unsigned long long*)kp.mark, kp.mark_words * workers, 0ull, s);
    if (e == hipSuccess) e = mbe::launch_dq_init(kp.dq, kp.n_slots, s);
    if (e == hipSuccess) e = mbe::launch_mbe(kp, workers, s);
    return e;
}

// A pass-0 launch prepared in advance (a batch plan's instantiated graph).
struct Pass0 {
    hipGraphExec_t exec = nullptr;
    mbe::KParams kp{};
    uint32_t workers = 0;
};

// Device side of a materialising solve (kernel records, see KParams::emit).
struct Emit {
    unsigned long long* pos = nullptr;  // 2 words, zeroed before the first pass
    uint64_t* buf = nullptr;
    uint64_t cap = 0;
};

// Run every root task of the uploaded graphs; roots whose local problem does
// not fit a worker arena are retried with fewer workers and larger arenas.
mbe_status run_tasks(Engine& E, Bufs& B, const Meta& m, int variant, const Tune& tu,
                     std::vector<uint64_t>& res, mbe_result& out, const Pass0* pre = nullptr,
                     const Emit* em = nullptr) {
    {
        mbe_status s = ensure_lds(tu.lds_words);
        if (s != MBE_OK) return s;
    }
    uint64_t n_tasks = m.total_tasks;
    uint64_t arena_words = tu.arena_words;
    uint32_t workers = tu.workers;
    std::vector<uint64_t> task_list;  // empty = range mode
    if (m.ng == 1 && (tu.root_begin || tu.root_end || tu.root_stride > 1)) {
        const uint64_t end = tu.root_end ? std::min<uint64_t>(tu.root_end, m.total_tasks) : m.total_tasks;
        for (uint64_t r = tu.root_begin; r < end; r += tu.root_stride) task_list.push_back(r);
        n_tasks = task_list.size();
    }
    float kernel_ms = 0;
    uint32_t launches = 0;
    const uint64_t budget_words = (uint64_t)48 << 30 >> 3;  // 48 GiB of arenas
    const double deadline = tu.time_limit_ms > 0 ? now_ms() + tu.time_limit_ms : 0;
    __atomic_store_n(E.abort_flag, 0ull, __ATOMIC_SEQ_CST);
    mbe::KParams kp{};
    for (int pass = 0; n_tasks > 0; pass++) {
        if (pass > 6) return fail(MBE_ENOMEM, "root subproblem does not fit any arena");
        uint32_t nw = workers;  // all start: those without a root task steal
        if (pass == 0 && pre) {
            kp = pre->kp;
            nw = pre->workers;
            memset(B.progress, 0, 2 * (size_t)nw * 8);
            HIPC(hipEventRecord(E.ev0, E.stream));
            HIPC(hipGraphLaunch(pre->exec, E.stream));
        } else {
            mbe_status s = setup_pass(B, E.stream, E.abort_flag, m, tu, variant, task_list, n_tasks,
                                      arena_words, nw, kp);
            if (s != MBE_OK) return s;
            if (em) {
                kp.emit_pos = em->pos;
                kp.emit = em->buf;
                kp.emit_cap = em->cap;
            }
            HIPC(hipEventRecord(E.ev0, E.stream));
            HIPC(enqueue_pass(kp, nw, pass == 0 ? m.ctl_words() : 2, E.stream));
        }
        HIPC(hipEventRecord(E.ev1, E.stream));
        HIPC(wait_with_progress(E, B, nw, n_tasks, pass, deadline));
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, E.ev0, E.ev1));
        kernel_ms += ms;
        launches++;
        if (kp.worker_busy) {
            g_busy.resize(nw);
            HIPC(hipMemcpy(g_busy.data(), kp.worker_busy, (size_t)nw * 8, hipMemcpyDeviceToHost));
        }
        out.workers = std::max(out.workers, nw);
        unsigned long long novf = 0;
        HIPC(hipMemcpy(&novf, kp.overflow_count, 8, hipMemcpyDeviceToHost));
        if (__atomic_load_n(E.abort_flag, __ATOMIC_SEQ_CST)) break;
        if (!novf) break;
        task_list.resize(novf);
        HIPC(hipMemcpy(task_list.data(), kp.overflow_list, novf * 8, hipMemcpyDeviceToHost));
        n_tasks = novf;
        arena_words *= 16;
        workers = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(workers, budget_words / arena_words));
    }
    if (!launches) {  // no root task at all: every graph has count 0
        res.assign(2 * (size_t)m.ng, 0);
        return MBE_OK;
    }
    std::vector<unsigned long long> st(mbe::kStatCount), rr(2 * (size_t)m.ng), sc(mbe::kSecCount);
    HIPC(hipMemcpy(st.data(), kp.stats, st.size() * 8, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(sc.data(), kp.stats + mbe::kStatCount, sc.size() * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < mbe::kSecCount; i++) out.sections[i] = sc[i];
    HIPC(hipMemcpy(rr.data(), kp.results, rr.size() * 8, hipMemcpyDeviceToHost));
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

size_t packed_bytes(const mbe::Prepared& p) {
    return upload_slot(p.offR.size() * 4) + upload_slot(p.adjR.size() * 4) +
           upload_slot(p.offL.size() * 4) + upload_slot(p.adjL.size() * 4) +
           upload_slot(p.hR.size() * 8) + upload_slot(p.hL.size() * 8);
}

Meta meta_of(const std::vector<const mbe::Prepared*>& ps, uint64_t total_tasks) {
    Meta m;
    m.ng = (uint32_t)ps.size();
    for (const mbe::Prepared* p : ps) m.max_nr = std::max(m.max_nr, p->nR);
    m.total_tasks = total_tasks;
    return m;
}

// Per-graph and total results of a batch from the result words.
void fill_results(const std::vector<uint64_t>& res, size_t n, mbe_result& agg,
                  mbe_result* per_graph) {
    for (size_t i = 0; i < n; i++) {
        agg.count += res[2 * i];
        agg.hash += res[2 * i + 1];
    }
    if (per_graph)
        for (size_t i = 0; i < n; i++) {
            per_graph[i] = agg;
            per_graph[i].count = res[2 * i];
            per_graph[i].hash = res[2 * i + 1];
        }
}

// Turn the kernel's records into the caller's: L ids from the hash terms,
// R' = the common neighbourhood of L' (checked against the reported size and
// hash sum), file coordinates, records sorted. count is the solve's count.
mbe_status decode_bicliques(const mbe::Prepared& p, const std::vector<uint64_t>& dev,
                            uint64_t count, uint32_t* out, uint64_t words) {
    const uint64_t salt_l = p.r_is_u ? mbe::kSaltV : mbe::kSaltU;
    std::vector<uint32_t> recs;
    std::vector<uint64_t> starts;
    std::vector<uint32_t> a, b, t, bid;
    recs.reserve(words);
    for (size_t at = 0; at < dev.size();) {
        const uint64_t nl = dev[at] & 0xffffffffu, nr = dev[at] >> 32, sr = dev[at + 1];
        if (nl == 0 || at + 2 + nl > dev.size()) return fail(MBE_EINTERNAL, "malformed biclique record");
        a.clear();
        for (uint64_t i = 0; i < nl; i++) {
            const uint64_t term = dev[at + 2 + i], id = mbe::unmix64(term) ^ salt_l;
            if (id >= p.nL || p.hL[id] != term) return fail(MBE_EINTERNAL, "unknown L hash term");
            a.push_back((uint32_t)id);
        }
        at += 2 + nl;
        std::sort(a.begin(), a.end());
        b.assign(p.adjL.begin() + p.offL[a[0]], p.adjL.begin() + p.offL[a[0] + 1]);
        for (size_t i = 1; i < a.size() && !b.empty(); i++) {
            t.clear();
            std::set_intersection(b.begin(), b.end(), p.adjL.begin() + p.offL[a[i]],
                                  p.adjL.begin() + p.offL[a[i] + 1], std::back_inserter(t));
            b.swap(t);
        }
        uint64_t h = 0;
        for (uint32_t r : b) h += p.hR[r];
        if (b.size() != nr || h != sr)
            return fail(MBE_EINTERNAL, "biclique's R side differs from the common neighbourhood");
        bid.clear();
        for (uint32_t r : b) bid.push_back(p.rank2id[r]);
        std::sort(bid.begin(), bid.end());
        const std::vector<uint32_t>& su = p.r_is_u ? bid : a;
        const std::vector<uint32_t>& sv = p.r_is_u ? a : bid;
        starts.push_back(recs.size());
        recs.push_back((uint32_t)su.size());
        recs.push_back((uint32_t)sv.size());
        recs.insert(recs.end(), su.begin(), su.end());
        recs.insert(recs.end(), sv.begin(), sv.end());
    }
    if (starts.size() != count || recs.size() != words)
        return fail(MBE_EINTERNAL, "biclique records do not match the count");
    auto key = [&](uint64_t s, int side) {
        const uint32_t* r = recs.data() + s;
        const uint32_t* first = r + 2 + (side ? r[0] : 0);
        return std::make_pair(first, first + r[side]);
    };
    std::sort(starts.begin(), starts.end(), [&](uint64_t x, uint64_t y) {
        auto ux = key(x, 0), uy = key(y, 0);
        if (!std::equal(ux.first, ux.second, uy.first, uy.second))
            return std::lexicographical_compare(ux.first, ux.second, uy.first, uy.second);
        auto vx = key(x, 1), vy = key(y, 1);
        return std::lexicographical_compare(vx.first, vx.second, vy.first, vy.second);
    });
    uint64_t o = 0;
    for (uint64_t s : starts) {
        const uint64_t len = 2 + (uint64_t)recs[s] + recs[s + 1];
        memcpy(out + o, recs.data() + s, len * 4);
        o += len;
    }
    return MBE_OK;
}

mbe_status ensure_lds(uint32_t lds_words) {
    static std::mutex mu;
    static uint32_t set = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (set >= lds_words) return MBE_OK;
    HIPC(mbe::set_lds_limit(lds_words));
    set = lds_words;
    return MBE_OK;
}

}  // namespace host
}  // namespace mbe

using namespace mbe::host;

// A batch kept resident on the device: graphs uploaded once, and the first
// pass (control clears, deque reset, persistent launch) captured as a graph
// that every run replays.
struct mbe_batch_plan {
    std::vector<mbe::Prepared> ps;
    Meta m;
    int variant = 0;
    Tune tu{};
    Bufs b;
    hipGraph_t graph = nullptr;
    Pass0 pass0;
    ~mbe_batch_plan() {
        if (pass0.exec) (void)hipGraphExecDestroy(pass0.exec);
        if (graph) (void)hipGraphDestroy(graph);
        b.release();
    }
};

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
        case MBE_ERANGE: return "output buffer too small";
        case MBE_EINTERNAL: return "internal self-check failed";
    }
    return "unknown status";
}

int mbe_last_hip_error(void) { return g_hip_err; }

uint32_t mbe_last_worker_busy(uint64_t* out, uint32_t cap) {
    const size_t n = std::min<size_t>(cap, g_busy.size());
    if (out && n) memcpy(out, g_busy.data(), n * 8);
    return (uint32_t)g_busy.size();
}
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
    std::vector<const mbe::Prepared*> pp(n);
    for (size_t i = 0; i < n; i++) {
        mbe::prepare(graphs[i]->g, tu.side, ps[i]);
        pp[i] = &ps[i];
    }
    std::vector<mbe::DevGraph> dg;
    uint64_t total_tasks = 0;
    mbe_result agg{};
    std::vector<uint64_t> res;
    if (n) {
        s = upload(E.b, E.stream, pp, dg, total_tasks);
        if (s != MBE_OK) return s;
        s = run_tasks(E, E.b, meta_of(pp, total_tasks), (int)variant, tu, res, agg);
        if (s != MBE_OK && s != MBE_ETIMEOUT) return s;  // a timeout keeps partial results
    }
    agg.total_ms = now_ms() - t0;
    fill_results(res, n, agg, per_graph);
    if (total) *total = agg;
    return s;
}

mbe_status mbe_solve_list(const mbe_graph* g, mbe_variant variant, const mbe_tuning* tuning,
                          uint32_t* out, uint64_t cap_words, uint64_t* used_words,
                          mbe_result* result) {
    if (!g || !used_words || (unsigned)variant > 3u) return fail(MBE_EINVAL, "bad argument");
    if (!out && cap_words) return fail(MBE_EINVAL, "null output with a nonzero capacity");
    if (cap_words > ((uint64_t)1 << 40)) return fail(MBE_ENOMEM, "output capacity above 2^40 words");
    const double t0 = now_ms();
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    mbe_status s = E.ensure_init();
    if (s != MBE_OK) return s;
    Tune tu = resolve(tuning, E.cus, E.lds_per_cu);
    if (tu.side < -1 || tu.side > 1) return fail(MBE_EINVAL, "bad side");
    mbe::Prepared p;
    mbe::prepare(g->g, tu.side, p);
    std::vector<const mbe::Prepared*> pp{&p};
    std::vector<mbe::DevGraph> dg;
    uint64_t total_tasks = 0;
    s = upload(E.b, E.stream, pp, dg, total_tasks);
    if (s != MBE_OK) return s;
    // A record takes 2 + |L'| u64 words on the device and 2 + |L'| + |R'|
    // (|R'| >= 1) u32 words for the caller, so cap_words device words hold
    // every record whenever the caller's buffer holds them all.
    struct Owned {
        DevBuf b;
        ~Owned() {
            if (b.p) (void)hipFree(b.p);
        }
    } eb;
    HIPC(eb.b.ensure((2 + cap_words) * 8));
    HIPC(hipMemsetAsync(eb.b.p, 0, 16, E.stream));
    const Emit em{eb.b.as<unsigned long long>(), eb.b.as<uint64_t>() + 2, cap_words};
    mbe_result agg{};
    std::vector<uint64_t> res;
    s = run_tasks(E, E.b, meta_of(pp, total_tasks), (int)variant, tu, res, agg, nullptr, &em);
    if (s != MBE_OK) return s;  // a watchdog stop leaves the list incomplete
    unsigned long long pos[2] = {0, 0};
    HIPC(hipMemcpy(pos, eb.b.p, 16, hipMemcpyDeviceToHost));
    *used_words = pos[1];
    fill_results(res, 1, agg, nullptr);
    if (out) {
        if (pos[1] > cap_words)
            return fail(MBE_ERANGE, "output holds " + std::to_string(cap_words) +
                                        " words; the bicliques take " + std::to_string(pos[1]));
        try {
            std::vector<uint64_t> dev(pos[0]);
            if (pos[0])
                HIPC(hipMemcpy(dev.data(), em.buf, pos[0] * 8, hipMemcpyDeviceToHost));
            s = decode_bicliques(p, dev, agg.count, out, pos[1]);
        } catch (const std::bad_alloc&) {
            return fail(MBE_ENOMEM, "host memory for the biclique records");
        }
        if (s != MBE_OK) return s;
    }
    agg.total_ms = now_ms() - t0;
    if (result) *result = agg;
    return MBE_OK;
}

mbe_status mbe_init(void) {
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    return E.ensure_init();
}

mbe_status mbe_graph_pack(const mbe_graph* g, int side, void* out, uint64_t cap_bytes,
                          mbe_packed_graph* info) {
    if (!g || !info || side < -1 || side > 1 || (!out && cap_bytes))
        return fail(MBE_EINVAL, "bad argument");
    mbe::Prepared p;
    mbe::prepare(g->g, side, p);
    const size_t sizes[6] = {p.offR.size() * 4, p.adjR.size() * 4, p.offL.size() * 4,
                             p.adjL.size() * 4, p.hR.size() * 8, p.hL.size() * 8};
    const void* src[6] = {p.offR.data(), p.adjR.data(), p.offL.data(),
                          p.adjL.data(), p.hR.data(), p.hL.data()};
    mbe_packed_graph inf{};
    inf.n_r = p.nR;
    inf.n_l = p.nL;
    inf.n_roots = p.n_roots;
    inf.r_is_u = p.r_is_u ? 1 : 0;
    uint64_t o = 0;
    for (int i = 0; i < 6; i++) {
        inf.off[i] = o;
        o += upload_slot(sizes[i]);
    }
    inf.bytes = o;
    *info = inf;
    if (!out) return MBE_OK;
    if (cap_bytes < inf.bytes)
        return fail(MBE_ERANGE, "the packed graph takes " + std::to_string(inf.bytes) + " bytes");
    memset(out, 0, inf.bytes);
    for (int i = 0; i < 6; i++)
        if (sizes[i]) memcpy((char*)out + inf.off[i], src[i], sizes[i]);
    return MBE_OK;
}

mbe_status mbe_solve_noalloc(const void* dev_graph, const mbe_packed_graph* info,
                             mbe_variant variant, const mbe_tuning* tuning, mbe_result* out) {
    if (!dev_graph || !info || !out || (unsigned)variant > 3u) return fail(MBE_EINVAL, "bad argument");
    for (int i = 0; i < 6; i++)
        if (info->off[i] % 256 || info->off[i] >= info->bytes)
            return fail(MBE_EINVAL, "packed graph offsets out of range");
    if (info->n_roots > info->n_r) return fail(MBE_EINVAL, "packed graph has more roots than vertices");
    const double t0 = now_ms();
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (!E.init) return fail(MBE_EINVAL, "mbe_solve_noalloc needs mbe_init() at set-up");
    Tune tu = resolve(tuning, E.cus, E.lds_per_cu);
    if (tu.flags & MBE_FLAG_VARIANCE)
        return fail(MBE_EINVAL, "per-worker busy times need a host buffer; not without allocation");
    if (!(tuning && tuning->arena_words)) tu.arena_words = kNoAllocArenaWords;
    const bool fixed_workers = tuning && tuning->workers;
    mbe_status s = ensure_lds(tu.lds_words);
    if (s != MBE_OK) return s;

    // root range of this call: ranks root_begin, +stride, ... below root_end
    const uint64_t rend = tu.root_end ? std::min<uint64_t>(tu.root_end, info->n_roots) : info->n_roots;
    uint64_t n_tasks = tu.root_begin < rend ? (rend - tu.root_begin + tu.root_stride - 1) / tu.root_stride : 0;
    mbe::DevGraph dg{};
    const char* base = (const char*)dev_graph;
    dg.offR = (const uint32_t*)(base + info->off[0]);
    dg.adjR = (const uint32_t*)(base + info->off[1]);
    dg.offL = (const uint32_t*)(base + info->off[2]);
    dg.adjL = (const uint32_t*)(base + info->off[3]);
    dg.hR = (const uint64_t*)(base + info->off[4]);
    dg.hL = (const uint64_t*)(base + info->off[5]);
    dg.nR = info->n_r;
    dg.nL = info->n_l;
    dg.n_roots = info->n_roots;
    dg.r_is_u = info->r_is_u;
    dg.task_base = 0;

    // scratch words: control block, descriptor, deque, two task lists (the
    // overflow list of one pass is the task list of the next), slots, then
    // per worker an arena, a mark bitmap and a rank table
    Meta m;
    m.ng = 1;
    m.max_nr = std::max<uint32_t>(1, info->n_r);
    m.total_tasks = n_tasks;
    auto up4 = [](uint64_t w) { return (w + 31) & ~(uint64_t)31; };  // 256-byte steps
    const uint64_t w_ctl = up4(m.ctl_words()), w_graph = up4(sizeof(mbe::DevGraph) / 8 + 1),
                   w_dq = up4(4 + kNoAllocSlots), w_list = up4(std::max<uint64_t>(n_tasks, 1));
    unsigned long long* ctl = E.scratch;
    mbe::DevGraph* gdev = (mbe::DevGraph*)(E.scratch + w_ctl);
    unsigned long long* dq = E.scratch + w_ctl + w_graph;
    uint64_t* lists = (uint64_t*)(dq + w_dq);
    const uint64_t fixed = w_ctl + w_graph + w_dq + 2 * w_list;
    if (n_tasks == 0) {
        *out = mbe_result{};
        out->total_ms = now_ms() - t0;
        return MBE_OK;
    }
    __atomic_store_n(E.abort_flag, 0ull, __ATOMIC_SEQ_CST);
    const double deadline = tu.time_limit_ms > 0 ? now_ms() + tu.time_limit_ms : 0;
    uint64_t arena_words = tu.arena_words;
    uint32_t workers = tu.workers;
    float kernel_ms = 0;
    uint32_t launches = 0, used_workers = 0;
    mbe::KParams kp{};
    for (int pass = 0; n_tasks > 0; pass++) {
        if (pass > 6) return fail(MBE_ERANGE, "root subproblem does not fit an arena of the scratch");
        const uint64_t slot_words = std::min<uint64_t>(std::max<uint64_t>(4096, arena_words), 1u << 19);
        const uint64_t w_slots = up4(kNoAllocSlots * slot_words);
        const uint64_t per_worker = up4(arena_words) + up4(m.mark_words()) + up4((m.cpos_len() + 1) / 2);
        const uint64_t fit = fixed + w_slots < mbe::kNoAllocWords
                                 ? (mbe::kNoAllocWords - fixed - w_slots) / per_worker : 0;
        if (fixed_workers && pass == 0 && workers > fit)
            return fail(MBE_ERANGE, std::to_string(workers) + " workers need " +
                                        std::to_string((fixed + w_slots + workers * per_worker) * 8) +
                                        " bytes of scratch; the allocation-free solve has " +
                                        std::to_string(mbe::kNoAllocBytes));
        const uint32_t nw = (uint32_t)std::min<uint64_t>(workers, fit);
        if (nw == 0)
            return fail(MBE_ERANGE, "the graph's buffers exceed the allocation-free solve's " +
                                        std::to_string(mbe::kNoAllocBytes) + " bytes of scratch");
        uint64_t* slots = lists + 2 * w_list;
        uint64_t* arena = slots + w_slots;
        uint64_t* mark = arena + (uint64_t)nw * up4(arena_words);
        uint32_t* cpos = (uint32_t*)(mark + (uint64_t)nw * up4(m.mark_words()));
        kp = mbe::KParams{};
        kp.graphs = gdev;
        kp.n_graphs = 1;
        kp.variant = (uint32_t)variant;
        kp.n_tasks = n_tasks;
        kp.task_list = pass == 0 ? nullptr : lists + ((pass - 1) & 1) * w_list;
        kp.task_first = tu.root_begin;
        kp.task_step = tu.root_stride;
        kp.cursor = ctl;
        kp.overflow_count = ctl + 1;
        kp.stats = ctl + 2;
        kp.sections = (tu.flags & MBE_FLAG_SECTIONS) ? ctl + 2 + mbe::kStatCount : nullptr;
        kp.results = ctl + 2 + mbe::kStatCount + mbe::kSecCount;
        kp.arena = arena;
        kp.arena_words = up4(arena_words);
        kp.mark = mark;
        kp.mark_words = up4(m.mark_words());
        kp.cpos = cpos;
        kp.cpos_len = 2 * up4((m.cpos_len() + 1) / 2);
        kp.lds_words = tu.lds_words;
        kp.list_cutoff = tu.list_cutoff;
        kp.compact_ratio = tu.compact_ratio;
        kp.sparse_ratio = tu.sparse_ratio;
        kp.split_nodes = tu.split_nodes;
        kp.twins = (tu.flags & MBE_FLAG_REFERENCE_TREE) ? 0 : 1;
        HIPC(hipHostGetDevicePointer((void**)&kp.abort, E.abort_flag, 0));
        kp.node_count = kp.stats + mbe::kStatNodeBudget;
        kp.node_budget = tu.max_nodes;
        kp.overflow_list = lists + (pass & 1) * w_list;
        kp.overflow_cap = n_tasks;
        kp.donate_depth = tu.donate_depth;
        kp.donate_interval = tu.donate_interval;
        kp.slots = slots;
        kp.slot_words = slot_words;
        kp.n_slots = kNoAllocSlots;
        kp.dq = dq;
        HIPC(hipEventRecord(E.ev0, E.stream));
        if (pass == 0) HIPC(mbe::launch_put_graph(gdev, dg, E.stream));
        HIPC(enqueue_pass(kp, nw, pass == 0 ? m.ctl_words() : 2, E.stream));
        HIPC(hipEventRecord(E.ev1, E.stream));
        for (;;) {
            const hipError_t q = hipEventQuery(E.ev1);
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady) HIPC(q);
            usleep(100);
            if (deadline > 0 && now_ms() > deadline) __atomic_store_n(E.abort_flag, 1ull, __ATOMIC_SEQ_CST);
        }
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, E.ev0, E.ev1));
        kernel_ms += ms;
        launches++;
        used_workers = std::max(used_workers, nw);
        unsigned long long novf = 0;
        HIPC(hipMemcpy(&novf, kp.overflow_count, 8, hipMemcpyDeviceToHost));
        if (__atomic_load_n(E.abort_flag, __ATOMIC_SEQ_CST) || !novf) break;
        n_tasks = novf;
        arena_words *= 16;
        workers = nw;
    }
    unsigned long long st[mbe::kStatCount], sc[mbe::kSecCount], rr[2];
    HIPC(hipMemcpy(st, kp.stats, sizeof(st), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(sc, kp.stats + mbe::kStatCount, sizeof(sc), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(rr, kp.results, sizeof(rr), hipMemcpyDeviceToHost));
    mbe_result r{};
    r.count = rr[0];
    r.hash = rr[1];
    r.kernel_ms = kernel_ms;
    r.nodes = st[mbe::kStatNodes];
    r.donated = st[mbe::kStatDonated];
    r.launches = launches;
    r.workers = used_workers;
    for (int i = 0; i < mbe::kSecCount; i++) r.sections[i] = sc[i];
    r.words = st[mbe::kStatWords];
    r.build_words = st[mbe::kStatBuildWords];
    r.total_ms = now_ms() - t0;
    *out = r;
    if (__atomic_load_n(E.abort_flag, __ATOMIC_SEQ_CST))
        return fail(MBE_ETIMEOUT, "watchdog: node budget or time limit reached; results partial");
    if (st[mbe::kStatDropped]) return fail(MBE_ENOMEM, "donated subtree did not fit a worker arena");
    return MBE_OK;
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

namespace {

// Capture the plan's first pass into a graph (again after an overflow pass
// reallocated buffers the previous capture points into).
mbe_status capture_pass0(Engine& E, mbe_batch_plan* p) {
    if (p->pass0.exec) (void)hipGraphExecDestroy(p->pass0.exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->pass0.exec = nullptr;
    p->graph = nullptr;
    mbe_status ls = ensure_lds(p->tu.lds_words);
    if (ls != MBE_OK) return ls;
    std::vector<uint64_t> task_list;
    uint64_t n_tasks = p->m.total_tasks;
    if (p->m.ng == 1 && (p->tu.root_begin || p->tu.root_end || p->tu.root_stride > 1)) {
        const uint64_t end = p->tu.root_end ? std::min<uint64_t>(p->tu.root_end, n_tasks) : n_tasks;
        for (uint64_t r = p->tu.root_begin; r < end; r += p->tu.root_stride) task_list.push_back(r);
        n_tasks = task_list.size();
    }
    mbe_status s = setup_pass(p->b, E.stream, E.abort_flag, p->m, p->tu, p->variant, task_list,
                              n_tasks, p->tu.arena_words, p->tu.workers, p->pass0.kp);
    if (s != MBE_OK) return s;
    p->pass0.workers = p->tu.workers;
    HIPC(hipStreamSynchronize(E.stream));
    HIPC(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_pass(p->pass0.kp, p->pass0.workers, p->m.ctl_words(), E.stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(E.stream, &g);
    HIPC(e);
    HIPC(e2);
    p->graph = g;
    HIPC(hipGraphInstantiate(&p->pass0.exec, p->graph, nullptr, nullptr, 0));
    return MBE_OK;
}

bool capture_current(const mbe_batch_plan* p) {
    const mbe::KParams& kp = p->pass0.kp;
    return p->pass0.exec && kp.arena == p->b.arena.as<uint64_t>() &&
           kp.mark == p->b.mark.as<uint64_t>() && kp.cpos == p->b.cpos.as<uint32_t>() &&
           kp.cursor == p->b.ctl.as<unsigned long long>() &&
           kp.overflow_list == p->b.ovf.as<uint64_t>() && kp.slots == p->b.slots.as<uint64_t>() &&
           kp.dq == p->b.dq.as<unsigned long long>();
}

}  // namespace

extern "C" {

mbe_status mbe_batch_plan_create(const mbe_graph* const* graphs, size_t n, mbe_variant variant,
                                 const mbe_tuning* tuning, mbe_batch_plan** out) {
    if (!out || !graphs || !n || (unsigned)variant > 3u) return fail(MBE_EINVAL, "bad argument");
    for (size_t i = 0; i < n; i++)
        if (!graphs[i]) return fail(MBE_EINVAL, "null graph in batch");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    mbe_status s = E.ensure_init();
    if (s != MBE_OK) return s;
    auto p = std::make_unique<mbe_batch_plan>();
    p->tu = resolve(tuning, E.cus, E.lds_per_cu);
    if (p->tu.side < -1 || p->tu.side > 1) return fail(MBE_EINVAL, "bad side");
    p->variant = (int)variant;
    p->ps.resize(n);
    std::vector<const mbe::Prepared*> pp(n);
    for (size_t i = 0; i < n; i++) {
        mbe::prepare(graphs[i]->g, p->tu.side, p->ps[i]);
        pp[i] = &p->ps[i];
    }
    std::vector<mbe::DevGraph> dg;
    uint64_t total_tasks = 0;
    s = upload(p->b, E.stream, pp, dg, total_tasks);
    if (s != MBE_OK) return s;
    p->m = meta_of(pp, total_tasks);
    s = capture_pass0(E, p.get());
    if (s != MBE_OK) return s;
    *out = p.release();
    return MBE_OK;
}

mbe_status mbe_batch_plan_run(mbe_batch_plan* p, mbe_result* per_graph, mbe_result* total) {
    if (!p) return fail(MBE_EINVAL, "null plan");
    const double t0 = now_ms();
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    mbe_status s = E.ensure_init();
    if (s != MBE_OK) return s;
    if (!capture_current(p)) {
        s = capture_pass0(E, p);
        if (s != MBE_OK) return s;
    }
    mbe_result agg{};
    std::vector<uint64_t> res;
    s = run_tasks(E, p->b, p->m, p->variant, p->tu, res, agg, &p->pass0);
    if (s != MBE_OK && s != MBE_ETIMEOUT) return s;
    agg.total_ms = now_ms() - t0;
    fill_results(res, p->m.ng, agg, per_graph);
    if (total) *total = agg;
    return s;
}

void mbe_batch_plan_free(mbe_batch_plan* p) {
    if (!p) return;
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    delete p;
}

}  // extern "C"
