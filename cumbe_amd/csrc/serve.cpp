// Serving loop of libmbe.so: admission by HBM footprint, continuous batching
// over concurrent launch lanes, per-request completion from the kernel's
// host-mapped completion records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unistd.h>

#include "host.h"

namespace {

using namespace mbe::host;

struct Req {
    uint64_t id = 0;
    mbe::Prepared prep;
    int variant = 0;
    uint64_t footprint = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    mbe_request_result res{};
    bool done = false;
};
using ReqP = std::shared_ptr<Req>;

// One lane: a stream, resident buffers, and the requests of its current
// launch. Lanes never free buffers while serving (a free would wait for the
// whole device); they only grow for an arena-overflow pass.
struct Lane {
    uint32_t index = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev1 = nullptr;
    Bufs b;
    DevBuf rem;                             // 2 words per graph: open tasks, errors
    unsigned long long* abort_host = nullptr;  // host-mapped
    unsigned long long* done_host = nullptr;   // host-mapped, 4 words per graph
    uint32_t done_cap = 0;
    std::vector<ReqP> reqs;
    std::vector<uint8_t> closed;
    std::vector<uint64_t> tasks;            // task list of an overflow pass
    Meta m;
    mbe::KParams kp{};
    int variant = 0, pass = 0;
    uint64_t arena_words = 0;
    uint32_t workers = 0;
    bool busy = false;
};

// Largest arena (u64 words) the local problem of any root of p can take:
// hash terms, both bit matrices, scratch lists and a full-depth frame stack
// (the kernel's layout), with the distance-two candidate count bounded by
// the sum of the root's neighbours' other edges.
uint64_t root_bound_words(const mbe::Prepared& p) {
    std::vector<uint32_t> degL(p.nL);
    for (uint32_t u = 0; u < p.nL; u++) degL[u] = p.offL[u + 1] - p.offL[u];
    uint64_t best = 0;
    for (uint32_t r = 0; r < p.n_roots; r++) {
        const uint64_t d = p.offR[r + 1] - p.offR[r];
        uint64_t nc = 0;
        for (uint32_t k = p.offR[r]; k < p.offR[r + 1]; k++) nc += degL[p.adjR[k]] - 1;
        nc = std::min<uint64_t>(nc, p.nR ? p.nR - 1 : 0);
        const uint64_t W = (d + 63) / 64, Wc = (nc + 63) / 64;
        const uint64_t data = d + nc + nc * W + d * Wc;
        const uint64_t scratch = (nc + 1) / 2 + 1 + (d + 1) / 2 + 1;
        const uint64_t frames = (std::min(d, nc) + 2) * (4 + W + 3 * Wc);
        best = std::max(best, data + scratch + frames);
    }
    return best;
}

}  // namespace

struct mbe_server {
    mbe_server_config cfg{};
    Tune tu{};
    int cus = 0;
    std::mutex mu;
    std::condition_variable done_cv;  // poll waiters
    std::condition_variable work_cv;  // scheduler wake-up
    std::deque<ReqP> queue;
    std::map<uint64_t, ReqP> live;    // submitted, not yet collected
    std::vector<std::unique_ptr<Lane>> lanes;
    uint64_t next_id = 1;
    mbe_server_stats st{};
    std::thread th;
    bool stop = false;

    void complete(Lane& L, size_t i, mbe_status s, uint64_t count, uint64_t hash) {
        const ReqP& r = L.reqs[i];
        L.closed[i] = 1;
        r->t_done = now_ms();
        r->res.status = s;
        r->res.lane = L.index;
        r->res.count = count;
        r->res.hash = hash;
        r->res.footprint = r->footprint;
        r->res.queued_ms = r->t_start - r->t_submit;
        r->res.latency_ms = r->t_done - r->t_submit;
        r->done = true;
        st.bytes_in_use -= r->footprint;
        st.in_flight--;
        st.completed++;
        if (s == MBE_OK) st.bicliques += count;
        done_cv.notify_all();
    }

    // Take the flags of finished requests; when the launch has ended, start
    // an overflow pass or free the lane. Returns true if anything happened.
    bool poll_lane(Lane& L) {
        bool any = false;
        volatile unsigned long long* d = L.done_host;
        for (size_t i = 0; i < L.reqs.size(); i++) {
            if (L.closed[i] || !d[4 * i + 2]) continue;
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
            complete(L, i, d[4 * i + 3] ? MBE_ENOMEM : MBE_OK, d[4 * i], d[4 * i + 1]);
            any = true;
        }
        const hipError_t q = hipEventQuery(L.ev1);
        if (q == hipErrorNotReady) return any;
        mbe_status s = MBE_OK;
        unsigned long long novf = 0;
        if (q != hipSuccess) {
            set_hip_error((int)q);
            s = fail(MBE_EHIP, std::string("serving launch: ") + hipGetErrorString(q));
        } else if (hipMemcpyAsync(&novf, L.kp.overflow_count, 8, hipMemcpyDeviceToHost,
                                  L.stream) != hipSuccess ||
                   hipStreamSynchronize(L.stream) != hipSuccess) {
            s = MBE_EHIP;
        }
        if (s == MBE_OK && novf && L.pass < 6) {
            // roots whose local problem exceeded the arena: rerun them with
            // 16x arenas (their requests stay open until then)
            L.tasks.resize(novf);
            if (hipMemcpyAsync(L.tasks.data(), L.kp.overflow_list, novf * 8,
                               hipMemcpyDeviceToHost, L.stream) == hipSuccess &&
                hipStreamSynchronize(L.stream) == hipSuccess) {
                L.pass++;
                L.arena_words *= 16;
                const uint64_t budget_words = (uint64_t)16 << 30 >> 3;  // 16 GiB per lane
                L.workers = (uint32_t)std::max<uint64_t>(
                    1, std::min<uint64_t>(L.workers, budget_words / L.arena_words));
                if (launch(L, L.tasks, 2) == MBE_OK) return true;
            }
            s = MBE_ENOMEM;
        }
        for (size_t i = 0; i < L.reqs.size(); i++) {
            if (L.closed[i]) continue;
            if (d[4 * i + 2]) {
                __atomic_thread_fence(__ATOMIC_ACQUIRE);
                complete(L, i, d[4 * i + 3] ? MBE_ENOMEM : MBE_OK, d[4 * i], d[4 * i + 1]);
            } else {
                complete(L, i, s == MBE_OK ? MBE_EHIP : s, 0, 0);
            }
        }
        L.reqs.clear();
        L.closed.clear();
        L.busy = false;
        return true;
    }

    // Enqueue one pass of the lane's current requests.
    mbe_status launch(Lane& L, const std::vector<uint64_t>& tasks, size_t ctl_clear) {
        mbe_status s = setup_pass(L.b, L.stream, L.abort_host, L.m, tu, L.variant, tasks,
                                  tasks.empty() ? L.m.total_tasks : tasks.size(), L.arena_words,
                                  L.workers, L.kp);
        if (s != MBE_OK) return s;
        L.kp.remaining = L.rem.as<unsigned long long>();
        HIPC(hipHostGetDevicePointer((void**)&L.kp.done, L.done_host, 0));
        HIPC(enqueue_pass(L.kp, L.workers, ctl_clear, L.stream));
        HIPC(hipEventRecord(L.ev1, L.stream));
        st.launches++;
        return MBE_OK;
    }

    // Admit queued requests (arrival order, one variant per launch) into a
    // free lane while the budget allows, upload them and launch.
    bool start_lane(Lane& L) {
        if (queue.empty()) return false;
        const int variant = queue.front()->variant;
        const uint64_t budget = cfg.hbm_budget;
        std::vector<ReqP> take;
        uint64_t bytes = 0;
        while (!queue.empty() && take.size() < cfg.max_batch) {
            const ReqP& r = queue.front();
            if (r->variant != variant) break;
            if (st.bytes_in_use + r->footprint > budget) break;
            const uint64_t pb = packed_bytes(r->prep);
            if (!take.empty() && bytes + pb > cfg.lane_data_bytes) break;
            bytes += pb;
            st.bytes_in_use += r->footprint;
            take.push_back(r);
            queue.pop_front();
        }
        if (take.empty()) return false;
        st.queued -= take.size();
        st.in_flight += take.size();
        const double t = now_ms();
        std::vector<const mbe::Prepared*> pp;
        for (auto& r : take) {
            r->t_start = t;
            pp.push_back(&r->prep);
        }
        L.reqs = take;
        L.closed.assign(take.size(), 0);
        L.variant = variant;
        L.pass = 0;
        L.arena_words = tu.arena_words;
        L.workers = tu.workers;
        L.busy = true;
        std::vector<mbe::DevGraph> dg;
        uint64_t total = 0;
        mbe_status s = upload(L.b, L.stream, pp, dg, total);
        // open tasks per graph = its roots; graphs without roots finish now
        std::vector<unsigned long long> rem(2 * take.size(), 0);
        for (size_t i = 0; i < take.size(); i++) rem[2 * i] = take[i]->prep.n_roots;
        if (s == MBE_OK && L.rem.ensure(rem.size() * 8) != hipSuccess) s = MBE_ENOMEM;
        if (s == MBE_OK && hipMemcpyAsync(L.rem.p, rem.data(), rem.size() * 8,
                                          hipMemcpyHostToDevice, L.stream) != hipSuccess)
            s = MBE_EHIP;
        if (s == MBE_OK && hipStreamSynchronize(L.stream) != hipSuccess) s = MBE_EHIP;
        memset(L.done_host, 0, (size_t)L.done_cap * 32);
        L.m = meta_of(pp, total);
        for (size_t i = 0; i < take.size(); i++)
            if (!take[i]->prep.n_roots) complete(L, i, s, 0, 0);
        if (s == MBE_OK && total) s = launch(L, {}, L.m.ctl_words());
        if (s != MBE_OK || !total) {
            for (size_t i = 0; i < take.size(); i++)
                if (!L.closed[i]) complete(L, i, s, 0, 0);
            L.reqs.clear();
            L.closed.clear();
            L.busy = false;
        }
        return true;
    }

    void run() {
        (void)hipSetDevice(0);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool any = false;
            bool busy = false;
            for (auto& L : lanes)
                if (L->busy) any |= poll_lane(*L);
            for (auto& L : lanes)
                if (!L->busy) any |= start_lane(*L);
            for (auto& L : lanes) busy |= L->busy;
            if (stop && queue.empty() && !busy) break;
            if (any) continue;
            if (busy) {
                // completions arrive through host-mapped flags: poll them
                lk.unlock();
                usleep(20);
                lk.lock();
            } else {
                work_cv.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
};

namespace {

void free_lane(Lane& L) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    L.b.release();
    if (L.rem.p) (void)hipFree(L.rem.p);
    L.rem.p = nullptr;
    if (L.abort_host) (void)hipHostFree(L.abort_host);
    if (L.done_host) (void)hipHostFree(L.done_host);
    if (L.ev1) (void)hipEventDestroy(L.ev1);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    L = Lane{};
}

mbe_status make_lane(mbe_server* S, uint32_t index, Lane& L) {
    L.index = index;
    HIPC(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&L.ev1, hipEventDisableTiming));
    HIPC(hipHostMalloc((void**)&L.abort_host, 64, hipHostMallocMapped | hipHostMallocCoherent));
    *L.abort_host = 0;
    L.done_cap = S->cfg.max_batch;
    HIPC(hipHostMalloc((void**)&L.done_host, (size_t)L.done_cap * 32,
                       hipHostMallocMapped | hipHostMallocCoherent));
    // reserve the resident buffers once, sized for a full launch of
    // C3-scale graphs; a larger launch grows them
    HIPC(L.b.data.ensure(S->cfg.lane_data_bytes));
    Meta m;
    m.ng = S->cfg.max_batch;
    m.max_nr = 1 << 13;
    m.total_tasks = 1 << 20;
    mbe::KParams kp{};
    mbe_status s = setup_pass(L.b, L.stream, L.abort_host, m, S->tu, 0, {}, 0, S->tu.arena_words,
                              S->tu.workers, kp);
    if (s != MBE_OK) return s;
    HIPC(L.rem.ensure((size_t)2 * S->cfg.max_batch * 8));
    HIPC(L.b.graphs.ensure(sizeof(mbe::DevGraph) * S->cfg.max_batch));
    return MBE_OK;
}

}  // namespace

extern "C" {

mbe_status mbe_server_create(const mbe_server_config* cfg, mbe_server** out) {
    if (!out) return fail(MBE_EINVAL, "null argument");
    auto S = std::make_unique<mbe_server>();
    if (cfg) S->cfg = *cfg;
    if (!S->cfg.lanes) S->cfg.lanes = 4;
    if (!S->cfg.max_batch) S->cfg.max_batch = 256;
    if (!S->cfg.queue_cap) S->cfg.queue_cap = 1u << 20;
    if (!S->cfg.hbm_budget) S->cfg.hbm_budget = (uint64_t)64 << 30;
    if (!S->cfg.lane_data_bytes) S->cfg.lane_data_bytes = (uint64_t)256 << 20;
    if (S->cfg.lanes > 16) return fail(MBE_EINVAL, "at most 16 lanes");
    HIPC(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, 0));
    S->cus = prop.multiProcessorCount;
    const size_t lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                    : 160 * 1024;
    S->tu = resolve(&S->cfg.tuning, S->cus, lds_per_cu);
    if (S->tu.side < -1 || S->tu.side > 1) return fail(MBE_EINVAL, "bad side");
    if (S->tu.root_begin || S->tu.root_end || S->tu.root_stride > 1)
        return fail(MBE_EINVAL, "root subsets are not served");
    mbe_status s = ensure_lds(S->tu.lds_words);
    if (s != MBE_OK) return s;
    for (uint32_t i = 0; i < S->cfg.lanes; i++) {
        S->lanes.push_back(std::make_unique<Lane>());
        s = make_lane(S.get(), i, *S->lanes.back());
        if (s != MBE_OK) {
            for (auto& L : S->lanes) free_lane(*L);
            return s;
        }
    }
    mbe_server* raw = S.release();
    raw->th = std::thread([raw] { raw->run(); });
    *out = raw;
    return MBE_OK;
}

mbe_status mbe_server_submit(mbe_server* S, const mbe_graph* g, mbe_variant variant,
                             uint64_t* ticket) {
    if (!S || !g || !ticket || (unsigned)variant > 3u) return fail(MBE_EINVAL, "bad argument");
    auto r = std::make_shared<Req>();
    r->t_submit = now_ms();
    mbe::prepare(g->g, S->tu.side, r->prep);
    r->variant = (int)variant;
    const uint64_t stack = root_bound_words(r->prep) * 8 *
                           std::min<uint64_t>(r->prep.n_roots, S->tu.workers);
    r->footprint = packed_bytes(r->prep) + stack;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(MBE_EINVAL, "server is shutting down");
    S->st.submitted++;
    if (r->footprint > S->cfg.hbm_budget) {
        S->st.rejected++;
        return fail(MBE_EOVERBUDGET, "request footprint exceeds the HBM budget");
    }
    if (S->queue.size() >= S->cfg.queue_cap) {
        S->st.rejected++;
        return fail(MBE_EBUSY, "server queue full");
    }
    r->id = S->next_id++;
    S->queue.push_back(r);
    S->live[r->id] = r;
    S->st.queued++;
    *ticket = r->id;
    S->work_cv.notify_one();
    return MBE_OK;
}

mbe_status mbe_server_poll(mbe_server* S, uint64_t ticket, int wait_ms, mbe_request_result* out) {
    if (!S || !out) return fail(MBE_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) return fail(MBE_EINVAL, "unknown ticket");
    ReqP r = it->second;
    if (!r->done && wait_ms != 0) {
        auto ready = [&] { return r->done; };
        if (wait_ms < 0)
            S->done_cv.wait(lk, ready);
        else
            S->done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    }
    if (!r->done) return MBE_ENOTREADY;
    *out = r->res;
    S->live.erase(ticket);
    return MBE_OK;
}

mbe_status mbe_server_get_stats(mbe_server* S, mbe_server_stats* out) {
    if (!S || !out) return fail(MBE_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    return MBE_OK;
}

void mbe_server_destroy(mbe_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->work_cv.notify_all();
    }
    if (S->th.joinable()) S->th.join();
    for (auto& L : S->lanes) free_lane(*L);
    delete S;
}

}  // extern "C"
