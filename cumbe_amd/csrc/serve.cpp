// Serving loop of libmbe.so: one long-lived persistent kernel per generation
// takes requests while it runs (continuous batching). The host admits
// requests against an HBM budget, packs each request's CSR into pinned
// staging memory, places it in a device data ring and publishes a mailbox
// record; an idle wave of the running kernel loads it and appends its roots
// to the shared task space, where they mix with every other request's work
// and donated subtrees. Each request completes when the kernel raises its
// host-mapped completion flag.
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
    std::unique_ptr<mbe_graph> graph;  // kept for requests served by the engine path
    int variant = 0;
    uint64_t footprint = 0;
    uint64_t words = 0;                // packed CSR size (u64 words)
    bool oversize = false;
    double t_submit = 0, t_start = 0, t_done = 0;
    uint64_t slot = 0, data_pos = 0, stage_pos = 0;
    mbe_request_result res{};
    bool done = false;
};
using ReqP = std::shared_ptr<Req>;

// Circular allocator in u64 words with in-order reclaim: blocks never wrap
// (a block that would cross the end starts at the next lap), and space comes
// back when the oldest live block is released.
struct Ring {
    uint64_t cap = 0, head = 0, tail = 0;
    struct Block {
        uint64_t start, end;
        bool freed;
    };
    std::deque<Block> blocks;
    bool alloc(uint64_t n, uint64_t& pos) {
        if (n > cap) return false;
        uint64_t start = head;
        if (start % cap + n > cap) start += cap - start % cap;
        if (start + n - tail > cap) return false;
        blocks.push_back({start, start + n, false});
        head = start + n;
        pos = start;
        return true;
    }
    void release(uint64_t pos) {
        for (auto& b : blocks)
            if (b.start == pos && !b.freed) {
                b.freed = true;
                break;
            }
        while (!blocks.empty() && blocks.front().freed) blocks.pop_front();
        tail = blocks.empty() ? head : blocks.front().start;
    }
};

// Largest arena (u64 words) the local problem of any root of p can take:
// the degree-prefix and candidate lists of root localisation, then hash
// terms, both bit matrices, scratch lists and a full-depth frame stack (the
// kernel's layout, before Q reduction shrinks it), and the rows of every
// candidate with the kept-Q list that localisation builds at the arena's
// tail, with the distance-two candidate count bounded by the sum of the
// root's neighbours' other edges.
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
        const uint64_t lists = (d + 2) / 2 + std::max<uint64_t>(d, (nc + 1) / 2) + 1;
        const uint64_t tail = nc * W + (nc + 1) / 2 + 1;
        best = std::max(best, lists + data + scratch + frames + tail);
    }
    return best;
}

template <class T>
hipError_t host_mapped(T*& p, size_t bytes) {
    return hipHostMalloc((void**)&p, bytes, hipHostMallocMapped | hipHostMallocCoherent);
}

}  // namespace

struct mbe_server {
    mbe_server_config cfg{};
    Tune tu{};
    std::mutex mu;
    std::condition_variable done_cv, work_cv;
    std::deque<ReqP> queue;              // admitted later, in arrival order
    std::map<uint64_t, ReqP> live;       // submitted, not yet collected
    std::map<uint64_t, ReqP> inflight;   // published to the kernel, by slot
    mbe_server_stats st{};
    std::thread th;
    bool stop = false;
    bool failed = false;
    std::string fail_msg;

    // device side
    hipStream_t stream = nullptr;
    hipEvent_t ev_end = nullptr;
    Bufs b;
    DevBuf gtable, rem, sctl, data;
    mbe::KParams kp{};
    bool running = false, closing = false;
    uint64_t seq = 0;                    // records published in this generation
    // host-mapped
    mbe::MailRec* mail = nullptr;
    unsigned long long *mail_pub = nullptr, *mail_taken = nullptr, *close_flag = nullptr;
    unsigned long long *done = nullptr, *abort_flag = nullptr;
    uint64_t* staging = nullptr;
    Ring data_ring, stage_ring;
    std::deque<std::pair<uint64_t, uint64_t>> staged;  // (record seq, staging pos)

    mbe_status launch_generation() {
        *mail_pub = 0;
        *mail_taken = 0;
        *close_flag = 0;
        *abort_flag = 0;
        memset(done, 0, (size_t)cfg.table_cap * 32);
        Meta m;
        m.ng = cfg.table_cap;
        m.max_nr = cfg.max_nr;
        m.total_tasks = 1u << 20;  // overflow-list capacity
        mbe_status s = setup_pass(b, stream, abort_flag, m, tu, 0, {}, 0, tu.arena_words,
                                  tu.workers, kp);
        if (s != MBE_OK) return s;
        kp.graphs = gtable.as<mbe::DevGraph>();
        kp.n_graphs = 0;
        kp.remaining = rem.as<unsigned long long>();
        kp.root_cursor = rem.as<unsigned long long>() + 2 * (size_t)cfg.table_cap;
        kp.lane_list = (uint32_t*)(rem.as<unsigned long long>() + 3 * (size_t)cfg.table_cap);
        kp.serve = sctl.as<mbe::ServeCtl>();
        kp.mail_cap = cfg.table_cap;
        kp.data = data.as<uint64_t>();
        HIPC(hipHostGetDevicePointer((void**)&kp.done, done, 0));
        HIPC(hipHostGetDevicePointer((void**)&kp.mail, mail, 0));
        HIPC(hipHostGetDevicePointer((void**)&kp.mail_pub, mail_pub, 0));
        HIPC(hipHostGetDevicePointer((void**)&kp.mail_taken, mail_taken, 0));
        HIPC(hipHostGetDevicePointer((void**)&kp.close, close_flag, 0));
        HIPC(hipHostGetDevicePointer((void**)&kp.staging, staging, 0));
        HIPC(mbe::launch_fill((unsigned long long*)sctl.p, sizeof(mbe::ServeCtl) / 8, 0ull, stream));
        HIPC(enqueue_pass(kp, tu.workers, m.ctl_words(), stream));
        HIPC(hipEventRecord(ev_end, stream));
        running = true;
        closing = false;
        seq = 0;
        staged.clear();
        stage_ring.head = stage_ring.tail = 0;
        stage_ring.blocks.clear();
        st.launches++;
        return MBE_OK;
    }

    void finish(const ReqP& r, mbe_status s, uint64_t count, uint64_t hash, uint32_t route) {
        r->t_done = now_ms();
        r->res.status = s;
        r->res.route = route;
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

    // Completion flags of published requests; staging space the loader has
    // copied out. Returns true if anything completed.
    bool reap() {
        bool any = false;
        volatile unsigned long long* d = done;
        for (auto it = inflight.begin(); it != inflight.end();) {
            const uint64_t g = it->first;
            if (!d[4 * g + 2]) {
                ++it;
                continue;
            }
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
            const ReqP r = it->second;
            finish(r, d[4 * g + 3] ? MBE_ENOMEM : MBE_OK, d[4 * g], d[4 * g + 1], 0);
            data_ring.release(r->data_pos);
            it = inflight.erase(it);
            any = true;
        }
        const uint64_t taken = __atomic_load_n(mail_taken, __ATOMIC_ACQUIRE);
        while (!staged.empty() && staged.front().first < taken) {
            stage_ring.release(staged.front().second);
            staged.pop_front();
        }
        return any;
    }

    // Publish queued requests to the running kernel while the budget, the
    // rings and the generation's table allow. The lock is released while a
    // request's CSR is packed into staging.
    bool publish(std::unique_lock<std::mutex>& lk) {
        bool any = false;
        while (!queue.empty() && running && !closing) {
            const ReqP r = queue.front();
            if (r->oversize) break;  // runs between generations
            if (st.bytes_in_use + r->footprint > cfg.hbm_budget) break;
            if (seq >= cfg.table_cap) break;  // generation full: rotate
            uint64_t dpos, spos;
            if (!data_ring.alloc(r->words, dpos)) break;
            if (!stage_ring.alloc(r->words, spos)) {
                data_ring.release(dpos);
                break;
            }
            queue.pop_front();
            st.queued--;
            st.in_flight++;
            st.bytes_in_use += r->footprint;
            r->slot = seq++;
            r->data_pos = dpos;
            r->stage_pos = spos;
            r->t_start = now_ms();
            lk.unlock();
            pack(*r);
            lk.lock();
            inflight[r->slot] = r;
            staged.emplace_back(r->slot, spos);
            __atomic_store_n(mail_pub, seq, __ATOMIC_RELEASE);
            any = true;
        }
        return any;
    }

    // Pack the prepared CSR into staging and fill its mailbox record.
    void pack(const Req& r) {
        const mbe::Prepared& p = r.prep;
        uint64_t* base = staging + r.stage_pos % stage_ring.cap;
        uint64_t o = 0;
        uint64_t offs[6];
        auto put = [&](int k, const void* src, size_t bytes) {
            offs[k] = o;
            if (bytes) memcpy(base + o, src, bytes);
            o += upload_slot(bytes) / 8;
        };
        put(0, p.offR.data(), p.offR.size() * 4);
        put(1, p.adjR.data(), p.adjR.size() * 4);
        put(2, p.offL.data(), p.offL.size() * 4);
        put(3, p.adjL.data(), p.adjL.size() * 4);
        put(4, p.hR.data(), p.hR.size() * 8);
        put(5, p.hL.data(), p.hL.size() * 8);
        mbe::MailRec& m = mail[r.slot % cfg.table_cap];
        m.src = r.stage_pos % stage_ring.cap;
        m.dst = r.data_pos % data_ring.cap;
        m.words = o;
        for (int k = 0; k < 6; k++) m.off[k] = offs[k];
        m.nR = p.nR;
        m.nL = p.nL;
        m.n_roots = p.n_roots;
        m.r_is_u = p.r_is_u ? 1 : 0;
        m.lane = p.adjR.size() >= cfg.large_edges ? 1 : 0;
    }

    void fail_all(mbe_status s) {
        for (auto& kv : inflight) finish(kv.second, s, 0, 0, 0);
        inflight.clear();
        while (!queue.empty()) {
            st.queued--;
            st.in_flight++;
            st.bytes_in_use += queue.front()->footprint;
            finish(queue.front(), s, 0, 0, 0);
            queue.pop_front();
        }
    }

    void run() {
        (void)hipSetDevice(0);
        std::unique_lock<std::mutex> lk(mu);
        double closed_at = 0;
        for (;;) {
            bool any = false;
            if (running) {
                any |= reap();
                any |= publish(lk);
                // close the generation to finish (stop), to start a new one
                // (table full) or to run an engine-path request
                const bool head_big = !queue.empty() && queue.front()->oversize;
                const bool drain = stop || head_big || (seq >= cfg.table_cap && !queue.empty());
                if (drain && inflight.empty() && !closing) {
                    // rotate the generation (table full) or finish (stop)
                    __atomic_store_n(close_flag, 1ull, __ATOMIC_RELEASE);
                    closing = true;
                    closed_at = now_ms();
                }
                if (closing) {
                    const hipError_t q = hipEventQuery(ev_end);
                    if (q != hipErrorNotReady) {
                        running = false;
                        if (q != hipSuccess) {
                            failed = true;
                            fail_msg = hipGetErrorString(q);
                            fail_all(MBE_EHIP);
                        }
                        any = true;
                    } else if (now_ms() - closed_at > 10000.0) {
                        __atomic_store_n(abort_flag, 1ull, __ATOMIC_RELEASE);
                    }
                } else {
                    const hipError_t q = hipEventQuery(ev_end);
                    if (q != hipErrorNotReady) {  // the kernel ended by itself: a fault
                        running = false;
                        failed = true;
                        fail_msg = q == hipSuccess ? "serving kernel exited" : hipGetErrorString(q);
                        fail_all(MBE_EHIP);
                    }
                }
            } else if (!failed && !queue.empty() && queue.front()->oversize) {
                run_engine(lk);
                any = true;
            } else if (!failed && !queue.empty()) {
                if (launch_generation() != MBE_OK) {
                    failed = true;
                    fail_msg = "serving kernel launch failed";
                    fail_all(MBE_EHIP);
                }
                any = true;
            }
            if (stop && !running && queue.empty() && inflight.empty()) break;
            if (any) continue;
            if (running) {
                lk.unlock();
                usleep(20);
                lk.lock();
            } else {
                work_cv.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }

    // A request the kernel does not serve (another variant, local problems
    // beyond the serving workers' arenas, or an enumerated side beyond the
    // kernel's tables) runs through mbe_solve while no generation is on the
    // device, so that engine buffers can grow without waiting on the kernel.
    void run_engine(std::unique_lock<std::mutex>& lk) {
        ReqP r = queue.front();
        queue.pop_front();
        st.queued--;
        st.in_flight++;
        st.bytes_in_use += r->footprint;
        r->t_start = now_ms();
        lk.unlock();
        mbe_tuning t{};
        t.side = tu.side;
        mbe_result out{};
        const mbe_status s = mbe_solve(r->graph.get(), (mbe_variant)r->variant, &t, &out);
        lk.lock();
        finish(r, s, out.count, out.hash, 1);
    }
};

namespace {

void free_server(mbe_server* S) {
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    S->b.release();
    for (DevBuf* d : {&S->gtable, &S->rem, &S->sctl, &S->data}) {
        if (d->p) (void)hipFree(d->p);
        d->p = nullptr;
    }
    for (void* h : {(void*)S->mail, (void*)S->mail_pub, (void*)S->mail_taken, (void*)S->close_flag,
                    (void*)S->done, (void*)S->abort_flag, (void*)S->staging})
        if (h) (void)hipHostFree(h);
    if (S->ev_end) (void)hipEventDestroy(S->ev_end);
    if (S->stream) (void)hipStreamDestroy(S->stream);
}

mbe_status make_server(mbe_server* S) {
    HIPC(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, 0));
    const size_t lds_per_cu = prop.maxSharedMemoryPerMultiProcessor
                                  ? prop.maxSharedMemoryPerMultiProcessor
                                  : 160 * 1024;
    // serving defaults: 8 MiB arenas per worker, so that every root of the
    // batched size distribution fits the kernel, and four 32 KiB workers per
    // CU (the large requests' local problems stay in LDS; six 24 KiB workers
    // raised their isolated latency by about 40 %)
    if (!S->cfg.tuning.arena_words) S->cfg.tuning.arena_words = 1u << 20;
    if (!S->cfg.tuning.lds_words) S->cfg.tuning.lds_words = 4096;
    if (!S->cfg.tuning.waves_per_cu && !S->cfg.tuning.workers) S->cfg.tuning.waves_per_cu = 4;
    S->tu = resolve(&S->cfg.tuning, prop.multiProcessorCount, lds_per_cu);
    if (S->tu.side < -1 || S->tu.side > 1) return fail(MBE_EINVAL, "bad side");
    if (S->tu.root_begin || S->tu.root_end || S->tu.root_stride > 1)
        return fail(MBE_EINVAL, "root subsets are not served");
    mbe_status s = ensure_lds(S->tu.lds_words);
    if (s != MBE_OK) return s;
    const uint64_t cap = S->cfg.table_cap;
    HIPC(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&S->ev_end, hipEventDisableTiming));
    HIPC(S->gtable.ensure(cap * sizeof(mbe::DevGraph)));
    // per position: open tasks, error bits, root cursor, two lane entries
    HIPC(S->rem.ensure(cap * 32));
    HIPC(S->sctl.ensure(sizeof(mbe::ServeCtl)));
    HIPC(S->data.ensure(S->cfg.data_bytes));
    HIPC(host_mapped(S->mail, cap * sizeof(mbe::MailRec)));
    HIPC(host_mapped(S->mail_pub, 64));
    HIPC(host_mapped(S->mail_taken, 64));
    HIPC(host_mapped(S->close_flag, 64));
    HIPC(host_mapped(S->abort_flag, 64));
    HIPC(host_mapped(S->done, cap * 32));
    HIPC(host_mapped(S->staging, S->cfg.staging_bytes));
    S->data_ring.cap = S->cfg.data_bytes / 8;
    S->stage_ring.cap = S->cfg.staging_bytes / 8;
    return MBE_OK;
}

}  // namespace

extern "C" {

mbe_status mbe_server_create(const mbe_server_config* cfg, mbe_server** out) {
    if (!out) return fail(MBE_EINVAL, "null argument");
    auto S = std::make_unique<mbe_server>();
    if (cfg) S->cfg = *cfg;
    if (!S->cfg.table_cap) S->cfg.table_cap = 1u << 18;
    if (!S->cfg.queue_cap) S->cfg.queue_cap = 1u << 20;
    if (!S->cfg.large_edges) S->cfg.large_edges = 1u << 18;
    if (!S->cfg.max_nr) S->cfg.max_nr = 1u << 17;
    if (!S->cfg.hbm_budget) S->cfg.hbm_budget = (uint64_t)64 << 30;
    if (!S->cfg.data_bytes) S->cfg.data_bytes = (uint64_t)4 << 30;
    if (!S->cfg.staging_bytes) S->cfg.staging_bytes = (uint64_t)1 << 30;
    mbe_status s = make_server(S.get());
    if (s != MBE_OK) {
        free_server(S.get());
        return s;
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
    // footprint: the packed CSR plus the worst-case stack, i.e. the largest
    // local problem (bit matrices and frame stack) one of its roots can build;
    // the per-worker arenas that hold those stacks are allocated once
    const uint64_t bound = root_bound_words(r->prep);
    r->footprint = packed_bytes(r->prep) + bound * 8;
    r->words = packed_bytes(r->prep) / 8;
    // the kernel serves one variant (cuMBE); requests for another variant,
    // or too big for its arenas and tables, take the engine path
    r->oversize = bound > S->tu.arena_words || r->prep.nR > S->cfg.max_nr ||
                  r->words * 8 > S->cfg.data_bytes / 4 || r->words * 8 > S->cfg.staging_bytes / 4 ||
                  variant != MBE_CUMBE;
    if (r->oversize) {
        r->graph = std::make_unique<mbe_graph>();
        r->graph->g = g->g;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(MBE_EINVAL, "server is shutting down");
    if (S->failed) return fail(MBE_EHIP, "serving kernel failed: " + S->fail_msg);
    S->st.submitted++;
    if (r->footprint > S->cfg.hbm_budget) {
        S->st.rejected++;
        return fail(MBE_EOVERBUDGET, "request footprint exceeds the HBM budget");
    }
    if (S->queue.size() >= S->cfg.queue_cap) {
        S->st.rejected++;
        return fail(MBE_EBUSY, "server queue full");
    }
    r->id = S->st.submitted;
    S->queue.push_back(r);
    S->live[r->id] = r;
    S->st.queued++;
    *ticket = r->id;
    S->work_cv.notify_all();
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
    free_server(S);
    delete S;
}

}  // extern "C"
