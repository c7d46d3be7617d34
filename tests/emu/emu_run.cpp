// Runs the engine's persistent enumeration kernel (cumbe_amd/csrc/
// mbe_kernel.hip, compiled for the CPU with wave_emu.h) on one graph --
// TEST INFRASTRUCTURE ONLY. Mirrors the host engine's pass loop (buffers,
// root tasks, overflow passes with larger arenas) in host memory and prints
//   count <n> hash <0x...> nodes <n> donated <n> stolen <n> launches <n>
// Usage:
//   emu_run [options] bi <path>
//   emu_run [options] cl <n_u> <n_v> <m> <gamma_u> <gamma_v> <max_deg> <seed>
//   emu_run [options] er <n_u> <n_v> <p> <n_planted> <lo> <hi> <top_frac> <seed>
//   emu_run [options] batch <seed>
//   emu_run [options] serve <request>...   (the serving kernel: a host thread
//           publishes each request's packed CSR to the kernel's mailbox, then
//           closes it; prints "req <i> count <n> hash <0x...> err <e>" per
//           request; a request is b<seed> (a batch graph) or
//           c<n_u>,<n_v>,<m>,<gamma_u>,<gamma_v>,<max_deg>,<seed>)
// Options: -v variant(0..3) -w workers -l lds_words -a arena_words -s slots
//          -c list_cutoff -r compact_ratio -i donate_interval -t twins(0|2)
//          -b root_begin -e root_end -k root_stride -S side(-1|0|1)
//          -p sparse_ratio (list-mode roots; 0 = never, 1 = nearly always)
//          -x split_nodes (long tasks donate while the deque is short; 0 = never)
//          -L large_edges (serve: requests with at least this many edges take
//             the large claim lane) -T table_cap (serve) -u publish gap (us)
//          -G seed of the garbage that fills uncleared device buffers
//          -P s: print the workers' progress every s seconds (stderr)
#include "wave_emu.h"
#include "../../cumbe_amd/csrc/mbe_kernel.hip"

#include <atomic>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../cumbe_amd/csrc/graph.h"

namespace mbe {
namespace {
__thread u64 lds_mem[20480];  // 160 KiB: the dynamic LDS of one wave
}
}  // namespace mbe

using namespace mbe;

struct Opt {
    int variant = 0, side = -1;
    uint32_t sparse = 0, split = 0;
    uint32_t large_edges = 1u << 18, table_cap = 64, gap_us = 0;
    uint32_t workers = 8, lds_words = 3072, slots = 64, cutoff = 0, ratio = 2, interval = 64,
             twins = 1, rb = 0, re = 0, rs = 1;
    uint64_t arena_words = 1u << 18;
};

template <int V>
static void wave_fn(void* p) { mbe_persistent<V>(*(KParams*)p); }

// Words of the emulated LDS past the launch's dynamic LDS hold this pattern
// while a wave runs; a write beyond the budget changes one and is reported.
constexpr u64 kLdsCanary = 0xA5C3F00DDEADBEEFull;
constexpr size_t kLdsEmuWords = sizeof(lds_mem) / sizeof(u64);

// With -P s: print the workers' finished roots and nodes (their progress
// slots) every s seconds while a pass runs, to tell a slow run from a stuck one.
static uint32_t g_progress_s = 0;

static void run_pass(KParams& kp, uint32_t workers) {
    void (*fn)(void*) = kp.variant == 1 ? wave_fn<1> : kp.variant == 2 ? wave_fn<2>
                      : kp.variant == 3 ? wave_fn<3> : wave_fn<0>;
    std::atomic<bool> pass_done{false};
    std::thread mon;
    if (g_progress_s && kp.progress)
        mon = std::thread([&] {
            const auto t0 = std::chrono::steady_clock::now();
            while (!pass_done.load()) {
                for (uint32_t i = 0; i < g_progress_s * 10 && !pass_done.load(); i++)
                    std::this_thread::sleep_for(std::chrono::milliseconds(100));
                if (pass_done.load()) break;
                unsigned long long r = 0, nd = 0;
                for (uint32_t w = 0; w < workers; w++) {
                    r += __atomic_load_n(&kp.progress[2 * w], __ATOMIC_RELAXED);
                    nd += __atomic_load_n(&kp.progress[2 * w + 1], __ATOMIC_RELAXED);
                }
                const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                fprintf(stderr, "progress %.0f s: roots %llu nodes %llu\n", s, r, nd);
            }
        });
    if (kp.lds_words > kLdsEmuWords) {
        fprintf(stderr, "emu_run: lds_words %u beyond the emulated %zu\n", kp.lds_words,
                kLdsEmuWords);
        exit(3);
    }
    std::vector<std::thread> th;
    for (uint32_t w = 0; w < workers; w++)
        th.emplace_back([&kp, w, workers, fn] {
            for (size_t i = kp.lds_words; i < kLdsEmuWords; i++) lds_mem[i] = kLdsCanary;
            emu::run_wave(w, workers, fn, &kp);
            for (size_t i = kp.lds_words; i < kLdsEmuWords; i++)
                if (lds_mem[i] != kLdsCanary) {
                    fprintf(stderr, "emu_run: wave %u wrote LDS word %zu beyond lds_words %u\n",
                            w, i, kp.lds_words);
                    abort();
                }
        });
    for (auto& t : th) t.join();
    pass_done.store(true);
    if (mon.joinable()) mon.join();
}

static void* zalloc(size_t bytes) {
    void* p = calloc(1, bytes + 256);
    if (!p) {
        fprintf(stderr, "emu_run: out of memory (%zu bytes)\n", bytes);
        exit(3);
    }
    return p;
}

// Device memory the host does not clear (hipMalloc'd buffers the kernels
// must write before they read): filled with pseudo-random words, so that a
// read of a word never written in this run sees garbage, as on the GPU.
static uint64_t g_fill_seed = 1;
static void* dalloc(size_t bytes) {
    uint64_t* p = (uint64_t*)zalloc(bytes);
    uint64_t x = g_fill_seed * 0x9E3779B97F4A7C15ull + bytes;
    for (size_t i = 0; i < (bytes + 7) / 8; i++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        p[i] = x;
    }
    return p;
}

static size_t up_slot(size_t n) { return (n + 8 + 255) & ~(size_t)255; }

// The serving kernel over a list of requests, mirroring the server's device
// set-up (cumbe_amd/csrc/serve.cpp: launch_generation, pack).
static int serve_main(const Opt& o, int n, char** specs) {
    std::vector<Prepared> ps(n);
    for (int r = 0; r < n; r++) {
        const std::string s = specs[r];
        Graph g;
        if (s[0] == 'b') {
            g = gen_batch_graph(atoll(s.c_str() + 1));
        } else if (s[0] == 'c') {
            unsigned nu, nv, md;
            unsigned long long m, seed;
            double gu, gv;
            if (sscanf(s.c_str() + 1, "%u,%u,%llu,%lf,%lf,%u,%llu", &nu, &nv, &m, &gu, &gv, &md,
                       &seed) != 7) {
                fprintf(stderr, "bad request %s\n", s.c_str());
                return 2;
            }
            g = gen_chung_lu(nu, nv, m, gu, gv, md, seed);
        } else {
            fprintf(stderr, "bad request %s\n", s.c_str());
            return 2;
        }
        prepare(g, o.side, ps[r]);
    }
    const uint32_t cap = o.table_cap;
    if ((uint32_t)n > cap) {
        fprintf(stderr, "more requests than table_cap\n");
        return 2;
    }
    // packed CSR words per request, in the server's layout
    std::vector<uint64_t> words(n), start(n);
    uint64_t total = 0, max_nr = 1;
    for (int r = 0; r < n; r++) {
        const Prepared& p = ps[r];
        words[r] = (up_slot(p.offR.size() * 4) + up_slot(p.adjR.size() * 4) +
                    up_slot(p.offL.size() * 4) + up_slot(p.adjL.size() * 4) +
                    up_slot(p.hR.size() * 8) + up_slot(p.hL.size() * 8)) / 8;
        start[r] = total;
        total += words[r];
        max_nr = std::max<uint64_t>(max_nr, p.nR);
    }
    uint64_t* staging = (uint64_t*)zalloc(total * 8);
    uint64_t* data = (uint64_t*)dalloc(total * 8);
    MailRec* mail = (MailRec*)zalloc(cap * sizeof(MailRec));
    unsigned long long* flags = (unsigned long long*)zalloc(8 * 8);  // pub, taken, close, abort
    unsigned long long* done = (unsigned long long*)zalloc((size_t)cap * 32);
    unsigned long long* rem = (unsigned long long*)dalloc((size_t)cap * 32);
    DevGraph* gtable = (DevGraph*)dalloc(cap * sizeof(DevGraph));
    ServeCtl* sctl = (ServeCtl*)zalloc(sizeof(ServeCtl));
    const uint64_t mark_words = (max_nr + 63) / 64 + 1, cpos_len = max_nr + 1;
    const size_t ctl_words = 2 + kStatCount + kSecCount + 2 * (size_t)cap;
    unsigned long long* ctl = (unsigned long long*)zalloc(ctl_words * 8);
    unsigned long long* dq = (unsigned long long*)zalloc((4 + o.slots) * 8);
    for (uint32_t s = 0; s < o.slots; s++) dq[4 + s] = s;
    const uint32_t workers = o.workers;
    const uint64_t arena_words = o.arena_words;
    const uint64_t slot_words = std::min<uint64_t>(std::max<uint64_t>(4096, arena_words), 1u << 19);
    const uint64_t ovf_cap = 1u << 20;
    KParams kp{};
    kp.graphs = gtable;
    kp.n_graphs = 0;
    kp.variant = (uint32_t)o.variant;
    kp.cursor = ctl;
    kp.overflow_count = ctl + 1;
    kp.stats = ctl + 2;
    kp.results = ctl + 2 + kStatCount + kSecCount;
    kp.arena = (uint64_t*)dalloc(arena_words * workers * 8);
    kp.arena_words = arena_words;
    kp.mark = (uint64_t*)zalloc(mark_words * workers * 8);
    kp.mark_words = mark_words;
    kp.cpos = (uint32_t*)dalloc(cpos_len * workers * 4);
    kp.cpos_len = cpos_len;
    kp.lds_words = o.lds_words;
    kp.list_cutoff = o.cutoff;
    kp.compact_ratio = o.ratio;
    kp.twins = o.twins;
    kp.sparse_ratio = o.sparse;
    kp.split_nodes = o.split;
    kp.abort = flags + 3;
    kp.node_count = kp.stats + kStatNodeBudget;
    kp.overflow_list = (uint64_t*)dalloc(ovf_cap * 8);
    kp.overflow_cap = ovf_cap;
    kp.slots = (uint64_t*)dalloc((size_t)o.slots * slot_words * 8);
    kp.slot_words = slot_words;
    kp.n_slots = o.slots;
    kp.donate_interval = o.interval;
    kp.dq = dq;
    kp.progress = (unsigned long long*)zalloc(workers * 16);
    kp.remaining = rem;
    kp.root_cursor = rem + 2 * (size_t)cap;
    kp.lane_list = (uint32_t*)(rem + 3 * (size_t)cap);
    kp.serve = sctl;
    kp.mail = mail;
    kp.mail_cap = cap;
    kp.mail_pub = flags;
    kp.mail_taken = flags + 1;
    kp.staging = staging;
    kp.data = data;
    kp.close = flags + 2;
    kp.done = done;
    std::thread host([&] {
        for (int r = 0; r < n; r++) {
            if (o.gap_us) std::this_thread::sleep_for(std::chrono::microseconds(o.gap_us));
            const Prepared& p = ps[r];
            uint64_t* base = staging + start[r];
            uint64_t off = 0, offs[6];
            auto put = [&](int k, const void* src, size_t bytes) {
                offs[k] = off;
                if (bytes) memcpy(base + off, src, bytes);
                off += up_slot(bytes) / 8;
            };
            put(0, p.offR.data(), p.offR.size() * 4);
            put(1, p.adjR.data(), p.adjR.size() * 4);
            put(2, p.offL.data(), p.offL.size() * 4);
            put(3, p.adjL.data(), p.adjL.size() * 4);
            put(4, p.hR.data(), p.hR.size() * 8);
            put(5, p.hL.data(), p.hL.size() * 8);
            MailRec& m = mail[r];
            m.src = start[r];
            m.dst = start[r];
            m.words = off;
            for (int k = 0; k < 6; k++) m.off[k] = offs[k];
            m.nR = p.nR;
            m.nL = p.nL;
            m.n_roots = p.n_roots;
            m.r_is_u = p.r_is_u ? 1 : 0;
            m.lane = p.adjR.size() >= o.large_edges ? 1 : 0;
            __atomic_store_n(flags, (unsigned long long)(r + 1), __ATOMIC_RELEASE);
        }
        while (__atomic_load_n(flags + 1, __ATOMIC_ACQUIRE) < (unsigned long long)n)
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        __atomic_store_n(flags + 2, 1ull, __ATOMIC_RELEASE);
    });
    run_pass(kp, workers);
    host.join();
    int bad = 0;
    for (int r = 0; r < n; r++) {
        const unsigned long long* d = done + 4 * r;
        printf("req %d count %llu hash 0x%016llx err %llu done %llu\n", r, d[0], d[1], d[3], d[2]);
        bad |= d[2] != 1 || d[3] != 0;
    }
    const unsigned long long* st = ctl + 2;
    printf("nodes %llu donated %llu stolen %llu dropped %llu\n", st[kStatNodes], st[kStatDonated],
           st[kStatStolen], st[kStatDropped]);
    return bad ? 5 : 0;
}

int main(int argc, char** argv) {
    Opt o;
    int i = 1;
    for (; i < argc && argv[i][0] == '-' && argv[i][1] && !isdigit(argv[i][1]); i += 2) {
        const char f = argv[i][1];
        const long long v = atoll(argv[i + 1]);
        switch (f) {
            case 'v': o.variant = (int)v; break;
            case 'w': o.workers = (uint32_t)v; break;
            case 'l': o.lds_words = (uint32_t)v; break;
            case 'a': o.arena_words = (uint64_t)v; break;
            case 's': o.slots = (uint32_t)v; break;
            case 'c': o.cutoff = (uint32_t)v; break;
            case 'r': o.ratio = (uint32_t)v; break;
            case 'i': o.interval = (uint32_t)v; break;
            case 't': o.twins = (uint32_t)v; break;
            case 'b': o.rb = (uint32_t)v; break;
            case 'e': o.re = (uint32_t)v; break;
            case 'k': o.rs = (uint32_t)v; break;
            case 'S': o.side = (int)v; break;
            case 'p': o.sparse = (uint32_t)v; break;
            case 'x': o.split = (uint32_t)v; break;
            case 'L': o.large_edges = (uint32_t)v; break;
            case 'T': o.table_cap = (uint32_t)v; break;
            case 'u': o.gap_us = (uint32_t)v; break;
            case 'G': g_fill_seed = (uint64_t)v; break;
            case 'P': g_progress_s = (uint32_t)v; break;
            default: fprintf(stderr, "unknown option -%c\n", f); return 2;
        }
    }
    if (i >= argc) return 2;
    const std::string kind = argv[i];
    if (kind == "serve") return serve_main(o, argc - i - 1, argv + i + 1);
    Graph g;
    std::string err;
    if (kind == "bi" && i + 1 < argc) {
        if (!read_bi(argv[i + 1], g, &err)) { fprintf(stderr, "%s\n", err.c_str()); return 2; }
    } else if (kind == "cl" && i + 7 < argc) {
        g = gen_chung_lu(atoi(argv[i + 1]), atoi(argv[i + 2]), atoll(argv[i + 3]), atof(argv[i + 4]),
                         atof(argv[i + 5]), atoi(argv[i + 6]), atoll(argv[i + 7]));
    } else if (kind == "er" && i + 8 < argc) {
        g = gen_er_planted(atoi(argv[i + 1]), atoi(argv[i + 2]), atof(argv[i + 3]), atoi(argv[i + 4]),
                           atoi(argv[i + 5]), atoi(argv[i + 6]), atof(argv[i + 7]), atoll(argv[i + 8]));
    } else if (kind == "batch" && i + 1 < argc) {
        g = gen_batch_graph(atoll(argv[i + 1]));
    } else {
        fprintf(stderr, "bad graph arguments\n");
        return 2;
    }
    Prepared p;
    prepare(g, o.side, p);
    DevGraph dg{};
    dg.offR = p.offR.data();
    dg.adjR = p.adjR.data();
    dg.offL = p.offL.data();
    dg.adjL = p.adjL.data();
    dg.hR = p.hR.data();
    dg.hL = p.hL.data();
    dg.nR = p.nR;
    dg.nL = p.nL;
    dg.n_roots = p.n_roots;
    dg.r_is_u = p.r_is_u;
    dg.task_base = 0;
    std::vector<uint64_t> tasks;
    const uint64_t end = o.re ? std::min<uint64_t>(o.re, p.n_roots) : p.n_roots;
    for (uint64_t r = o.rb; r < end; r += o.rs ? o.rs : 1) tasks.push_back(r);

    const uint64_t mark_words = (p.nR + 63) / 64 + 1, cpos_len = (uint64_t)p.nR + 1;
    const size_t ctl_words = 2 + kStatCount + kSecCount + 2;
    unsigned long long* ctl = (unsigned long long*)zalloc(ctl_words * 8);
    unsigned long long* dq = (unsigned long long*)zalloc((4 + o.slots) * 8);
    unsigned long long abort_flag = 0;
    uint64_t arena_words = o.arena_words;
    uint32_t workers = o.workers;
    unsigned launches = 0;
    KParams kp{};
    while (!tasks.empty()) {
        if (launches > 6) { fprintf(stderr, "root does not fit any arena\n"); return 4; }
        const uint64_t slot_words = std::min<uint64_t>(std::max<uint64_t>(4096, arena_words), 1u << 19);
        uint64_t* arena = (uint64_t*)dalloc(arena_words * workers * 8);
        uint64_t* mark = (uint64_t*)zalloc(mark_words * workers * 8);
        uint32_t* cpos = (uint32_t*)dalloc(cpos_len * workers * 4);
        uint64_t* slots = (uint64_t*)dalloc((size_t)o.slots * slot_words * 8);
        uint64_t* ovf = (uint64_t*)dalloc(tasks.size() * 8);
        unsigned long long* progress = (unsigned long long*)zalloc(workers * 16);
        for (uint32_t s = 0; s < o.slots; s++) dq[4 + s] = s;
        dq[0] = dq[1] = dq[2] = dq[3] = 0;
        ctl[0] = ctl[1] = 0;  // cursor, overflow count (stats and results accumulate)
        kp = KParams{};
        kp.graphs = &dg;
        kp.n_graphs = 1;
        kp.variant = (uint32_t)o.variant;
        kp.n_tasks = tasks.size();
        kp.task_list = tasks.data();
        kp.cursor = ctl;
        kp.overflow_count = ctl + 1;
        kp.stats = ctl + 2;
        kp.results = ctl + 2 + kStatCount + kSecCount;
        kp.arena = arena;
        kp.arena_words = arena_words;
        kp.mark = mark;
        kp.mark_words = mark_words;
        kp.cpos = cpos;
        kp.cpos_len = cpos_len;
        kp.lds_words = o.lds_words;
        kp.list_cutoff = o.cutoff;
        kp.compact_ratio = o.ratio;
        kp.twins = o.twins;
        kp.sparse_ratio = o.sparse;
        kp.split_nodes = o.split;
        kp.abort = &abort_flag;
        kp.node_count = kp.stats + kStatNodeBudget;
        kp.overflow_list = ovf;
        kp.overflow_cap = tasks.size();
        kp.slots = slots;
        kp.slot_words = slot_words;
        kp.n_slots = o.slots;
        kp.donate_interval = o.interval;
        kp.dq = dq;
        kp.progress = progress;
        run_pass(kp, workers);
        launches++;
        const uint64_t novf = ctl[1];
        std::vector<uint64_t> next(ovf, ovf + novf);
        free(arena);
        free(mark);
        free(cpos);
        free(slots);
        free(ovf);
        free(progress);
        tasks.swap(next);
        arena_words *= 16;
        workers = std::max<uint32_t>(1, std::min<uint32_t>(workers, 4));
    }
    const unsigned long long* st = ctl + 2;
    const unsigned long long* res = ctl + 2 + kStatCount + kSecCount;
    printf("count %llu hash 0x%016llx nodes %llu donated %llu stolen %llu dropped %llu launches %u\n",
           res[0], res[1], st[kStatNodes], st[kStatDonated], st[kStatStolen], st[kStatDropped],
           launches);
    return st[kStatDropped] ? 5 : 0;
}
