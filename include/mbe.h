/*
 * mbe.h -- C ABI of the MI355X maximal biclique enumeration engine (libmbe.so).
 *
 * Plain pointers and sizes only; no torch or HIP types cross this boundary.
 * Status codes are returned, never thrown; the library never aborts on bad
 * input. The caller owns every input and result buffer; the library owns
 * every device buffer.
 *
 * Reference interfaces replaced (cuMBE, paths relative to its repository):
 *   mbe_graph_load_bi ........ .bi reader in main, src/main.cu:38-48
 *   mbe_graph_write_bi ....... .bi writer, data/gen_bi.cpp:72-79
 *   mbe_graph_from_edge_list . edge list -> CSR relabel/dedupe, data/gen_bi.cpp:34-51
 *   mbe_solve ................ side swap + CUDA_TRANSPOSE + root sort + the
 *                              cooperative CUDA_MBE_<variant> launch and the
 *                              num_mb readback, src/main.cu:65-207
 *   mbe_variant .............. the argv[2] algorithm switch, src/main.cu:121-122
 *   mbe_result.count ......... the "maximal bicliques: N" line, src/main.cu:207
 *   mbe_result.kernel_ms ..... cudaEvent span around the MBE kernel,
 *                              src/main.cu:185-204
 */
#ifndef MBE_H
#define MBE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MBE_OK = 0,
    MBE_EINVAL = -1,       /* bad argument or malformed input */
    MBE_ENOMEM = -2,       /* host or device allocation failed */
    MBE_EHIP = -3,         /* a HIP call failed; see mbe_last_hip_error() */
    MBE_ETIMEOUT = -4,     /* a watchdog or poll deadline expired */
    MBE_EOVERBUDGET = -5,  /* request exceeds the admission budget */
    MBE_EIO = -6,          /* file could not be read or written */
    MBE_EBUSY = -7,        /* server queue full; retry later */
    MBE_ENOTREADY = -8,    /* poll: request not finished yet */
    MBE_ERANGE = -9,       /* output buffer too small; the size needed is returned */
    MBE_EINTERNAL = -10    /* the engine's results failed a self-check */
} mbe_status;

/* The four algorithm variants of the reference (src/main.cu:121-122). */
typedef enum {
    MBE_CUMBE = 0, /* reverse scans + early-stop selection + work stealing */
    MBE_NORS = 1,  /* forward scans over N(v) for the Q check and expansion */
    MBE_NOES = 2,  /* candidate selection scans all of P, no early stop */
    MBE_NOWS = 3   /* root pool only, no stealing below the root */
} mbe_variant;

typedef struct mbe_graph mbe_graph;   /* host CSR, sorted, deduplicated */

/* Launch tuning. Zero fields take the library default (or the autotuner's
 * cached choice when one is loaded). */
typedef struct {
    uint32_t workers;        /* persistent wave64 workers; 0 = auto */
    uint32_t waves_per_cu;   /* occupancy target used when workers == 0 */
    uint32_t lds_words;      /* per-worker LDS stack budget in u64 words */
    uint32_t arena_words;    /* per-worker HBM arena in u64 words */
    uint32_t list_cutoff;    /* P/Q size at or below which nodes use row
                                (sorted-list style) scans instead of
                                column bitmaps; 0 = default */
    uint32_t donate_depth;   /* deepest level a worker donates from; 0 = any */
    uint32_t compact_ratio;  /* re-localise a child subproblem when its universe
                                shrank by this factor; 0 = default (2),
                                UINT32_MAX = never */
    uint32_t donate_interval;/* nodes between a worker's donation checks; 0 = 64 */
    int32_t side;            /* -1 reference side choice, 0 enumerate U, 1 V */
    uint32_t flags;          /* MBE_FLAG_* */
    /* Root subset: ranks r with root_begin <= r < root_end (0 = all) and
     * (r - root_begin) % root_stride == 0 (0 = 1). Interleaving roots with a
     * stride shards one graph across devices or processes: the partial
     * counts and hashes of the shards add up to the whole. Single graph only. */
    uint32_t root_begin, root_end, root_stride;
    /* Watchdog: stop after about max_nodes tree nodes or time_limit_ms of
     * wall time (0 = none); the solve then returns MBE_ETIMEOUT and the
     * result holds the partial count, hash and node total. */
    uint64_t max_nodes;
    double time_limit_ms;
    /* List-versus-bitmap cutoff of root localisation: a root with at most 8
     * P candidates whose bit matrices would take more than sparse_ratio
     * 64-bit words per adjacency entry of its local graph is kept as sorted
     * adjacency lists instead (the reference's list scans,
     * src/mbe_cuMBE.cuh:188-236,345-390): bit rows only for its P
     * candidates, the candidates of each L vertex and the L vertices of each
     * candidate as lists. Such a root's tree (at most 256 nodes) is searched
     * by one worker. 0 = default (8), UINT32_MAX = never. */
    uint32_t sparse_ratio;
    /* A task (root or stolen subtree) that has run this many tree nodes
     * hands part of its work to the deque at its donation checks while fewer
     * than 64 tasks wait there, even when no worker is idle, so that heavy
     * roots spread before the tail. 0 = default (4096), UINT32_MAX = only
     * when workers are idle. */
    uint32_t split_nodes;
} mbe_tuning;

typedef struct {
    uint64_t count;        /* maximal bicliques */
    uint64_t hash;         /* order-independent 64-bit hash: the sum mod 2^64
                              over bicliques (A, B) of mix64(S_U ^ rotl(S_V, 29)
                              ^ (|A| K1 + |B| K2)), S_U and S_V the sums of the
                              file ids' terms (cumbe_amd/csrc/hash.h) */
    double kernel_ms;      /* HIP-event time of the enumeration kernels */
    double total_ms;       /* host wall time incl. preprocessing and copies */
    uint64_t nodes;        /* enumeration-tree nodes visited */
    uint64_t donated;      /* subtree tasks donated through the deque */
    uint32_t launches;     /* kernel launches used */
    uint32_t workers;      /* workers used */
    /* Time per kernel section summed over workers, in 10 ns ticks, filled
     * when tuning->flags has MBE_FLAG_SECTIONS, in the slots of the
     * reference's DESECTION build (src/header.cuh:25-32) as
     * utils/plot_section.py:7-8 groups them: 0 other, 1 re-localisation,
     * 2 donation (work stealing), 3 L' construction, 4 maximality check of
     * non-maximal nodes, 5 check + expansion of maximal nodes, 6 report,
     * 7 candidate selection, 8 root claim and localisation (subtree
     * fetching), 9 stolen-subtree load (work stealing), 10 idle wait. */
    uint64_t sections[11];
    /* Algorithmic work: 64-bit words of the local bit matrices and set
     * bitsets that the visited nodes read (from the set sizes, not counted by
     * hardware), and adjacency entries read while localising roots. */
    uint64_t words;
    uint64_t build_words;
} mbe_result;

#define MBE_FLAG_SECTIONS 1u
/* Also record each worker's busy time (all sections but the idle wait), the
 * reference's DEVARIANCE build (src/mbe_noES.cuh:49-55); read it with
 * mbe_last_worker_busy. Implies MBE_FLAG_SECTIONS. */
#define MBE_FLAG_VARIANCE 2u
/* Walk exactly the reference's search tree. By default a candidate whose
 * neighbourhood within the current L equals that of the vertex just
 * selected is dropped from P (iMBEA's rule, which the reference leaves
 * out): each such candidate would only yield a non-maximal leaf. The
 * bicliques reported are the same either way; the tree and the time are
 * not. In the serving loop, set it in mbe_server_config.tuning.flags. */
#define MBE_FLAG_REFERENCE_TREE 4u

typedef struct mbe_batch_plan mbe_batch_plan;  /* resident batch + captured launch */

const char* mbe_status_str(mbe_status s);
int mbe_last_hip_error(void);       /* hipError_t of the last MBE_EHIP */
const char* mbe_last_error(void);   /* human-readable detail, thread-local */
/* Per-worker busy ticks (10 ns) of the last launch of the last solve on
 * this thread that ran with MBE_FLAG_VARIANCE: copies up to cap values into
 * out and returns the number of workers. */
uint32_t mbe_last_worker_busy(uint64_t* out, uint32_t cap);

/* ------------------------------------------------------------- graphs -- */
mbe_status mbe_graph_load_bi(const char* path, mbe_graph** out);
mbe_status mbe_graph_write_bi(const mbe_graph* g, const char* path);
mbe_status mbe_graph_from_edge_list(const char* path, int skip_words, int skip_per_edge,
                                    mbe_graph** out);
/* Copies the CSR over U (rows): off_u[n_u + 1], adj_u[off_u[n_u]] with V ids.
 * Sorts and deduplicates; rejects out-of-range ids with MBE_EINVAL. */
mbe_status mbe_graph_from_csr(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                              const uint32_t* adj_u, mbe_graph** out);
mbe_status mbe_graph_info(const mbe_graph* g, uint32_t* n_u, uint32_t* n_v, uint64_t* edges);
/* Borrow the canonical CSR (valid until mbe_graph_free). */
mbe_status mbe_graph_csr(const mbe_graph* g, const uint32_t** off_u, const uint32_t** adj_u);
void mbe_graph_free(mbe_graph* g);

/* Seeded synthetic generators (SplitMix64-seeded xoshiro256**). */
mbe_status mbe_gen_chung_lu(uint32_t n_u, uint32_t n_v, uint64_t m, double gamma_u,
                            double gamma_v, uint32_t max_deg, uint64_t seed, mbe_graph** out);
mbe_status mbe_gen_er_planted(uint32_t n_u, uint32_t n_v, double p, uint32_t n_planted,
                              uint32_t lo, uint32_t hi, double top_frac, uint64_t seed,
                              mbe_graph** out);
/* One graph of the batched workload's size distribution for this seed. */
mbe_status mbe_gen_batch_graph(uint64_t seed, mbe_graph** out);

/* -------------------------------------------------------------- solve -- */
/* One graph, one stream, synchronous. tuning may be NULL. */
mbe_status mbe_solve(const mbe_graph* g, mbe_variant variant, const mbe_tuning* tuning,
                     mbe_result* out);

/* Materialised bicliques. Enumerates like mbe_solve and returns every
 * maximal biclique (A, B), A a set of rows (U) and B of columns (V) of the
 * graph, as records of uint32 words: |A|, |B|, the ids of A ascending, the
 * ids of B ascending. Records are sorted (by A's ids, then B's), so a list
 * does not depend on the variant, the tuning or the run. *used_words receives
 * the words all records take. With out == NULL (and cap_words 0) only that
 * size is computed; with cap_words < *used_words nothing is written and
 * MBE_ERANGE is returned. Each record is checked on the host: B must be the
 * common neighbourhood of A, with the size and hash the kernel reported
 * (MBE_EINTERNAL otherwise). The device holds cap_words u64 words while it
 * runs (cap_words at most 2^40, else MBE_ENOMEM). result may be NULL. No reference counterpart: the reference only
 * counts (src/mbe_cuMBE.cuh:295-300). */
mbe_status mbe_solve_list(const mbe_graph* g, mbe_variant variant, const mbe_tuning* tuning,
                          uint32_t* out, uint64_t cap_words, uint64_t* used_words,
                          mbe_result* result);

/* Initialise the engine on device 0 (stream, events, the watchdog flag and
 * the address of the allocation-free solve's scratch). Every entry point
 * but mbe_solve_noalloc does this on its first call; a caller that cannot
 * allocate later calls it once at set-up. */
mbe_status mbe_init(void);

/* A prepared graph packed into one buffer: the enumerated side R relabelled
 * by root order, both CSRs with sorted lists and the per-vertex hash terms,
 * each array at a 256-byte aligned byte offset. */
typedef struct {
    uint32_t n_r, n_l;       /* enumerated side, other side */
    uint32_t n_roots;        /* ranks [0, n_roots) have a nonzero degree */
    uint32_t r_is_u;         /* 1 when R holds the file's rows */
    uint64_t off[6];         /* byte offsets of offR, adjR, offL, adjL, hR, hL */
    uint64_t bytes;          /* bytes the packed graph takes */
} mbe_packed_graph;

/* Prepare g (side: -1 the reference's choice, 0 enumerate U, 1 V) and write
 * it packed into out. With out NULL and cap_bytes 0 only *info is filled;
 * MBE_ERANGE if cap_bytes < info->bytes. The caller copies the buffer to
 * device memory it owns for mbe_solve_noalloc. */
mbe_status mbe_graph_pack(const mbe_graph* g, int side, void* out, uint64_t cap_bytes,
                          mbe_packed_graph* info);

/* Bytes of the allocation-free solve's scratch, one static array per device. */
#define MBE_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* mbe_solve without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * dev_graph is device memory holding what mbe_graph_pack wrote (described
 * by *info); the caller owns it and the result. Every other buffer the
 * kernels use -- control words, the deque and its slots, task lists, the
 * per-worker arenas, marks and rank tables -- is carved out of one static
 * device array of MBE_NOALLOC_SCRATCH_BYTES, so the call makes no hipMalloc,
 * hipFree or other allocating call. mbe_init must have run (MBE_EINVAL
 * otherwise). Defaults: 32 K-word (256 KiB) arenas and as many workers as
 * fit, up to mbe_solve's count; roots whose local problem exceeds an arena
 * run in follow-up passes with 16 times larger arenas and fewer workers
 * inside the same array. A tuning or graph whose buffers do not fit the
 * array returns MBE_ERANGE; MBE_FLAG_VARIANCE returns MBE_EINVAL. The array
 * is shared by every call on a device and keeps the last call's contents
 * (each call clears what must start at zero); calls are serialised by the
 * engine's lock within a process, and calls that overlap on one device
 * from several processes race. */
mbe_status mbe_solve_noalloc(const void* dev_graph, const mbe_packed_graph* info,
                             mbe_variant variant, const mbe_tuning* tuning, mbe_result* out);

/* Many graphs in one persistent launch sharing one task pool; per_graph[i]
 * receives graph i's count and hash (timing fields are the batch's).
 * No reference counterpart: the reference runs one graph per process. */
mbe_status mbe_batch_solve(const mbe_graph* const* graphs, size_t n, mbe_variant variant,
                           const mbe_tuning* tuning, mbe_result* per_graph,
                           mbe_result* total);

/* Batch plan: the graphs are prepared and uploaded once into a device arena
 * owned by the plan, and the first pass (control clears, deque reset, the
 * persistent launch over every root of every graph) is captured into a HIP
 * graph that each run replays. Roots whose local problem exceeds a worker's
 * arena are finished by eager follow-up launches, as in mbe_batch_solve.
 * Runs of one plan, and plans with solves, are serialised. No reference
 * counterpart (the reference's per-graph launch is src/main.cu:185-204). */
mbe_status mbe_batch_plan_create(const mbe_graph* const* graphs, size_t n, mbe_variant variant,
                                 const mbe_tuning* tuning, mbe_batch_plan** out);
mbe_status mbe_batch_plan_run(mbe_batch_plan* plan, mbe_result* per_graph, mbe_result* total);
void mbe_batch_plan_free(mbe_batch_plan* plan);

/* ------------------------------------------------------------ serving -- */
/* A request loop with continuous batching. No reference counterpart (the
 * reference solves one graph per process, src/main.cu:26-234).
 *
 * One persistent kernel (one generation) runs while requests arrive. Submit
 * prepares the graph on the caller's thread and queues it. A scheduler
 * thread admits queued requests in arrival order while the sum of admitted
 * footprints stays within hbm_budget -- footprint = packed CSR bytes plus
 * worst-case stack (the largest local problem, bit matrices and frame stack,
 * that any of its roots can build) -- packs the CSR
 * into pinned staging memory, reserves its place in the device data ring and
 * publishes a mailbox record. An idle wave of the running kernel copies the
 * CSR in and appends the request's roots to the shared task space, so a
 * request starts while earlier ones are still running and its subtrees share
 * the work-stealing deque with theirs. A request completes when its own last
 * task does: the kernel raises its host-mapped completion flag. Requests for
 * another variant than cuMBE, or whose local problems exceed the serving
 * workers' arenas, run through mbe_solve in their turn, between two kernel
 * generations. Results are bit-identical to mbe_solve. The server object is
 * internally synchronised; submit and poll may be called from any thread. */
typedef struct mbe_server mbe_server;

typedef struct {
    uint32_t table_cap;      /* requests per kernel generation; 0 = 1 << 18 */
    uint32_t queue_cap;      /* queued (unadmitted) requests; 0 = 1 << 20 */
    uint32_t max_nr;         /* largest enumerated side the kernel serves; 0 = 1 << 17 */
    uint32_t large_edges;    /* requests with at least this many edges take the
                                large claim lane: waves take roots of small
                                requests first, and one wave in four prefers the
                                large lane; 0 = 1 << 18, UINT32_MAX = one lane */
    uint64_t hbm_budget;     /* admission budget in bytes; 0 = 64 GiB */
    uint64_t data_bytes;     /* device CSR ring; 0 = 4 GiB */
    uint64_t staging_bytes;  /* pinned staging ring; 0 = 1 GiB */
    mbe_tuning tuning;       /* launch tuning of the serving kernel (0 fields =
                                defaults: arena_words 1 << 20, lds_words 4096,
                                waves_per_cu 4) */
} mbe_server_config;

typedef struct {
    mbe_status status;       /* MBE_OK, or the request's failure */
    uint32_t route;          /* 0 = serving kernel, 1 = mbe_solve path */
    uint64_t count;
    uint64_t hash;
    uint64_t footprint;      /* admitted bytes */
    double queued_ms;        /* submit -> published to the kernel */
    double latency_ms;       /* submit -> completion seen by the scheduler */
} mbe_request_result;

typedef struct {
    uint64_t submitted, completed, rejected, launches;
    uint64_t queued, in_flight, bytes_in_use;
    uint64_t bicliques;      /* sum of completed counts */
} mbe_server_stats;

mbe_status mbe_server_create(const mbe_server_config* cfg, mbe_server** out);
/* Queue a graph (copied; the caller may free it afterwards). Returns
 * MBE_EOVERBUDGET if its footprint alone exceeds the budget, MBE_EBUSY when
 * the queue is full. */
mbe_status mbe_server_submit(mbe_server* s, const mbe_graph* g, mbe_variant variant,
                             uint64_t* ticket);
/* Wait up to wait_ms (0 = just check, < 0 = forever) for a request; on
 * MBE_OK the result is moved out and the ticket is retired. MBE_ENOTREADY
 * if it has not finished; MBE_EINVAL for an unknown ticket. */
mbe_status mbe_server_poll(mbe_server* s, uint64_t ticket, int wait_ms, mbe_request_result* out);
mbe_status mbe_server_get_stats(mbe_server* s, mbe_server_stats* out);
/* Finish every admitted and queued request, close the kernel's mailbox,
 * wait for the kernel to drain, release the device buffers. */
void mbe_server_destroy(mbe_server* s);

#ifdef __cplusplus
}
#endif
#endif /* MBE_H */
