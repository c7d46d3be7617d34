// Structures shared by the host engine (api.cpp) and the gfx950 kernels
// (mbe_kernel.hip). Everything here is plain data laid out for HBM.
#pragma once
#include <cstddef>
#include <cstdint>

namespace mbe {

// One prepared graph resident in device memory. R = enumerated side, ids are
// root-order ranks; L keeps file ids. Both CSRs have sorted lists.
struct DevGraph {
    const uint32_t* offR;   // nR + 1
    const uint32_t* adjR;   // L ids
    const uint32_t* offL;   // nL + 1
    const uint32_t* adjL;   // R ranks
    const uint64_t* hR;     // hash term per rank (file id of the R side)
    const uint64_t* hL;     // hash term per L id
    uint32_t nR, nL;
    uint32_t n_roots;       // ranks [0, n_roots) have nonzero degree
    uint32_t r_is_u;        // 1 when R holds the file's rows
    uint64_t task_base;     // first global task id of this graph
};

// Counters the kernel accumulates for profiling (per launch, summed).
enum KStat : int {
    kStatNodes = 0,      // enumeration-tree nodes processed (one per chosen x)
    kStatRoots = 1,      // root tasks completed
    kStatDonated = 2,    // subtree tasks pushed to the deque
    kStatStolen = 3,     // subtree tasks taken from the deque
    kStatOverflow = 4,   // root tasks deferred for a larger arena
    kStatBuildWords = 5, // adjacency entries read while building local problems
    kStatDropped = 6,    // slots discarded (snapshot did not fit the thief's arena)
    kStatNodeBudget = 7, // nodes charged against the watchdog budget
    kStatWords = 8,      // bit-matrix operand words the nodes read (algorithmic)
    kStatCount = 12
};

// Kernel time sections, in the slots of the reference's DESECTION clocks
// (src/header.cuh:25-32) as its plot script groups them
// (utils/plot_section.py:7-8): 3 L' construction, 4 maximality checking,
// 5 maximality expansion, 7 candidate selection, 8 subtree fetching, 2 + 9
// work stealing, the rest "others".
enum KSec : int {
    kSecOther = 0,    // loop control, frame bookkeeping
    kSecCompact = 1,  // child re-localisation
    kSecDonate = 2,   // work stealing: donation (snapshot + push)
    kSecLPrime = 3,   // L' construction
    kSecCheck = 4,    // maximality check of non-maximal nodes
    kSecExpand = 5,   // maximality check + expansion of maximal nodes
    kSecReport = 6,   // biclique hash and count
    kSecSelect = 7,   // candidate selection (ES)
    kSecBuild = 8,    // subtree fetching: root claim and localisation
    kSecFetch = 9,    // work stealing: stolen subtree load
    kSecWait = 10,    // idle: deque empty, waiting for work or termination
    kSecCount = 11
};

// Device scratch of the allocation-free solve (mbe_solve_noalloc): one
// static array per device, partitioned per call into control words, the
// deque, task lists and per-worker arenas, marks and rank tables.
constexpr uint64_t kNoAllocBytes = 256ull << 20;
constexpr uint64_t kNoAllocWords = kNoAllocBytes / 8;

// Fixed per-worker LDS header (ES minima, nest stack, section counters) in
// u64 words; a worker's LDS budget is at least this plus one frame.
constexpr uint32_t kLdsHeaderWords = 344;

// Deque slot header for donated subtrees (see mbe_kernel.hip).
constexpr uint32_t kSlotHeaderWords = 8;

// Serving mode: requests arrive while the kernel runs. The host writes a
// request's packed CSR into pinned staging memory and a record into the
// mailbox, then bumps the published count; an idle wave (the loader) copies
// the CSR into the device data buffer, appends the graph to the table and
// its roots to the task space.
struct MailRec {
    uint64_t src;        // u64 offset of the packed CSR in staging
    uint64_t dst;        // u64 offset in the device data buffer
    uint64_t words;      // packed size in u64 words
    uint64_t off[6];     // u64 offsets of offR, adjR, offL, adjL, hR, hL in the block
    uint32_t nR, nL, n_roots, r_is_u;
    uint32_t lane;       // claim lane: 0 small, 1 large
};

struct ServeCtl {                      // device memory, zeroed before launch
    unsigned long long tasks_avail;    // root tasks in the task space
    unsigned long long graphs_avail;   // graphs in the table
    unsigned long long taken;          // mailbox records loaded
    unsigned long long lock;           // loader lock
    unsigned long long next_poll;      // wall clock of the next mailbox poll
    unsigned long long pub_seen;       // mirrors of the host's published count,
    unsigned long long close_seen;     // close flag and abort flag, refreshed by
    unsigned long long abort_seen;     // the polling loader
    unsigned long long lane_avail[2];  // graphs appended to each claim lane
    unsigned long long lane_first[2];  // lane entries below this have every root claimed
};

struct KParams {
    const DevGraph* graphs;
    uint32_t n_graphs;
    uint32_t variant;           // mbe_variant
    uint64_t n_tasks;           // tasks in this launch
    const uint64_t* task_list;  // nullable: explicit global task ids
    uint64_t task_first;        // without a list: task t is task_first + t * task_step
    uint64_t task_step;         // (0 reads as 1)
    unsigned long long* cursor; // next task index (atomic)
    unsigned long long* results;// 2 per graph: count, hash
    unsigned long long* stats;  // kStatCount
    // per-worker resources
    uint64_t* arena;            // workers * arena_words
    uint64_t arena_words;
    uint64_t* mark;             // workers * mark_words (bitmap over R)
    uint64_t mark_words;
    uint32_t* cpos;             // workers * cpos_len (R rank -> local index)
    uint64_t cpos_len;
    uint32_t lds_words;         // dynamic LDS per worker (u64 words)
    uint32_t list_cutoff;       // |P|+|Q| at or below which a node tests rows (forward form)
    uint32_t sparse_ratio;      // list-mode roots: bitmap words > ratio * local adjacency (0 = never)
    uint32_t compact_ratio;     // re-localise a child whose universe shrank this much (0 = never)
    uint32_t twins;             // nonzero: drop twin candidates of the selected vertex from P
    // roots that did not fit the arena
    unsigned long long* overflow_count;
    uint64_t* overflow_list;
    uint64_t overflow_cap;
    // work-stealing deque (donated subtrees)
    uint64_t* slots;            // n_slots * slot_words
    uint64_t slot_words;
    uint32_t n_slots;
    uint32_t donate_depth;
    uint32_t donate_interval;   // nodes between a worker's donation checks
    uint32_t split_nodes;       // a task past this many nodes donates while the deque is short (0 = never)
    unsigned long long* dq;     // [0] head, [1] tail, [2] idle, [3] pending, [4..] seq per slot
    // host-mapped progress: 2 slots per worker (roots done, nodes), plain stores
    unsigned long long* progress;
    unsigned long long* sections;   // kSecCount tick sums, or null
    // watchdog: abort flag (host-mapped, system scope) and node budget
    unsigned long long* abort;
    unsigned long long* node_count;
    uint64_t node_budget;           // 0 = unlimited
    // per-graph completion (nullable): 2 words per graph in device memory
    // (open tasks, error bits) and a host-mapped record of 4 words per graph
    // (count, hash, done flag, error bits)
    unsigned long long* remaining;
    unsigned long long* done;
    // per-worker busy ticks at exit (nullable; needs sections)
    unsigned long long* worker_busy;
    // serving mode (serve != null): graphs is the growing table, results,
    // remaining and done are indexed by table position
    ServeCtl* serve;
    unsigned long long* root_cursor;      // per table position: roots claimed
    uint32_t* lane_list;                  // 2 x mail_cap: table positions per claim lane
    const MailRec* mail;                  // host-mapped ring of mail_cap records
    uint64_t mail_cap;
    const unsigned long long* mail_pub;   // host-mapped: records published
    unsigned long long* mail_taken;       // host-mapped: records loaded
    const uint64_t* staging;              // host-mapped staging buffer
    uint64_t* data;                       // device data buffer
    const unsigned long long* close;      // host-mapped: no more requests
    // biclique materialisation (emit_pos != null): [0] u64 words reserved in
    // emit, [1] u32 words the records take on the host
    unsigned long long* emit_pos;
    uint64_t* emit;
    uint64_t emit_cap;                    // u64 words
};

}  // namespace mbe
