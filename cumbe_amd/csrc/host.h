// Host engine pieces shared by the solve, batch and serving entry points:
// device buffers, launch tuning, graph upload and one enumeration pass.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mbe.h"
#include "engine.h"
#include "graph.h"

// A graph handle of the C ABI: the canonical host CSR.
struct mbe_graph {
    mbe::Graph g;
};

namespace mbe {
hipError_t launch_mbe(const KParams& kp, uint32_t workers, hipStream_t stream);
hipError_t launch_dq_init(unsigned long long* dq, uint32_t n_slots, hipStream_t stream);
hipError_t launch_fill(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t stream);
hipError_t set_lds_limit(uint32_t lds_words);
hipError_t noalloc_scratch_address(void** ptr);
hipError_t launch_put_graph(DevGraph* dst, const DevGraph& g, hipStream_t stream);

namespace host {

// Record an error (thread-local detail for mbe_last_error) and return s.
mbe_status fail(mbe_status s, const std::string& msg);
void set_hip_error(int e);
double now_ms();

#define HIPC(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::mbe::host::set_hip_error((int)e_);                                                           \
            return ::mbe::host::fail(MBE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
        }                                                                                  \
    } while (0)

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

// Device buffers of one launch context: the engine's (reused by every solve)
// or a batch plan's (owned by the plan, whose captured graph points into them).
struct Bufs {
    DevBuf arena, mark, cpos, ctl, graphs, data, ovf, slots, dq, busy;
    unsigned long long* progress = nullptr;  // host-mapped, 2 per worker
    size_t progress_slots = 0;
    hipError_t ensure_progress(size_t workers) {
        if (progress_slots >= 2 * workers) return hipSuccess;
        if (progress) (void)hipHostFree(progress);
        progress = nullptr;
        progress_slots = 0;
        hipError_t e = hipHostMalloc((void**)&progress, 2 * workers * 8,
                                     hipHostMallocMapped | hipHostMallocCoherent);
        if (e == hipSuccess) progress_slots = 2 * workers;
        return e;
    }
    void release() {
        for (DevBuf* b : {&arena, &mark, &cpos, &ctl, &graphs, &data, &ovf, &slots, &dq, &busy}) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->bytes = 0;
        }
        if (progress) (void)hipHostFree(progress);
        progress = nullptr;
        progress_slots = 0;
    }
};

struct Tune {
    uint32_t workers, lds_words, list_cutoff, donate_depth, compact_ratio, donate_interval;
    uint64_t arena_words;
    int side;
    uint32_t flags;
    uint32_t root_begin, root_end, root_stride;
    uint64_t max_nodes;
    double time_limit_ms;
    uint32_t sparse_ratio;  // 0 = never list mode
    uint32_t split_nodes;   // 0 = only when workers are idle
};

// What a launch needs to know about the uploaded graphs.
struct Meta {
    uint32_t ng = 0;
    uint32_t max_nr = 1;
    uint64_t total_tasks = 0;
    uint64_t mark_words() const { return (max_nr + 63) / 64 + 1; }
    uint64_t cpos_len() const { return (uint64_t)max_nr + 1; }
    // control block: cursor, overflow count, stats, section ticks, results
    size_t ctl_words() const { return 2 + mbe::kStatCount + mbe::kSecCount + 2 * (size_t)ng; }
};

constexpr uint32_t kSlots = 1024;  // deque slots
// allocation-free solve: deque slots and default arena (u64 words)
constexpr uint32_t kNoAllocSlots = 64;
constexpr uint64_t kNoAllocArenaWords = 1u << 15;

Tune resolve(const mbe_tuning* t, int cus, size_t lds_per_cu);

// Device bytes of one array in a packed upload: its bytes plus 8 spare,
// rounded up to 256 (one rule for sizing and placement).
inline size_t upload_slot(size_t n) { return (n + 8 + 255) & ~(size_t)255; }
// Device bytes of one prepared graph in a packed upload.
size_t packed_bytes(const Prepared& p);

// Pack prepared graphs into B.data and their descriptors into B.graphs
// (copies on `stream`, synchronised before returning).
mbe_status upload(Bufs& B, hipStream_t stream, const std::vector<const Prepared*>& ps,
                  std::vector<DevGraph>& dg, uint64_t& total_tasks);

Meta meta_of(const std::vector<const Prepared*>& ps, uint64_t total_tasks);

// Size the buffers of one pass and fill its kernel parameters; abort_host is
// the host-mapped watchdog flag the kernel polls.
mbe_status setup_pass(Bufs& B, hipStream_t stream, unsigned long long* abort_host, const Meta& m,
                      const Tune& tu, int variant, const std::vector<uint64_t>& task_list,
                      uint64_t n_tasks, uint64_t arena_words, uint32_t workers, KParams& kp);

// Enqueue one pass with kernels only (capturable): clear control words and
// marks, reset the deque, launch the persistent workers.
hipError_t enqueue_pass(const KParams& kp, uint32_t workers, size_t ctl_clear, hipStream_t s);

// Raise the kernels' dynamic LDS limit to at least lds_words (process-wide).
mbe_status ensure_lds(uint32_t lds_words);

}  // namespace host
}  // namespace mbe
