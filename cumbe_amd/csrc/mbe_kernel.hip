// gfx950 kernels of the MBE engine.
//
// Execution model: a persistent grid of single-wave workgroups; each wave64
// is one worker (the reference's "block", src/mbe_noWS.cuh:101-615). A
// worker claims root tasks from a global cursor (the reference's P_ptr,
// src/mbe_noWS.cuh:123-133), localises the root's subproblem into bit
// matrices and walks its set-enumeration subtree with an explicit stack of
// frames. Idle workers take donated subtrees from a deque in HBM.
//
// Local subproblem of root x (rank i): the universe of L is N(x) (d bits,
// W = ceil(d/64) words); the candidates are the R vertices at distance two
// (nc of them, Wc = ceil(nc/64) words), numbered in rank order, so the
// candidates below position kx are the root's P (ranks < i) and the rest its
// Q (ranks > i) -- the root-level sets of src/mbe_noWS.cuh:135-154.
//   rows[k]  = N(cand k)  n N(x)   as a d-bit set      (forward view)
//   cols[j]  = N(u_j) n cands      as an nc-bit set    (reverse view)
// A frame at level k holds x, the running R hash sum, |R|, |P| and |Q|, then
// L_k (W words), P_k, Q_k and X_k (Wc words each). X_k holds candidates that
// follow P_k in this level's order but were handed to another worker: they
// take part in expansion (absorption, P') but are never selected here. When everything fits, the whole
// problem (hash terms, both bit matrices and all frames) lives in LDS;
// otherwise the frames take the LDS and the rest sits in the worker's HBM
// arena. All accesses go through generic pointers.
//
// Per node (src/mbe_cuMBE.cuh:345-553):
//   candidate selection  x = argmin_{v in P} |row(v) & L|, scanned in
//                        descending candidate order with early stop at the
//                        level's previous minimum (ES);
//   L' = row(x) & L, and the list of its set bits;
//   maximality + expansion, reverse scan (RS): A = AND_{j in L'} cols[j],
//                        O = OR_{j in L'} cols[j]; non-maximal iff A & Q;
//                        absorbed = A & P; P' = (O & P) \ A; Q' = O & Q;
//   forward form (noRS, and small P u Q): the same sets from row tests.
// Q' drops candidates with no neighbour in L' (they can never cover a
// descendant's L); a descent whose universe shrank by compact_ratio is
// re-localised into a nested, compacted problem; root localisation and every
// re-localisation keep only Q candidates whose rows are not contained in
// another kept Q candidate's row (reduce_q), which shrinks the candidate
// universe several-fold on skewed graphs. None of these removes a tree node.
#ifndef MBE_EMULATE
#include <hip/hip_runtime.h>
#endif

#include "engine.h"
#include "hash.h"

namespace mbe {

using u64 = unsigned long long;

namespace {

constexpr uint32_t kInf = 1u << 30;              // src/header.cuh:8
constexpr uint32_t kFrameHdr = 4;                // x, SR, |R|, |P| | |Q| << 32
// Fixed per-worker LDS header: ES minima per level, the nest stack of
// suspended problems, section counters. Problems use LDS after it.
constexpr uint32_t kPreMinLevels = 256;          // deeper levels share the last
constexpr uint32_t kNestMax = 48;                // nested compacted problems
constexpr uint32_t kLdsPreMin = 0;                            // u32[256]
constexpr uint32_t kLdsNest = kLdsPreMin + kPreMinLevels / 2;  // u64[4 * kNestMax]
constexpr uint32_t kLdsSec = kLdsNest + 4 * kNestMax;          // u64[kSecCount]
constexpr uint32_t kLdsHdr = (kLdsSec + kSecCount + 7) & ~7u;
static_assert(kLdsHdr <= kLdsHeaderWords, "LDS header larger than the host reserves");
constexpr uint32_t kArenaHdr = 0;                // arena words before problems
constexpr uint32_t kLdsFramesMin = 24;           // frames an LDS tier must hold
constexpr uint64_t kSplitQueue = 64;             // deque length below which long tasks split
constexpr uint32_t kSparseMaxP = 8;              // list mode only for roots with |P| <= this

enum Variant : int { kCuMBE = 0, kNoRS = 1, kNoES = 2, kNoWS = 3 };

__device__ __forceinline__ uint32_t lane() { return threadIdx.x; }
__device__ __forceinline__ u64 uni64(u64 v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t rdlane32(uint32_t v, uint32_t l) {
    return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ u64 rdlane64(u64 v, uint32_t l) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 lanemask_lt() { return (1ull << lane()) - 1; }
__device__ __forceinline__ u64 shfl_xor64(u64 v, uint32_t m) {
    const uint32_t lo = __shfl_xor((int)(uint32_t)v, (int)m);
    const uint32_t hi = __shfl_xor((int)(uint32_t)(v >> 32), (int)m);
    return ((u64)hi << 32) | lo;
}

// Wave reductions with DPP: row_shr 1,2,4,8 builds an inclusive scan inside
// each 16-lane row, row_bcast:15 and row_bcast:31 carry it across rows, and
// lane 63 holds the total.
#define MBE_DPP(old, v, ctrl, rm) \
    ((uint32_t)__builtin_amdgcn_update_dpp((int)(old), (int)(v), ctrl, rm, 0xf, false))
#define MBE_DPP64(old, v, ctrl, rm)                                                 \
    ((((u64)MBE_DPP((uint32_t)((old) >> 32), (uint32_t)((v) >> 32), ctrl, rm)) << 32) | \
     (u64)MBE_DPP((uint32_t)(old), (uint32_t)(v), ctrl, rm))

__device__ __forceinline__ uint32_t scan_add32(uint32_t v) {  // inclusive
    v += MBE_DPP(0, v, 0x111, 0xf);
    v += MBE_DPP(0, v, 0x112, 0xf);
    v += MBE_DPP(0, v, 0x114, 0xf);
    v += MBE_DPP(0, v, 0x118, 0xf);
    v += MBE_DPP(0, v, 0x142, 0xa);
    v += MBE_DPP(0, v, 0x143, 0xc);
    return v;
}
__device__ __forceinline__ uint32_t red_add32(uint32_t v) { return rdlane32(scan_add32(v), 63); }
__device__ __forceinline__ uint32_t red_min32(uint32_t v) {
    v = min(v, MBE_DPP(0xffffffffu, v, 0x111, 0xf));
    v = min(v, MBE_DPP(0xffffffffu, v, 0x112, 0xf));
    v = min(v, MBE_DPP(0xffffffffu, v, 0x114, 0xf));
    v = min(v, MBE_DPP(0xffffffffu, v, 0x118, 0xf));
    v = min(v, MBE_DPP(0xffffffffu, v, 0x142, 0xa));
    v = min(v, MBE_DPP(0xffffffffu, v, 0x143, 0xc));
    return rdlane32(v, 63);
}
__device__ __forceinline__ u64 red_add64(u64 v) {
    v += MBE_DPP64(0ull, v, 0x111, 0xf);
    v += MBE_DPP64(0ull, v, 0x112, 0xf);
    v += MBE_DPP64(0ull, v, 0x114, 0xf);
    v += MBE_DPP64(0ull, v, 0x118, 0xf);
    v += MBE_DPP64(0ull, v, 0x142, 0xa);
    v += MBE_DPP64(0ull, v, 0x143, 0xc);
    return rdlane64(v, 63);
}

__device__ __forceinline__ void wsync() { __syncthreads(); }

// Watchdog: the host (time limit) or any worker (node budget) raises
// *abort; workers then drop their work and drain the deque. Checked every
// kAbortEvery nodes.
constexpr uint32_t kAbortEvery = 256;
__device__ __forceinline__ bool aborted(const KParams& kp) {
    if (!kp.abort) return false;
    u64 a = 0;
    if (lane() == 0) a = __hip_atomic_load(kp.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return uni64(a) != 0;
}
__device__ __forceinline__ bool charge_nodes(const KParams& kp, u64 n) {
    u64 total = 0;
    if (lane() == 0 && kp.node_budget) {
        total = atomicAdd(kp.node_count, n) + n;
        if (total >= kp.node_budget)
            __hip_atomic_store(kp.abort, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return aborted(kp);
}

// Per-section time accounting (the reference's DESECTION clocks,
// src/header.cuh:25-32), enabled when KParams::sections is set. Ticks are
// wall_clock64() units (100 MHz) summed over workers.
struct Sec {
    u64* t;  // kSecCount counters in the worker's LDS header
    u64 last;
    bool on;
    __device__ __forceinline__ void init(bool en, u64* lds) {
        on = en;
        t = lds + kLdsSec;
        if (lane() < kSecCount) t[lane()] = 0;
        last = en ? (u64)wall_clock64() : 0;
    }
    __device__ __forceinline__ void mark(int i) {
        if (on) {
            u64 now = (u64)wall_clock64();
            if (lane() == 0) t[i] += now - last;
            last = now;
        }
    }
};

// Graph of global task t among the first n graphs (task_base ascending).
__device__ __forceinline__ uint32_t find_graph(const KParams& kp, uint32_t n, uint64_t t) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (kp.graphs[mid].task_base <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// largest j in [0, d) with pre[j] <= f (pre strictly increasing)
__device__ __forceinline__ uint32_t find_j(const uint32_t* pre, uint32_t d, uint32_t f) {
    uint32_t lo = 0, hi = d;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= f) lo = mid; else hi = mid;
    }
    return lo;
}

// One local subproblem: a root's, a stolen snapshot's or a compacted child's.
struct Local {
    uint32_t x, d, W, nc, Wc, kx, fs, nlds, maxdepth;
    uint32_t lds_off, lds_end;   // LDS words [lds_off, lds_end) in use
    bool in_lds;                 // whole problem in LDS
    // list mode (a sparse root problem, layout_sparse): rows holds only the
    // root's P candidates (positions below kx); the candidates of each L
    // bit and the L bits of each candidate are lists instead of bit matrices
    bool sparse;
    uint32_t* lofs;              // d + 1: L bit j's candidates are lcand[lofs[j], lofs[j+1])
    uint32_t* lcand;
    uint32_t* rofs;              // nc + 1: candidate k's L bits are rl[rofs[k], rofs[k+1])
    uint32_t* rl;
    size_t base, end;            // arena words [base, end) in use
    const uint32_t* Nx;
    uint32_t* pre;
    uint32_t* cand;
    u64* hl;                     // hash term per L bit
    u64* hc;                     // hash term per candidate
    u64* rows;
    u64* cols;
    uint32_t* list;              // scratch: candidate compaction
    uint32_t* lbits;             // scratch: set bits of the newest L'
    u64* hstack;                 // frames beyond the LDS ones
    u64* fbase;                  // LDS frames

    __device__ __forceinline__ u64* frame(uint32_t k) const {
        uint32_t i = k - 1;
        return i < nlds ? fbase + (size_t)i * fs : hstack + (size_t)(i - nlds) * fs;
    }
};

__device__ __forceinline__ uint32_t pcount(const u64* F) { return (uint32_t)F[3]; }
__device__ __forceinline__ u64 pq(uint32_t p, uint32_t q) { return (u64)p | ((u64)q << 32); }

// Place a problem of size (d, nc) from LDS word lds_off / arena word off:
// everything in LDS when it fits; else the scratch lists and the frames in
// LDS and the data (hash terms, bit matrices) in the arena; else only as many
// frames as fit in LDS.
__device__ bool layout(const KParams& kp, Local& S, u64* arena, u64* lds, size_t off,
                       uint32_t lds_off) {
    S.sparse = false;
    S.W = (S.d + 63) >> 6;
    S.Wc = (S.nc + 63) >> 6;
    S.maxdepth = min(S.d, S.nc) + 2;
    S.fs = kFrameHdr + S.W + 3 * S.Wc;
    S.lds_off = lds_off;
    S.base = off;
    const size_t data = (size_t)S.d + S.nc + (size_t)S.nc * S.W + (size_t)S.d * S.Wc;
    const size_t scratch = (S.nc + 1) / 2 + 1 + (S.d + 1) / 2 + 1;
    // LDS frames wanted: the real depth is far below min(d, nc) + 2, so
    // only the first kLdsFramesMin frames count when choosing a tier
    const size_t frames = (size_t)min(S.maxdepth, kLdsFramesMin) * S.fs;
    const size_t room = lds_off < kp.lds_words ? kp.lds_words - lds_off : 0;
    S.in_lds = data + scratch + frames <= room;
    const bool scratch_lds = S.in_lds || scratch + frames <= room;
    u64* lp = lds + lds_off;  // LDS cursor
    u64* ap = arena + off;    // arena cursor
    u64* p = S.in_lds ? lp : ap;
    S.hl = p;
    p += S.d;
    S.hc = p;
    p += S.nc;
    S.rows = p;
    p += (size_t)S.nc * S.W;
    S.cols = p;
    p += (size_t)S.d * S.Wc;
    if (S.in_lds) lp = p; else ap = p;
    p = scratch_lds ? lp : ap;
    S.list = (uint32_t*)p;
    p += (S.nc + 1) / 2 + 1;
    S.lbits = (uint32_t*)p;
    p += (S.d + 1) / 2 + 1;
    if (scratch_lds) lp = p; else ap = p;
    const size_t lroom = room - (size_t)(lp - (lds + lds_off));
    S.nlds = (uint32_t)min((size_t)S.maxdepth, lroom / S.fs);
    S.fbase = lp;
    lp += (size_t)S.nlds * S.fs;
    S.hstack = ap;
    ap += (size_t)(S.maxdepth - S.nlds) * S.fs;
    S.lds_end = (uint32_t)(lp - lds);
    S.end = (size_t)(ap - arena);
    return S.end <= kp.arena_words;
}

// Place a list-mode root problem (d, nc, kx set; ntl = entries of each of
// the two lists, or ~0 to read it back from lofs[d]) from arena word off:
// hash terms, the P candidates' rows, both list offsets, then the scratch
// lists and as many frames as fit in LDS, the other frames, then the lists.
// P and X only ever hold the root's P candidates, so a search is at most
// min(d, kx) + 2 levels deep.
__device__ bool layout_sparse(const KParams& kp, Local& S, u64* arena, u64* lds, size_t off,
                              uint32_t lds_off, uint32_t ntl) {
    S.sparse = true;
    S.in_lds = false;
    S.W = (S.d + 63) >> 6;
    S.Wc = (S.nc + 63) >> 6;
    S.maxdepth = min(S.d, S.kx) + 2;
    S.fs = kFrameHdr + S.W + 3 * S.Wc;
    S.lds_off = lds_off;
    S.base = off;
    S.cols = nullptr;
    u64* ap = arena + off;
    S.hl = ap;
    ap += S.d;
    S.hc = ap;
    ap += S.nc;
    S.rows = ap;
    ap += (size_t)S.kx * S.W;
    S.lofs = (uint32_t*)ap;
    ap += (S.d + 2) / 2;
    S.rofs = (uint32_t*)ap;
    ap += (S.nc + 2) / 2;
    if (ntl == 0xFFFFFFFFu) ntl = S.lofs[S.d];
    const size_t room = lds_off < kp.lds_words ? kp.lds_words - lds_off : 0;
    const size_t scratch = (S.nc + 1) / 2 + 1 + (S.d + 1) / 2 + 1;
    u64* lp = lds + lds_off;
    const bool scratch_lds = scratch <= room;
    u64* p = scratch_lds ? lp : ap;
    S.list = (uint32_t*)p;
    p += (S.nc + 1) / 2 + 1;
    S.lbits = (uint32_t*)p;
    p += (S.d + 1) / 2 + 1;
    if (scratch_lds) lp = p; else ap = p;
    const size_t lroom = room - (size_t)(lp - (lds + lds_off));
    S.nlds = (uint32_t)min((size_t)S.maxdepth, lroom / S.fs);
    S.fbase = lp;
    lp += (size_t)S.nlds * S.fs;
    S.hstack = ap;
    ap += (size_t)(S.maxdepth - S.nlds) * S.fs;
    S.lcand = (uint32_t*)ap;
    ap += (ntl + 1) / 2;
    S.rl = (uint32_t*)ap;
    ap += (ntl + 1) / 2;
    S.lds_end = (uint32_t)(lp - lds);
    S.end = (size_t)(ap - arena);
    return S.end <= kp.arena_words;
}

// cols = transpose(rows) with 64 ballots per 64x64 bit block.
__device__ void build_cols(const Local& S) {
    for (uint32_t b = 0; b < S.Wc; b++) {
        const uint32_t k = b * 64 + lane();
        const bool valid = k < S.nc;
        for (uint32_t w = 0; w < S.W; w++) {
            u64 rw = valid ? S.rows[(size_t)k * S.W + w] : 0;
            u64 mine = 0;
            if (__ballot(rw != 0)) {
                for (uint32_t t = 0; t < 64; t++) {
                    u64 bal = __ballot((rw >> t) & 1);
                    if (lane() == t) mine = bal;
                }
            }
            const uint32_t j = w * 64 + lane();
            if (j < S.d) S.cols[(size_t)j * S.Wc + b] = mine;
        }
    }
    wsync();
}

// Sum of hash terms over the first nl entries of the L' bit list.
__device__ __forceinline__ u64 sum_l(const Local& S, uint32_t nl) {
    u64 s = 0;
    for (uint32_t c0 = 0; c0 < nl; c0 += 64) {
        uint32_t l = c0 + lane();
        if (l < nl) s += S.hl[S.lbits[l]];
    }
    return red_add64(s);
}

__device__ __forceinline__ void record(const DevGraph& G, u64 sl, u64 sr, u64 nl, u64 nr, u64& cnt,
                                       u64& hs) {
    u64 h = G.r_is_u ? biclique_hash(sr, sl, nr, nl) : biclique_hash(sl, sr, nl, nr);
    cnt++;
    hs += h;
}

// Materialisation: one record per reported biclique, [nl | nr << 32, R hash
// sum, the nl L' hash terms]. The terms identify the L ids (mix64 is a
// bijection), R' = N(L') follows from the graph and is checked against the
// sum and size. lbits: the L' bit list, or null when L' is every L bit.
// Records past cap are counted, not written. Kept out of line: it runs only
// in materialising launches and must not disturb the node loop's registers.
__device__ __noinline__ void emit_biclique(unsigned long long* pos, uint64_t* out, uint64_t cap,
                                           const u64* hl, const uint32_t* lbits, uint32_t nl,
                                           u64 nr, u64 sr) {
    u64 at = 0;
    if (lane() == 0) {
        at = atomicAdd(&pos[0], (u64)nl + 2);
        atomicAdd(&pos[1], (u64)nl + nr + 2);
    }
    at = uni64(at);
    if (at + nl + 2 > cap) return;
    uint64_t* o = out + at;
    if (lane() == 0) {
        o[0] = (u64)nl | (nr << 32);
        o[1] = sr;
    }
    for (uint32_t l = lane(); l < nl; l += 64) o[2 + l] = hl[lbits ? lbits[l] : l];
}

// Candidate selection at a level >= 1 (src/mbe_cuMBE.cuh:345-390): scan P in
// descending candidate order (words high to low, one lane per set bit) and
// keep the first strict minimum of |row(v) & L|; with ES stop as soon as the
// running minimum equals the level's previous minimum
// (src/mbe_cuMBE.cuh:352,365-366; noES scans all, src/mbe_noES.cuh:352).
// Removes the chosen candidate from P and returns it.
template <int V>
__device__ uint32_t select_x(const Local& S, u64* F, uint32_t* premin_lvl) {
    u64* L = F + kFrameHdr;
    u64* P = L + S.W;
    const uint32_t pm = *premin_lvl;
    uint32_t best = kInf, arg = 0;
    bool nostop = (V == kNoES);
    bool done = false;
    // P words 64 at a time, highest first (lane i holds word top - i); the
    // nonzero ones are taken four per step so that their row gathers are in
    // flight together, and the ES scan walks them in descending order
    for (int top = (int)S.Wc - 1; top >= 0 && !done; top -= 64) {
        const int myw = top - (int)lane();
        const u64 mine = myw >= 0 ? P[myw] : 0;
        u64 nz = __ballot(mine != 0);
        while (nz && !done) {
            u64 pw[4];
            int wi[4];
            uint32_t c[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                if (nz) {
                    const uint32_t li = __builtin_ctzll(nz);
                    nz &= nz - 1;
                    pw[s] = rdlane64(mine, li);
                    wi[s] = top - (int)li;
                } else {
                    pw[s] = 0;
                    wi[s] = 0;
                }
                c[s] = kInf;
            }
#pragma unroll
            for (int s = 0; s < 4; s++) {
                if ((pw[s] >> lane()) & 1) {
                    const u64* r = S.rows + (size_t)(wi[s] * 64 + lane()) * S.W;
                    uint32_t cc = 0;
                    for (uint32_t i = 0; i < S.W; i++) cc += __popcll(r[i] & L[i]);
                    c[s] = cc;
                }
            }
            // ES: the first candidate in scan order (words in order, high
            // lanes first) whose count is at most the level's previous
            // minimum stops the scan if it equals it; one below it ends early
            // stopping. Counts of absent candidates are kInf.
            if (!nostop) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const u64 m = __ballot(c[s] <= pm);
                    if (!m) continue;
                    const uint32_t sl = 63 - __builtin_clzll(m);
                    if (rdlane32(c[s], sl) == pm) {  // earlier words hold no count <= pm
                        done = true;
                        if (best > pm) {
                            best = pm;
                            arg = wi[s] * 64 + sl;
                        }
                    } else {
                        nostop = true;
                    }
                    break;
                }
            }
            if (!done) {
                // one reduction for the group; ties go to the first word in
                // scan order and its highest lane
                uint32_t mn = c[0];
#pragma unroll
                for (int s = 1; s < 4; s++) mn = min(mn, c[s]);
                const uint32_t cm = red_min32(mn);
                if (cm < best) {
                    best = cm;
#pragma unroll
                    for (int s = 3; s >= 0; s--) {
                        const u64 m = __ballot(c[s] == cm);
                        if (m) arg = wi[s] * 64 + (63 - __builtin_clzll(m));
                    }
                }
            }
        }
    }
    if (lane() == 0) {
        *premin_lvl = best;
        P[arg >> 6] &= ~(1ull << (arg & 63));
        F[3] -= 1;  // |P|
    }
    wsync();
    return arg;
}

struct NodeOut {
    bool maximal;
    bool descend;
    uint32_t ab;    // absorbed count
    u64 abh;        // absorbed hash sum
    uint32_t pn;    // |P'|
    uint32_t qn;    // |Q'|
    uint32_t un;    // |P' u Q'|
    u64 tw;         // Wc <= 64 reverse scan: lane w < Wc holds word w of P & A
    bool has_tw;
};

// L' = row(x) & L into the child frame and its set-bit list; returns |L'|.
__device__ uint32_t make_lprime(const Local& S, const u64* L, uint32_t x, u64* Lc) {
    const u64* r = S.rows + (size_t)x * S.W;
    uint32_t nl = 0;
    for (uint32_t w0 = 0; w0 < S.W; w0 += 64) {
        const uint32_t w = w0 + lane();
        u64 lw = 0;
        if (w < S.W) {
            lw = r[w] & L[w];
            Lc[w] = lw;
        }
        const uint32_t nw = min(64u, S.W - w0);
        for (uint32_t i = 0; i < nw; i++) {
            const u64 word = rdlane64(lw, i);
            if ((word >> lane()) & 1)
                S.lbits[nl + __popcll(word & lanemask_lt())] = (w0 + i) * 64 + lane();
            nl += __popcll(word);
        }
    }
    wsync();
    return nl;
}

// Reverse-scan maximality check and expansion (src/mbe_noWS.cuh:215-289):
// AND/OR of the L' columns. Wc <= 64: lanes split into groups of L' bits
// times column words, then a butterfly over the groups; larger Wc: one lane
// per column word, walking the L' bits.
__device__ NodeOut process_rs(const Local& S, u64* F, u64* Fc, uint32_t nl) {
    const u64* P = F + kFrameHdr + S.W;
    const u64* Q = P + S.Wc;
    const u64* X = Q + S.Wc;
    u64* Pc = Fc + kFrameHdr + S.W;
    u64* Qc = Pc + S.Wc;
    u64* Xc = Qc + S.Wc;
    NodeOut o{true, false, 0, 0, 0, 0, 0, 0, false};
    u64 abh = 0;
    if (S.Wc <= 64) {
        // lane = group * Wp + word (Wp = Wc rounded up to a power of two):
        // each of the 64 / Wp groups AND/ORs every (64 / Wp)-th L' column
        // with independent loads, then a butterfly over the groups leaves
        // the totals in every lane of the word
        const uint32_t Wc = S.Wc;
        const uint32_t lg = Wc > 1 ? 32 - __builtin_clz(Wc - 1) : 0;
        const uint32_t Wp = 1u << lg, G = 64u >> lg;
        const uint32_t w = lane() & (Wp - 1), gi = lane() >> lg;
        const bool valid = w < Wc;
        u64 a = ~0ull, ob = 0;
        if (valid) {
            uint32_t l = gi;
            for (; l + 3 * G < nl; l += 4 * G) {
                const uint32_t j0 = S.lbits[l], j1 = S.lbits[l + G];
                const uint32_t j2 = S.lbits[l + 2 * G], j3 = S.lbits[l + 3 * G];
                const u64 c0 = S.cols[(size_t)j0 * Wc + w], c1 = S.cols[(size_t)j1 * Wc + w];
                const u64 c2 = S.cols[(size_t)j2 * Wc + w], c3 = S.cols[(size_t)j3 * Wc + w];
                a &= c0 & c1 & c2 & c3;
                ob |= c0 | c1 | c2 | c3;
            }
            for (; l < nl; l += G) {
                const u64 c = S.cols[(size_t)S.lbits[l] * Wc + w];
                a &= c;
                ob |= c;
            }
        }
        for (uint32_t m = Wp; m < 64; m <<= 1) a &= shfl_xor64(a, m);
        const bool mine = valid && gi == 0;
        const u64 q = mine ? Q[w] : 0;
        o.tw = mine ? a & P[w] : 0;  // candidates of P containing L'
        o.has_tw = true;
        const bool nonmax = mine && (a & q) != 0;
        o.maximal = __ballot(nonmax) == 0;
        if (o.maximal) {  // O is combined only where it is used
            for (uint32_t m = Wp; m < 64; m <<= 1) ob |= shfl_xor64(ob, m);
            const u64 p = mine ? (P[w] | X[w]) : 0;
            const u64 absd = a & p, pn = ob & p & ~a, qn = ob & q;
            if (mine) {
                Pc[w] = pn;
                Qc[w] = qn;
                Xc[w] = 0;
            }
            // four counts of at most 4096 each, packed into one reduction
            const u64 packed = (u64)__popcll(absd) | ((u64)__popcll(pn) << 16) |
                               ((u64)__popcll(qn) << 32) | ((u64)__popcll(pn | qn) << 48);
            const u64 tot = red_add64(packed);
            o.ab = (uint32_t)(tot & 0xFFFF);
            o.pn = (uint32_t)((tot >> 16) & 0xFFFF);
            o.qn = (uint32_t)((tot >> 32) & 0xFFFF);
            o.un = (uint32_t)(tot >> 48);
            u64 am = __ballot(mine && absd != 0);
            while (am) {  // absorbed candidates' hash terms, one word at a time
                const uint32_t ww = __builtin_ctzll(am);
                am &= am - 1;
                const u64 aw = rdlane64(absd, ww);
                if ((aw >> lane()) & 1) abh += S.hc[ww * 64 + lane()];
            }
        }
    } else {
        // First pass: is some Q candidate adjacent to all of L'? L' is
        // usually tiny at non-maximal nodes and covering candidates are
        // plentiful, so this stops at the first 64-word chunk holding one.
        bool nonmax = false;
        for (uint32_t w0 = 0; w0 < S.Wc && !nonmax; w0 += 64) {
            const uint32_t w = w0 + lane();
            u64 a = w < S.Wc ? Q[w] : 0;
            uint32_t l = 0;
            for (; l + 4 <= nl && a; l += 4) {
                const uint32_t j0 = S.lbits[l], j1 = S.lbits[l + 1];
                const uint32_t j2 = S.lbits[l + 2], j3 = S.lbits[l + 3];
                a &= S.cols[(size_t)j0 * S.Wc + w] & S.cols[(size_t)j1 * S.Wc + w] &
                     S.cols[(size_t)j2 * S.Wc + w] & S.cols[(size_t)j3 * S.Wc + w];
            }
            for (; l < nl && a; l++) a &= S.cols[(size_t)S.lbits[l] * S.Wc + w];
            nonmax = __ballot(a != 0) != 0;
        }
        o.maximal = !nonmax;
        // second pass, maximal nodes only: A and O over every word
        uint32_t ab = 0, pnc = 0, qnc = 0, unc = 0;
        for (uint32_t w0 = 0; w0 < S.Wc && o.maximal; w0 += 64) {
            const uint32_t w = w0 + lane();
            const bool valid = w < S.Wc;
            u64 a = ~0ull, ob = 0;
            uint32_t l = 0;
            for (; l + 4 <= nl; l += 4) {
                const uint32_t j0 = S.lbits[l], j1 = S.lbits[l + 1];
                const uint32_t j2 = S.lbits[l + 2], j3 = S.lbits[l + 3];
                if (valid) {
                    const u64 c0 = S.cols[(size_t)j0 * S.Wc + w], c1 = S.cols[(size_t)j1 * S.Wc + w];
                    const u64 c2 = S.cols[(size_t)j2 * S.Wc + w], c3 = S.cols[(size_t)j3 * S.Wc + w];
                    a &= c0 & c1 & c2 & c3;
                    ob |= c0 | c1 | c2 | c3;
                }
            }
            for (; l < nl; l++) {
                const uint32_t j = S.lbits[l];
                if (valid) {
                    const u64 c = S.cols[(size_t)j * S.Wc + w];
                    a &= c;
                    ob |= c;
                }
            }
            const u64 q = valid ? Q[w] : 0, p = valid ? (P[w] | X[w]) : 0;
            const u64 absd = a & p, pn = ob & p & ~a, qn = ob & q;
            if (valid) {
                Pc[w] = pn;
                Qc[w] = qn;
                Xc[w] = 0;
            }
            ab += __popcll(absd);
            pnc += __popcll(pn);
            qnc += __popcll(qn);
            unc += __popcll(pn | qn);
            u64 t = absd;
            while (t) {
                const uint32_t b = __builtin_ctzll(t);
                t &= t - 1;
                abh += S.hc[w * 64 + b];
            }
        }
        if (o.maximal) {
            o.ab = red_add32(ab);
            o.pn = red_add32(pnc);
            o.qn = red_add32(qnc);
            o.un = red_add32(unc);
        }
    }
    o.abh = red_add64(abh);
    o.descend = o.maximal && o.pn != 0;
    wsync();
    return o;
}

// Forward form (noRS, src/mbe_noRS.cuh:216-287, and nodes whose P u Q is
// small): each candidate's row is tested against L', one lane per set bit.
__device__ NodeOut process_fwd(const Local& S, u64* F, u64* Fc, uint32_t nl) {
    const u64* P = F + kFrameHdr + S.W;
    const u64* Q = P + S.Wc;
    const u64* X = Q + S.Wc;
    const u64* Lc = Fc + kFrameHdr;
    u64* Pc = Fc + kFrameHdr + S.W;
    u64* Qc = Pc + S.Wc;
    u64* Xc = Qc + S.Wc;
    NodeOut o{true, false, 0, 0, 0, 0, 0, 0, false};
    for (uint32_t w = 0; w < S.Wc; w++) {
        const u64 qw = uni64(Q[w]);
        u64 qn = 0;
        if (qw) {
            const bool valid = (qw >> lane()) & 1;
            bool cov = false, touch = false;
            if (valid) {
                const u64* r = S.rows + (size_t)(w * 64 + lane()) * S.W;
                cov = true;
                for (uint32_t i = 0; i < S.W; i++) {
                    const u64 l = Lc[i], x = r[i] & l;
                    cov &= x == l;
                    touch |= x != 0;
                }
            }
            if (__ballot(cov)) {
                o.maximal = false;
                break;
            }
            qn = __ballot(touch);
        }
        if (lane() == 0) Qc[w] = qn;
        o.qn += __popcll(qn);
    }
    if (!o.maximal) {
        wsync();
        return o;
    }
    wsync();
    u64 abh = 0;
    for (uint32_t w = 0; w < S.Wc; w++) {
        const u64 pw = uni64(P[w] | X[w]);
        u64 pn = 0, absd = 0;
        if (pw) {
            const bool valid = (pw >> lane()) & 1;
            uint32_t c = 0;
            if (valid) {
                const u64* r = S.rows + (size_t)(w * 64 + lane()) * S.W;
                for (uint32_t i = 0; i < S.W; i++) c += __popcll(r[i] & Lc[i]);
            }
            absd = __ballot(valid && c == nl);
            pn = __ballot(valid && c != 0 && c != nl);
            if ((absd >> lane()) & 1) abh += S.hc[w * 64 + lane()];
        }
        if (lane() == 0) {
            Pc[w] = pn;
            Xc[w] = 0;
        }
        o.ab += __popcll(absd);
        o.pn += __popcll(pn);
        o.un += __popcll(pn | uni64(Qc[w]));
    }
    o.abh = red_add64(abh);
    o.descend = o.pn != 0;
    wsync();
    return o;
}

// List-mode maximality check and expansion (a sparse root problem): the
// reference's list scans (src/mbe_cuMBE.cuh:216-316) narrowed by two facts.
// A Q candidate containing L' is adjacent to the L' bit j* with the fewest
// candidates, so only j*'s candidates are tested, each by a wave-wide scan
// of its L-bit list against L'. Absorption and P' concern the root's P
// candidates only, whose rows are bitmaps. Q' is Q (the check of a
// descendant only looks at candidates adjacent to its own L'). With prune
// set, twins of the selected vertex (rows within L equal to L', see
// prune_twins) leave P in the same pass. `scanned` adds the list entries
// and row words read.
__device__ __noinline__ NodeOut process_sparse(const Local& S, u64* F, u64* Fc, uint32_t nl,
                                               bool prune, u64& scanned) {
    const u64* L = F + kFrameHdr;
    u64* P = F + kFrameHdr + S.W;
    const u64* Q = P + S.Wc;
    const u64* X = Q + S.Wc;
    const u64* Lc = Fc + kFrameHdr;
    u64* Pc = Fc + kFrameHdr + S.W;
    u64* Qc = Pc + S.Wc;
    u64* Xc = Qc + S.Wc;
    NodeOut o{true, false, 0, 0, 0, 0, 0, 0, false};
    uint32_t bl = 0xFFFFFFFFu, bj = 0;
    for (uint32_t l = lane(); l < nl; l += 64) {
        const uint32_t j = S.lbits[l];
        const uint32_t len = S.lofs[j + 1] - S.lofs[j];
        if (len < bl) {
            bl = len;
            bj = j;
        }
    }
    const uint32_t mn = red_min32(bl);
    const uint32_t js = rdlane32(bj, __builtin_ctzll(__ballot(bl == mn)));
    const uint32_t c0 = S.lofs[js], c1 = S.lofs[js + 1];
    u64 sc = c1 - c0;
    bool nonmax = false;
    for (uint32_t t = c0; t < c1 && !nonmax; t++) {
        const uint32_t q = S.lcand[t];
        if (!((Q[q >> 6] >> (q & 63)) & 1)) continue;
        const uint32_t r0 = S.rofs[q], r1 = S.rofs[q + 1];
        if (r1 - r0 < nl) continue;
        sc += r1 - r0;
        uint32_t hits = 0;
        for (uint32_t e0 = r0; e0 < r1; e0 += 64) {
            const uint32_t e = e0 + lane();
            bool hit = false;
            if (e < r1) {
                const uint32_t j = S.rl[e];
                hit = (Lc[j >> 6] >> (j & 63)) & 1;
            }
            hits += __popcll(__ballot(hit));
        }
        nonmax = hits == nl;
    }
    o.maximal = !nonmax;
    if (nonmax && !prune) {
        scanned += sc;
        wsync();
        return o;
    }
    // the root's P candidates: absorption and P' (maximal nodes), twins
    const uint32_t Wp = (S.kx + 63) >> 6;
    uint32_t ab = 0, pnc = 0, removed = 0;
    u64 abh = 0;
    for (uint32_t w = 0; w < Wp; w++) {
        const u64 pw = uni64(P[w]), all = pw | uni64(X[w]);
        u64 pn = 0, absd = 0, twins = 0;
        if (all) {
            const bool valid = (all >> lane()) & 1;
            uint32_t c = 0, cl = 0;
            if (valid) {
                const u64* r = S.rows + (size_t)(w * 64 + lane()) * S.W;
                for (uint32_t i = 0; i < S.W; i++) {
                    const u64 rw = r[i];
                    c += __popcll(rw & Lc[i]);
                    cl += __popcll(rw & L[i]);
                }
            }
            absd = __ballot(valid && c == nl);
            pn = __ballot(valid && c != 0 && c != nl);
            twins = prune ? __ballot(valid && c == nl && cl == nl && ((pw >> lane()) & 1)) : 0;
            if ((absd >> lane()) & 1) abh += S.hc[w * 64 + lane()];
            sc += (u64)__popcll(all) * S.W;
        }
        if (lane() == 0) {
            if (!nonmax) Pc[w] = pn;
            if (twins) P[w] = pw & ~twins;
        }
        ab += __popcll(absd);
        pnc += __popcll(pn);
        removed += __popcll(twins);
    }
    if (nonmax) {
        if (lane() == 0 && removed) F[3] -= removed;
        scanned += sc;
        wsync();
        return o;
    }
    uint32_t qn = 0;
    for (uint32_t w = lane(); w < S.Wc; w += 64) {
        const u64 q = Q[w];
        Qc[w] = q;
        Xc[w] = 0;
        if (w >= Wp) Pc[w] = 0;
        qn += __popcll(q);
    }
    if (lane() == 0 && removed) F[3] -= removed;  // |P| is the low half
    o.ab = ab;
    o.abh = red_add64(abh);
    o.pn = pnc;
    o.qn = red_add32(qn);
    o.un = pnc + o.qn;
    o.descend = pnc != 0;
    scanned += sc + Wp * S.W;
    wsync();
    return o;
}

// Twin pruning (the candidate removal of iMBEA, Zhang et al., BMC
// Bioinformatics 2014, which the reference leaves out). A candidate v left in
// P with N(v) n L == L' = N(x) n L would, selected later at this level, give
// the same L', which x (by then in Q) covers: a non-maximal leaf. A node
// below a later sibling that holds v in R has its L inside L', which x covers
// too, so v is no witness that x is not. Such v leave P and enter neither Q
// nor X: the tree loses exactly these leaves and reports the same bicliques.
// T = P & AND of the L' columns holds the candidates containing L'; its
// twins have no neighbour in L \ L'. The scratch list (free until the next
// compaction) holds T.
// Kept out of line, with plain arguments, so that the node loop's inlining
// and registers stay as they are; returns |T| << 32 | removed.
// tw: when has_tw, lane w < Wc already holds word w of T (the reverse scan
// computed it), which saves the column loads.
__device__ __noinline__ u64 prune_twins(u64* F, const u64* Lc, const u64* cols, const u64* rows,
                                        const uint32_t* lbits, uint32_t* list, uint32_t W,
                                        uint32_t Wc, uint32_t nl, u64 tw, bool has_tw) {
    const u64* L = F + kFrameHdr;
    u64* P = F + kFrameHdr + W;
    uint32_t nt = 0;
    for (uint32_t w0 = 0; w0 < Wc; w0 += 64) {
        const uint32_t w = w0 + lane();
        u64 a = has_tw ? tw : (w < Wc ? P[w] : 0);
        if (!__ballot(a != 0)) continue;
        uint32_t l = has_tw ? nl : 0;
        for (; l + 4 <= nl && a; l += 4) {
            const uint32_t j0 = lbits[l], j1 = lbits[l + 1];
            const uint32_t j2 = lbits[l + 2], j3 = lbits[l + 3];
            a &= cols[(size_t)j0 * Wc + w] & cols[(size_t)j1 * Wc + w] &
                 cols[(size_t)j2 * Wc + w] & cols[(size_t)j3 * Wc + w];
        }
        for (; l < nl && a; l++) a &= cols[(size_t)lbits[l] * Wc + w];
        const uint32_t c = __popcll(a);
        const uint32_t inc = scan_add32(c);
        uint32_t k = nt + inc - c;
        while (a) {
            const uint32_t b = __builtin_ctzll(a);
            a &= a - 1;
            list[k++] = w * 64 + b;
        }
        nt += rdlane32(inc, 63);
    }
    if (nt == 0) return 0;
    wsync();
    uint32_t removed = 0;
    if (W <= 64) {
        // lane = group * Wp + word: 64 / Wp candidates per pass
        const uint32_t lg = W > 1 ? 32 - __builtin_clz(W - 1) : 0;
        const uint32_t Wp = 1u << lg;
        const uint32_t G = 64u >> lg;
        const uint32_t w = lane() & (Wp - 1), gi = lane() >> lg;
        const u64 m = w < W ? (L[w] & ~Lc[w]) : 0;
        const u64 gmask = (Wp == 64 ? ~0ull : ((1ull << Wp) - 1)) << (gi * Wp);
        for (uint32_t b0 = 0; b0 < nt; b0 += G) {
            const uint32_t i = b0 + gi;
            const uint32_t c = i < nt ? list[i] : 0;
            const bool hit = i < nt && w < W && (rows[(size_t)c * W + w] & m) != 0;
            const u64 bad = __ballot(hit);
            const bool twin = i < nt && w == 0 && (bad & gmask) == 0;
            if (twin) atomicAnd(&P[c >> 6], ~(1ull << (c & 63)));
            removed += __popcll(__ballot(twin));
        }
    } else {
        for (uint32_t i = 0; i < nt; i++) {
            const uint32_t c = list[i];
            bool hit = false;
            for (uint32_t w = lane(); w < W; w += 64)
                hit |= (rows[(size_t)c * W + w] & L[w] & ~Lc[w]) != 0;
            if (!__ballot(hit)) {
                if (lane() == 0) P[c >> 6] &= ~(1ull << (c & 63));
                removed++;
            }
        }
    }
    if (lane() == 0 && removed) F[3] -= removed;  // |P| is the low half
    wsync();
    return ((u64)nt << 32) | removed;
}

// ------------------------------------------------------------------ deque
// Donated subtrees travel through a bounded multi-producer multi-consumer
// ring of fixed-size slots in HBM (per-slot sequence numbers, Vyukov's
// bounded queue). dq[0] head, dq[1] tail, dq[2] idle workers, dq[3] pending
// tasks (queued or running), dq[4 + i] sequence number of slot i.
// Slot layout (u64 words): [0] graph, [1] d, [2] nc, [3] R hash sum, [4] |R|,
// then hl[d], hc[nc], rows[nc * W], P[Wc], Q[Wc]: a LEVEL(L, R, P, Q) call
// over a compacted universe (L = all d bits).

__device__ __forceinline__ u64 ld_rlx(unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(unsigned long long* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(unsigned long long* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Acquire-only fence: waits for this wave's memory operations and
// invalidates its stale cache lines, without writing the L2 back (a release
// or a full fence would write back every dirty line of the XCD's L2).
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Claim a slot to write (lane 0 decides); returns its position or -1 if full.
__device__ int64_t dq_claim_push(const KParams& kp) {
    int64_t r = -1;
    if (lane() == 0) {
        unsigned long long* tail = kp.dq + 1;
        u64 pos = ld_rlx(tail);
        for (int it = 0; it < 64; it++) {
            u64 seq = ld_rlx(kp.dq + 4 + (pos % kp.n_slots));
            int64_t dif = (int64_t)(seq - pos);
            if (dif == 0) {
                u64 old = atomicCAS(tail, pos, pos + 1);
                if (old == pos) { r = (int64_t)pos; break; }
                pos = old;
            } else if (dif < 0) {
                break;  // full
            } else {
                pos = ld_rlx(tail);
            }
        }
    }
    r = (int64_t)uni64((u64)r);
    if (r >= 0) fence_acq();  // the slot's last reader is done with it
    return r;
}

__device__ void dq_publish(const KParams& kp, u64 pos) {
    __threadfence();
    wsync();
    if (lane() == 0) st_rel(kp.dq + 4 + (pos % kp.n_slots), pos + 1);
}

// Claim a filled slot to read; returns its position or -1 if empty.
__device__ int64_t dq_claim_pop(const KParams& kp) {
    int64_t r = -1;
    if (lane() == 0) {
        unsigned long long* head = kp.dq;
        u64 pos = ld_rlx(head);
        for (int it = 0; it < 64; it++) {
            u64 seq = ld_rlx(kp.dq + 4 + (pos % kp.n_slots));
            int64_t dif = (int64_t)(seq - (pos + 1));
            if (dif == 0) {
                u64 old = atomicCAS(head, pos, pos + 1);
                if (old == pos) { r = (int64_t)pos; break; }
                pos = old;
            } else if (dif < 0) {
                break;  // empty
            } else {
                pos = ld_rlx(head);
            }
        }
    }
    r = (int64_t)uni64((u64)r);
    // an empty probe touches no cache; a claimed slot is read after an
    // acquire (its writer published it with a release)
    if (r >= 0) fence_acq();
    return r;
}

// The reader hands the slot back once its loads are complete (their values
// are already in the worker's memory); it wrote nothing others need, so no
// L2 write-back.
__device__ void dq_release(const KParams& kp, u64 pos) {
    __builtin_amdgcn_s_waitcnt(0);
    wsync();
    if (lane() == 0) st_rlx(kp.dq + 4 + (pos % kp.n_slots), pos + kp.n_slots);
}

// ------------------------------------------------------------ Q reduction
// Q (candidates already handled) only ever answers one question: does some
// Q candidate's row contain L'? Every L' of this problem and of its
// descendants lies in the problem's universe, and a descendant's Q rows are
// restrictions of these rows, so a Q candidate whose row is contained in
// another kept Q candidate's row never decides anything. Dropping it leaves
// every maximality decision, the tree and the bicliques unchanged. At the
// configuration-1 roots of rank ~400, for example, 8,782 Q candidates hold
// 1,730 distinct rows, and 339 of them are not contained in another.
//
// Rows are taken by popcount, highest first, so that a row can only be
// contained in rows taken before it; each row is tested against the kept
// rows through the kept rows' columns (kcols[j] = kept rows holding bit j,
// kKeepCap of them), a batch of 64 at a time. Rows of one batch are not
// tested against each other, so a few contained rows may stay: the result
// is a valid reduction, not always the smallest. Once kKeepCap rows are
// kept, later survivors are kept without being indexed.
constexpr uint32_t kKeepCap = 1024;
constexpr uint32_t kKW = kKeepCap / 64;  // kcols words per universe bit
constexpr uint32_t kReduceMinQ = 64;     // smaller Q sets are kept whole

// After a wave's own global atomics: make them visible to its later loads.
__device__ __forceinline__ void sync_mem() {
    fence_acq();
    wsync();
}

// Words of scratch reduce_q needs for nq rows over a d-bit universe.
__host__ __device__ constexpr size_t reduce_scratch_words(uint32_t d, uint32_t nq) {
    return (size_t)(d + 2) / 2 + (nq + 1) / 2 + (size_t)d * kKW + 2;
}

// rows: candidate rows of W words over a d-bit universe, by position; the
// Q candidates are positions qs[0..nq) (or q0 + i when qs is null). Sets
// the bit of every kept Q position in keep (cleared by the caller) and
// returns how many were kept. scratch: reduce_scratch_words(d, nq) words.
__device__ uint32_t reduce_q(const u64* rows, uint32_t W, uint32_t d, const uint32_t* qs,
                             uint32_t q0, uint32_t nq, u64* keep, u64* scratch) {
    uint32_t* hist = (uint32_t*)scratch;                       // d + 1 bins
    uint32_t* ord = (uint32_t*)(scratch + (d + 2) / 2);        // nq positions
    u64* kcols = scratch + (d + 2) / 2 + (nq + 1) / 2;         // d * kKW
    for (uint32_t i = lane(); i <= d; i += 64) hist[i] = 0;
    for (size_t i = lane(); i < (size_t)d * kKW; i += 64) kcols[i] = 0;
    sync_mem();
    // counting sort of the positions by popcount, descending
    for (uint32_t i = lane(); i < nq; i += 64) {
        const uint32_t q = qs ? qs[i] : q0 + i;
        uint32_t pc = 0;
        for (uint32_t w = 0; w < W; w++) pc += __popcll(rows[(size_t)q * W + w]);
        atomicAdd(&hist[d - pc], 1u);
    }
    sync_mem();
    uint32_t carry = 0;
    for (uint32_t b = 0; b <= d; b += 64) {
        const uint32_t i = b + lane();
        const uint32_t v = i <= d ? hist[i] : 0;
        const uint32_t inc = scan_add32(v);
        if (i <= d) hist[i] = carry + inc - v;
        carry += rdlane32(inc, 63);
    }
    sync_mem();
    for (uint32_t i = lane(); i < nq; i += 64) {
        const uint32_t q = qs ? qs[i] : q0 + i;
        uint32_t pc = 0;
        for (uint32_t w = 0; w < W; w++) pc += __popcll(rows[(size_t)q * W + w]);
        ord[atomicAdd(&hist[d - pc], 1u)] = q;
    }
    sync_mem();
    uint32_t nk = 0, kept = 0;
    for (uint32_t b = 0; b < nq; b += 64) {
        const uint32_t i = b + lane();
        const bool valid = i < nq;
        const uint32_t q = valid ? ord[i] : 0;
        const u64* r = rows + (size_t)q * W;
        bool dom = false;
        if (valid && nk) {
            const uint32_t nkw = (nk + 63) >> 6;
            for (uint32_t w = 0; w < nkw && !dom; w++) {
                u64 a = (w == nkw - 1 && (nk & 63)) ? ((1ull << (nk & 63)) - 1) : ~0ull;
                for (uint32_t rw = 0; rw < W && a; rw++) {
                    u64 bits = r[rw];
                    while (bits && a) {
                        const uint32_t j = rw * 64 + __builtin_ctzll(bits);
                        bits &= bits - 1;
                        a &= kcols[(size_t)j * kKW + w];
                    }
                }
                dom = a != 0;
            }
        }
        const bool surv = valid && !dom;
        const u64 sm = __ballot(surv);
        const uint32_t cnt = __popcll(sm);
        if (surv) atomicOr(&keep[q >> 6], 1ull << (q & 63));
        if (nk + cnt <= kKeepCap) {
            if (surv) {
                const uint32_t t = nk + __popcll(sm & lanemask_lt());
                for (uint32_t rw = 0; rw < W; rw++) {
                    u64 bits = r[rw];
                    while (bits) {
                        const uint32_t j = rw * 64 + __builtin_ctzll(bits);
                        bits &= bits - 1;
                        atomicOr(&kcols[(size_t)j * kKW + (t >> 6)], 1ull << (t & 63));
                    }
                }
            }
            nk += cnt;
        }
        kept += cnt;
        sync_mem();
    }
    return kept;
}

// Source sets of a compaction: the candidates of P at index >= t stay in P,
// those below t join Q (they precede the handed-over part), X stays X, and
// candidate xq (the level's in-progress x, or none) joins Q.
struct SplitSrc {
    const u64 *L, *P, *Q, *X;
    uint32_t t, xq;
    __device__ __forceinline__ u64 pw(uint32_t w) const {
        const u64 m = w * 64 >= t ? ~0ull : (w * 64 + 64 <= t ? 0ull : ~0ull << (t & 63));
        return P[w] & m;
    }
    __device__ __forceinline__ u64 qw(uint32_t w) const {
        u64 v = Q[w] | (P[w] & ~pw_mask(w));
        if (w == (xq >> 6)) v |= 1ull << (xq & 63);
        return v;
    }
    __device__ __forceinline__ u64 pw_mask(uint32_t w) const {
        return w * 64 >= t ? ~0ull : (w * 64 + 64 <= t ? 0ull : ~0ull << (t & 63));
    }
    __device__ __forceinline__ u64 cw(uint32_t w) const { return pw(w) | qw(w) | X[w]; }
};

// |L| and the number of candidates of a compaction source.
__device__ __forceinline__ void frame_sizes(const Local& S, const SplitSrc& src, uint32_t& dn,
                                            uint32_t& n) {
    uint32_t a = 0, b = 0;
    for (uint32_t w = lane(); w < S.W; w += 64) a += __popcll(src.L[w]);
    for (uint32_t w = lane(); w < S.Wc; w += 64) b += __popcll(src.cw(w));
    dn = red_add32(a);
    n = red_add32(b);
}

// Scratch of a compaction at the arena's tail, above every live problem:
// each candidate's row over the new universe, the keep bitmap, the Q
// positions and reduce_q's scratch.
struct CompactTmp {
    u64* trow;     // n * Wn
    u64* keep;     // (n + 63) / 64
    uint32_t* qs;  // n
    u64* rq;       // reduce_scratch_words(dn, n)
    size_t words;
};
__device__ __forceinline__ size_t compact_tmp_words(uint32_t dn, uint32_t n) {
    return (size_t)n * ((dn + 63) >> 6) + (n + 63) / 64 + (n + 1) / 2 + reduce_scratch_words(dn, n);
}
// Place the scratch at the tail of the arena; false if it would reach down
// to word `floor` (the end of the arena words in use).
__device__ bool compact_tmp(const KParams& kp, u64* arena, size_t floor, uint32_t dn, uint32_t n,
                            CompactTmp& t) {
    t.words = compact_tmp_words(dn, n);
    if (floor + t.words > kp.arena_words) return false;
    u64* p = arena + (kp.arena_words - t.words);
    t.trow = p;
    p += (size_t)n * ((dn + 63) >> 6);
    t.keep = p;
    p += (n + 63) / 64;
    t.qs = (uint32_t*)p;
    p += (n + 1) / 2;
    t.rq = p;
    return true;
}

// Compaction of a source of problem S into a smaller universe, phase 1: the
// source's candidates (S.list, ascending) and L bits (S.lbits), each
// candidate's row over the L bits renumbered 0..dn-1 (tmp.trow), and the
// candidates kept (tmp.keep): every P and X candidate and the Q candidates
// reduce_q keeps. Returns how many are kept. S.list and S.lbits are
// overwritten (scratch).
__device__ uint32_t compact_prepare(const Local& S, const SplitSrc& src, uint32_t dn, uint32_t n,
                                    const CompactTmp& tmp) {
    const uint32_t Wn = (dn + 63) >> 6;
    uint32_t n0 = 0;
    for (uint32_t w = 0; w < S.Wc; w++) {
        const u64 cw = uni64(src.cw(w));
        if ((cw >> lane()) & 1) S.list[n0 + __popcll(cw & lanemask_lt())] = w * 64 + lane();
        n0 += __popcll(cw);
    }
    uint32_t d0 = 0;
    for (uint32_t w = 0; w < S.W; w++) {
        const u64 word = uni64(src.L[w]);
        if ((word >> lane()) & 1) S.lbits[d0 + __popcll(word & lanemask_lt())] = w * 64 + lane();
        d0 += __popcll(word);
    }
    wsync();
    if (S.sparse) {
        // list mode: each candidate's row from the candidate lists of the new
        // universe's L bits (S.list is ascending: binary search)
        for (size_t i = lane(); i < (size_t)n * Wn; i += 64) tmp.trow[i] = 0;
        sync_mem();
        for (uint32_t t = lane(); t < dn; t += 64) {
            const uint32_t j = S.lbits[t];
            for (uint32_t e = S.lofs[j]; e < S.lofs[j + 1]; e++) {
                const uint32_t k = S.lcand[e];
                uint32_t lo = 0, hi = n;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (S.list[mid] < k) lo = mid + 1; else hi = mid;
                }
                if (lo < n && S.list[lo] == k)
                    atomicOr(&tmp.trow[(size_t)lo * Wn + (t >> 6)], 1ull << (t & 63));
            }
        }
        sync_mem();
    }
    uint32_t nq = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
        const uint32_t kn = c0 + lane();
        const bool valid = kn < n;
        const uint32_t k = valid ? S.list[kn] : 0;
        const u64* r = S.rows + (size_t)k * S.W;
        for (uint32_t w2 = 0; w2 < Wn && !S.sparse; w2++) {
            u64 acc = 0;
            const uint32_t lim = min(64u, dn - w2 * 64);
            for (uint32_t t = 0; t < lim; t++) {
                const uint32_t j = S.lbits[w2 * 64 + t];
                if (valid) acc |= ((r[j >> 6] >> (j & 63)) & 1ull) << t;
            }
            if (valid) tmp.trow[(size_t)kn * Wn + w2] = acc;
        }
        const bool inQ = valid && ((src.qw(k >> 6) >> (k & 63)) & 1);
        const u64 bq = __ballot(inQ);
        const u64 bk = __ballot(valid && !inQ);
        if (lane() == 0) tmp.keep[c0 >> 6] = bk;
        if (inQ) tmp.qs[nq + __popcll(bq & lanemask_lt())] = kn;
        nq += __popcll(bq);
    }
    sync_mem();
    if (nq < kReduceMinQ) {
        for (uint32_t w = lane(); w < (n + 63) / 64; w += 64)
            tmp.keep[w] = (w == (n - 1) / 64 && (n & 63)) ? ((1ull << (n & 63)) - 1) : ~0ull;
        sync_mem();
        return n;
    }
    return n - nq + reduce_q(tmp.trow, Wn, dn, tmp.qs, 0, nq, tmp.keep, tmp.rq);
}

// Phase 2: the kept candidates renumbered 0..nk-1 in order: hl2 (dn terms),
// hc2, rows2 (Wn words each), and P2, Q2, X2 (Wcn words each, cleared here).
__device__ void compact_write(const Local& S, const SplitSrc& src, uint32_t dn, uint32_t n,
                              uint32_t nk, const CompactTmp& tmp, u64* hl2, u64* hc2, u64* rows2,
                              u64* P2, u64* Q2, u64* X2) {
    const uint32_t Wn = (dn + 63) >> 6, Wcn = (nk + 63) >> 6;
    for (uint32_t i = lane(); i < dn; i += 64) hl2[i] = S.hl[S.lbits[i]];
    for (uint32_t w = lane(); w < Wcn; w += 64) {
        P2[w] = 0;
        Q2[w] = 0;
        X2[w] = 0;
    }
    sync_mem();
    uint32_t at = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
        const uint32_t kn = c0 + lane();
        const u64 kw = uni64(tmp.keep[c0 >> 6]);
        if ((kw >> lane()) & 1) {
            const uint32_t to = at + __popcll(kw & lanemask_lt());
            const uint32_t k = S.list[kn];
            hc2[to] = S.hc[k];
            for (uint32_t w2 = 0; w2 < Wn; w2++)
                rows2[(size_t)to * Wn + w2] = tmp.trow[(size_t)kn * Wn + w2];
            const uint32_t w = k >> 6, b = k & 63;
            u64* dst = ((src.pw(w) >> b) & 1) ? P2 : ((src.qw(w) >> b) & 1) ? Q2 : X2;
            atomicOr(&dst[to >> 6], 1ull << (to & 63));
        }
        at += __popcll(kw);
    }
    sync_mem();
}

// Write a compaction source into a slot as a LEVEL call over its own
// universe: [0] graph, [1] d, [2] nc, [3] R hash sum, [4] |R|, then hl[d],
// hc[nc], rows[nc * W], P[Wc], Q[Wc], X[Wc]. The scratch goes to the tail of
// the donor's arena (room checked by snapshot_words).
__device__ void write_snapshot(const KParams& kp, const Local& S, const SplitSrc& src, u64* arena,
                               size_t floor, u64 sr, u64 nr, uint32_t g, u64* slot) {
    uint32_t dn, n;
    frame_sizes(S, src, dn, n);
    CompactTmp tmp;
    compact_tmp(kp, arena, floor, dn, n, tmp);
    const uint32_t nk = compact_prepare(S, src, dn, n, tmp);
    const uint32_t Wn = (dn + 63) >> 6, Wcn = (nk + 63) >> 6;
    u64* hl2 = slot + 8;
    u64* hc2 = hl2 + dn;
    u64* rows2 = hc2 + nk;
    u64* P2 = rows2 + (size_t)nk * Wn;
    u64* Q2 = P2 + Wcn;
    u64* X2 = Q2 + Wcn;
    compact_write(S, src, dn, n, nk, tmp, hl2, hc2, rows2, P2, Q2, X2);
    if (lane() == 0) {
        slot[0] = g;
        slot[1] = dn;
        slot[2] = nk;
        slot[3] = sr;
        slot[4] = nr;
    }
}

__device__ __forceinline__ void fill_full_l(const Local& S, u64* L1) {
    for (uint32_t w = lane(); w < S.W; w += 64)
        L1[w] = (w == S.W - 1 && (S.d & 63)) ? ((1ull << (S.d & 63)) - 1) : ~0ull;
}

// Load a slot into the worker's memory as a ready frame 1; false if too big.
__device__ bool load_snapshot(const KParams& kp, Local& S, u64* arena, u64* lds, const u64* slot) {
    S.d = (uint32_t)slot[1];
    S.nc = (uint32_t)slot[2];
    S.x = 0xFFFFFFFFu;
    S.kx = 0;  // kx is a root problem's field; nest records keep it
    if (!layout(kp, S, arena, lds, kArenaHdr, kLdsHdr)) return false;
    const u64* hl2 = slot + 8;
    const u64* hc2 = hl2 + S.d;
    const u64* rows2 = hc2 + S.nc;
    const u64* P2 = rows2 + (size_t)S.nc * S.W;
    const u64* Q2 = P2 + S.Wc;
    const u64* X2 = Q2 + S.Wc;
    for (uint32_t i = lane(); i < S.d; i += 64) S.hl[i] = hl2[i];
    for (uint32_t i = lane(); i < S.nc; i += 64) S.hc[i] = hc2[i];
    const size_t nr = (size_t)S.nc * S.W;
    for (size_t i = lane(); i < nr; i += 64) S.rows[i] = rows2[i];
    wsync();
    build_cols(S);
    u64* F1 = S.frame(1);
    u64* L1 = F1 + kFrameHdr;
    u64* P1 = L1 + S.W;
    u64* Q1 = P1 + S.Wc;
    u64* X1 = Q1 + S.Wc;
    fill_full_l(S, L1);
    uint32_t pc = 0, qc = 0;
    for (uint32_t w = lane(); w < S.Wc; w += 64) {
        P1[w] = P2[w];
        Q1[w] = Q2[w];
        X1[w] = X2[w];
        pc += __popcll(P2[w]);
        qc += __popcll(Q2[w]);
    }
    pc = red_add32(pc);
    qc = red_add32(qc);
    if (lane() == 0) {
        F1[1] = slot[3];
        F1[2] = slot[4];
        F1[3] = pq(pc, qc);
    }
    wsync();
    return true;
}

// Replace the pending descent into child frame Fc of S by a nested problem
// S2 over the child's own universe (L', P' u Q'), placed after S in the
// arena and after S's live frames in LDS. Returns false if it does not fit.
__device__ bool compact_child(const KParams& kp, const Local& S, uint32_t k, const u64* Fc,
                              Local& S2, u64* arena, u64* lds, uint32_t dn, uint32_t n,
                              uint32_t pn) {
    const u64* L = Fc + kFrameHdr;
    const u64* P = L + S.W;
    const u64* Q = P + S.Wc;
    const SplitSrc src{L, P, Q, Q + S.Wc, 0, 0xFFFFFFFFu};
    CompactTmp tmp;
    if (!compact_tmp(kp, arena, S.end, dn, n, tmp)) return false;
    const uint32_t nk = compact_prepare(S, src, dn, n, tmp);
    S2.d = dn;
    S2.nc = nk;
    S2.x = 0xFFFFFFFFu;
    S2.kx = 0;
    // after S's LDS data and scratch and its live frames (1..k+1)
    const uint32_t lds_from =
        (uint32_t)(S.fbase - lds) + min(k + 1, S.nlds) * S.fs;
    if (!layout(kp, S2, arena, lds, S.end, lds_from) || S2.end + tmp.words > kp.arena_words)
        return false;
    u64* F1 = S2.frame(1);
    u64* L1 = F1 + kFrameHdr;
    u64* P1 = L1 + S2.W;
    u64* Q1 = P1 + S2.Wc;
    u64* X1 = Q1 + S2.Wc;
    const u64 sr = Fc[1], nr = Fc[2];
    compact_write(S, src, dn, n, nk, tmp, S2.hl, S2.hc, S2.rows, P1, Q1, X1);
    build_cols(S2);
    fill_full_l(S2, L1);
    if (lane() == 0) {
        F1[1] = sr;
        F1[2] = nr;
        F1[3] = pq(pn, nk - pn);
    }
    wsync();
    return true;
}

// Size in words of a snapshot of the given source (before Q reduction), or
// ~0 when the snapshot or the problem it becomes would not fit a slot or a
// thief's arena, or its compaction scratch would not fit above word `floor`
// of the donor's arena.
__device__ size_t snapshot_words(const KParams& kp, const Local& S, const SplitSrc& src,
                                 size_t floor) {
    uint32_t dn, n;
    frame_sizes(S, src, dn, n);
    if (!dn || floor + compact_tmp_words(dn, n) > kp.arena_words) return ~(size_t)0;
    const size_t Wn = (dn + 63) >> 6, Wcn = (n + 63) >> 6;
    const size_t words = 8 + (size_t)dn + n + (size_t)n * Wn + 3 * Wcn;
    // the thief's layout with nothing in LDS (see layout())
    const size_t thief = (size_t)dn + n + (size_t)n * Wn + (size_t)dn * Wcn + (n + 1) / 2 + 1 +
                         (dn + 1) / 2 + 1 + (size_t)(min(dn, n) + 2) * (kFrameHdr + Wn + 3 * Wcn);
    if (words > kp.slot_words || kArenaHdr + thief > kp.arena_words) return ~(size_t)0;
    return words;
}

// Rebuild a suspended problem from its nest record.
__device__ __forceinline__ void unpack(const KParams& kp, const u64* rec, Local& T, u64* arena,
                                       u64* lds, uint32_t& k, uint32_t& depth0) {
    T.d = (uint32_t)rec[1];
    T.nc = (uint32_t)(rec[1] >> 32);
    T.x = 0xFFFFFFFFu;
    T.kx = (uint32_t)(rec[3] >> 32) & 0x7FFFFFFFu;
    if (rec[3] >> 63)
        layout_sparse(kp, T, arena, lds, (size_t)rec[0], (uint32_t)rec[2], 0xFFFFFFFFu);
    else
        layout(kp, T, arena, lds, (size_t)rec[0], (uint32_t)rec[2]);
    k = (uint32_t)(rec[2] >> 32);
    depth0 = (uint32_t)rec[3];
}

// Index t of the (h+1)-th set bit of P (so h bits lie below t).
__device__ uint32_t split_point(const Local& S, const u64* P, uint32_t h) {
    uint32_t seen = 0;
    for (uint32_t w = 0; w < S.Wc; w++) {
        const u64 pw = uni64(P[w]);
        const uint32_t c = __popcll(pw);
        if (seen + c > h) {
            u64 m = pw;
            for (uint32_t i = seen; i < h; i++) m &= m - 1;
            return w * 64 + __builtin_ctzll(m);
        }
        seen += c;
    }
    return S.Wc * 64;
}

// Hand part of a level to the deque: the shallowest level (suspended outer
// problems first, then the current one) that still has unselected work.
// A level whose x is in progress gives away its remaining P whole when it
// holds one candidate, else its upper half; the current level (no x yet)
// gives its upper half when it holds two or more. The donor keeps the lower
// part A and moves the given part B into X; the thief gets P = B,
// Q = Q u A (u x), X = X: the order "x, A, B, X" is a valid level order, so
// both sides enumerate exactly the nodes of the sequential walk.
__device__ bool try_donate(const KParams& kp, const Local& S, uint32_t cur, uint32_t g,
                           u64* arena, u64* lds, const u64* nest, uint32_t np) {
    for (uint32_t i = 0; i <= np; i++) {
        Local T;
        uint32_t top, d0;
        if (i < np) {
            unpack(kp, nest + 4 * i, T, arena, lds, top, d0);
        } else {
            T = S;
            top = cur;
        }
        if (kp.donate_depth) top = min(top, kp.donate_depth);
        for (uint32_t k = 1; k <= top; k++) {
            u64* F = T.frame(k);
            const uint32_t pc = pcount(F);
            const bool running = i < np || k < cur;  // x[k] in progress
            if (pc == 0 || (!running && pc < 2)) continue;
            u64* L = F + kFrameHdr;
            u64* P = L + T.W;
            u64* Q = P + T.Wc;
            u64* X = Q + T.Wc;
            const uint32_t keep = pc / 2;  // A: the lowest `keep` candidates
            const uint32_t t = keep ? split_point(T, P, keep) : 0;
            const SplitSrc src{L, P, Q, X, t, running ? (uint32_t)F[0] : 0xFFFFFFFFu};
            if (snapshot_words(kp, T, src, S.end) == ~(size_t)0) continue;
            if (lane() == 0) {  // pending and the graph's open tasks, before publishing
                atomicAdd(kp.dq + 3, 1ull);
                if (kp.remaining) atomicAdd(kp.remaining + 2 * (size_t)g, 1ull);
            }
            int64_t pos = dq_claim_push(kp);
            if (pos < 0) {
                if (lane() == 0) {
                    atomicAdd(kp.dq + 3, ~0ull);
                    if (kp.remaining) atomicAdd(kp.remaining + 2 * (size_t)g, ~0ull);
                }
                return false;
            }
            u64* slot = (u64*)kp.slots + (size_t)((u64)pos % kp.n_slots) * kp.slot_words;
            write_snapshot(kp, T, src, arena, S.end, F[1], F[2], g, slot);
            dq_publish(kp, (u64)pos);
            for (uint32_t w = lane(); w < T.Wc; w += 64) {
                const u64 b = src.pw(w);
                X[w] |= b;
                P[w] &= ~b;
            }
            if (lane() == 0) F[3] = (F[3] & ~0xFFFFFFFFull) | keep;  // |P| = |A|
            wsync();
            return true;
        }
    }
    return false;
}

// Iterative DFS over levels >= 1 from a ready frame 1 of S
// (src/mbe_noWS.cuh:337-615), with donation to the deque in place of
// src/mbe_cuMBE.cuh's level-1 publish/steal, and descents into children
// whose universe shrank by compact_ratio re-localised as nested problems.
// SP: the outermost problem may be a list-mode root (its children are bit
// matrices); without it the node loop holds no list-mode code.
template <int V, bool SP>
__device__ void dfs(const KParams& kp, const DevGraph& G, uint32_t g, Local& S, u64* arena,
                    u64* lds, uint32_t* premin, u64& cnt, u64& hs, u64& nodes, u64& donated,
                    u64& words, Sec& sec) {
    u64* nest = lds + kLdsNest;
    uint32_t np = 0, depth0 = 0, k = 1;
    const u64 split_at = kp.split_nodes ? nodes + kp.split_nodes : ~0ull;
    u64 next_check = nodes + kp.donate_interval;
    u64 next_abort = nodes + kAbortEvery;
    for (;;) {
        if (nodes >= next_abort) {
            next_abort = nodes + kAbortEvery;
            if (charge_nodes(kp, kAbortEvery)) break;
        }
        u64* F = S.frame(k);
        const u64 cnts = uni64(F[3]);
        const uint32_t pc = (uint32_t)cnts, qc = (uint32_t)(cnts >> 32);
        if (pc == 0) {
            k--;
            if (k == 0) {
                if (np == 0) break;
                np--;
                unpack(kp, nest + 4 * np, S, arena, lds, k, depth0);
            }
            // ascend: Q[k] <- Q[k] u {x[k]} (src/mbe_noWS.cuh:590-601)
            u64* Fp = S.frame(k);
            if (lane() == 0) {
                const uint32_t xp = (uint32_t)Fp[0];
                u64* Qp = Fp + kFrameHdr + S.W + S.Wc;
                Qp[xp >> 6] |= 1ull << (xp & 63);
                Fp[3] += 1ull << 32;  // |Q|
            }
            wsync();
            continue;
        }
        // donate when workers idle and the deque holds fewer tasks than
        // there are idle workers, or when this task has run split_nodes
        // nodes and fewer than kSplitQueue tasks wait; at most once per
        // donate_interval nodes
        if (V != kNoWS && (k > 1 || np > 0) && nodes >= next_check) {
            next_check = nodes + kp.donate_interval;
            u64 want = 0;
            if (lane() == 0) {
                const u64 idle = ld_rlx(kp.dq + 2);
                const u64 queued = ld_rlx(kp.dq + 1) - ld_rlx(kp.dq);
                want = idle > queued || (nodes >= split_at && queued < kSplitQueue);
            }
            if (uni64(want) && try_donate(kp, S, k, g, arena, lds, nest, np)) {
                donated++;
                sec.mark(kSecDonate);
                continue;  // this level may have been handed over
            }
        }
        sec.mark(kSecOther);
        u64* L = F + kFrameHdr;
        uint32_t* pm = premin + min(depth0 + k, kPreMinLevels - 1);
        const uint32_t xk = select_x<V>(S, F, pm);
        nodes++;
        if (lane() == 0) F[0] = xk;
        sec.mark(kSecSelect);
        u64* Fc = S.frame(k + 1);
        const uint32_t nl = make_lprime(S, L, xk, Fc + kFrameHdr);
        sec.mark(kSecLPrime);
        const bool fwd = V == kNoRS || pc - 1 + qc <= kp.list_cutoff;
        NodeOut o;
        if (SP && S.sparse) {  // list-mode root problem: the same check in every variant
            u64 sc = 0;
            o = process_sparse(S, F, Fc, nl, pc > 1 && kp.twins, sc);
            words += (u64)pc * S.W + S.Wc + 2 * S.W + nl + sc + 3 * S.Wc;
        } else {
            o = fwd ? process_fwd(S, F, Fc, nl) : process_rs(S, F, Fc, nl);
            // operand words of this node: selection rows and P words, x's row
            // and L, the L' columns (or the P and Q rows) with P/Q/X, L' terms
            words += (u64)pc * S.W + S.Wc + 2 * S.W + nl +
                     (fwd ? (u64)(pc - 1 + qc) * S.W + 3 * S.Wc : (u64)S.Wc * (nl + 3));
        }
        if (pc > 1 && kp.twins && !(SP && S.sparse)) {
            const u64 r = prune_twins(F, Fc + kFrameHdr, S.cols, S.rows, S.lbits, S.list, S.W,
                                      S.Wc, nl, o.tw, o.has_tw);
            words += (o.has_tw ? 0 : (u64)S.Wc * (nl + 1)) + (r >> 32) * S.W;
        }
        sec.mark(o.maximal ? kSecExpand : kSecCheck);
        if (o.maximal) {
            const u64 sl = sum_l(S, nl);
            const u64 sr = F[1] + S.hc[xk] + o.abh;
            const u64 nr = F[2] + 1 + o.ab;
            record(G, sl, sr, nl, nr, cnt, hs);
            if (kp.emit_pos) emit_biclique(kp.emit_pos, kp.emit, kp.emit_cap, S.hl, S.lbits, nl, nr, sr);
            sec.mark(kSecReport);
            if (o.descend) {
                if (lane() == 0) {
                    Fc[1] = sr;
                    Fc[2] = nr;
                    Fc[3] = pq(o.pn, o.qn);
                }
                wsync();
                const uint32_t Wcn = (o.un + 63) >> 6, Wn = (nl + 63) >> 6;
                if (kp.compact_ratio && np < kNestMax && o.pn >= 2 &&
                    (Wcn * kp.compact_ratio <= S.Wc || Wn * kp.compact_ratio <= S.W)) {
                    Local S2;
                    const bool okc = compact_child(kp, S, k, Fc, S2, arena, lds, nl, o.un, o.pn);
                    sec.mark(kSecCompact);
                    if (okc) {
                        if (lane() == 0) {
                            u64* r = nest + 4 * np;
                            r[0] = S.base;
                            r[1] = (u64)S.d | ((u64)S.nc << 32);
                            r[2] = (u64)S.lds_off | ((u64)k << 32);
                            r[3] = depth0 | ((u64)(S.kx & 0x7FFFFFFFu) << 32) |
                                   ((u64)S.sparse << 63);
                        }
                        wsync();
                        np++;
                        depth0 += k;
                        S = S2;
                        k = 1;
                        continue;
                    }
                }
                k++;
                continue;
            }
        }
        if (lane() == 0) {
            u64* Q = L + S.W + S.Wc;
            Q[xk >> 6] |= 1ull << (xk & 63);
            F[3] += 1ull << 32;  // |Q|
        }
        wsync();
    }
}

// List-mode localisation of root x (the candidates are numbered, S.cand and
// cpos hold them): the candidates of each L bit (N(u_j) without x, in rank
// order, from the degree prefix), the L bits of each candidate (a counting
// transpose), the bit rows of the root's P candidates, then the root node.
// Returns like build_root.
__device__ __noinline__ int build_root_sparse(const KParams& kp, const DevGraph& G, uint32_t x,
                                              u64* arena, u64* lds, const uint32_t* cpos,
                                              Local& S, uint32_t T, uint32_t nc, uint32_t kx,
                                              size_t off, u64& cnt, u64& hs, u64& nodes) {
    S.nc = nc;
    S.kx = kx;
    const uint32_t ntl = T - S.d;  // every N(u_j) holds x once
    if (!layout_sparse(kp, S, arena, lds, off, kLdsHdr, ntl)) return -1;
    const uint32_t W = S.W;
    for (uint32_t j = lane(); j < S.d; j += 64) S.hl[j] = G.hL[S.Nx[j]];
    for (uint32_t k = lane(); k < nc; k += 64) S.hc[k] = G.hR[S.cand[k]];
    for (uint32_t j = lane(); j <= S.d; j += 64) S.lofs[j] = S.pre[j] - j;
    for (size_t i = lane(); i < (size_t)kx * W; i += 64) S.rows[i] = 0;
    for (uint32_t k = lane(); k <= nc; k += 64) S.list[k] = 0;  // per-candidate counts
    sync_mem();
    for (uint32_t f0 = 0; f0 < T; f0 += 64) {
        const uint32_t f = f0 + lane();
        if (f < T) {
            const uint32_t j = find_j(S.pre, S.d, f);
            const uint32_t u = S.Nx[j];
            const uint32_t t = f - S.pre[j];
            const uint32_t v = G.adjL[G.offL[u] + t];
            if (v != x) {
                const uint32_t k = cpos[v];
                S.lcand[S.lofs[j] + (v < x ? t : t - 1)] = k;
                atomicAdd(&S.list[k], 1u);
                if (k < kx) atomicOr(&S.rows[(size_t)k * W + (j >> 6)], 1ull << (j & 63));
            }
        }
    }
    sync_mem();
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 <= nc; c0 += 64) {
        const uint32_t k = c0 + lane();
        const uint32_t v = k < nc ? S.list[k] : 0;
        const uint32_t inc = scan_add32(v);
        if (k <= nc) {
            S.rofs[k] = carry + inc - v;
            S.list[k] = carry + inc - v;  // scatter cursor
        }
        carry += rdlane32(inc, 63);
    }
    sync_mem();
    for (uint32_t f0 = 0; f0 < T; f0 += 64) {
        const uint32_t f = f0 + lane();
        if (f < T) {
            const uint32_t j = find_j(S.pre, S.d, f);
            const uint32_t u = S.Nx[j];
            const uint32_t v = G.adjL[G.offL[u] + f - S.pre[j]];
            if (v != x) S.rl[atomicAdd(&S.list[cpos[v]], 1u)] = j;
        }
    }
    sync_mem();
    // root node: L' = N(x); a candidate adjacent to all of it is full
    bool nonmax = false;
    uint32_t ab = 0;
    u64 abh = 0;
    for (uint32_t c0 = 0; c0 < nc && !nonmax; c0 += 64) {
        const uint32_t k = c0 + lane();
        const bool full = k < nc && S.rofs[k + 1] - S.rofs[k] == S.d;
        if (__ballot(full && k >= kx)) nonmax = true;
        if (full && k < kx) {
            ab++;
            abh += S.hc[k];
        }
    }
    nodes++;
    if (nonmax) return 0;
    ab = red_add32(ab);
    abh = red_add64(abh);
    const u64 sr = G.hR[x] + abh, nr = 1 + ab;
    u64 sl = 0;
    for (uint32_t j = lane(); j < S.d; j += 64) sl += S.hl[j];
    sl = red_add64(sl);
    record(G, sl, sr, S.d, nr, cnt, hs);
    if (kp.emit_pos) emit_biclique(kp.emit_pos, kp.emit, kp.emit_cap, S.hl, nullptr, S.d, nr, sr);
    if (kx == ab) return 0;  // P' empty
    u64* F1 = S.frame(1);
    u64* L1 = F1 + kFrameHdr;
    u64* P1 = L1 + S.W;
    u64* Q1 = P1 + S.Wc;
    u64* X1 = Q1 + S.Wc;
    fill_full_l(S, L1);
    for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
        const uint32_t k = c0 + lane();
        const bool full = k < kx && S.rofs[k + 1] - S.rofs[k] == S.d;
        const u64 bp = __ballot(k < kx && !full), bq = __ballot(k >= kx && k < nc);
        if (lane() == 0) {
            P1[c0 >> 6] = bp;
            Q1[c0 >> 6] = bq;
            X1[c0 >> 6] = 0;
        }
    }
    if (lane() == 0) {
        F1[1] = sr;
        F1[2] = nr;
        F1[3] = pq(kx - ab, nc - kx);
    }
    wsync();
    return 1;
}

// The search of a list-mode root, out of line: the bit-matrix search that
// every other task runs keeps its registers.
template <int V>
__device__ __noinline__ void dfs_sparse(const KParams& kp, const DevGraph& G, uint32_t g, Local& S,
                                        u64* arena, u64* lds, uint32_t* premin, u64& cnt, u64& hs,
                                        u64& nodes, u64& donated, u64& words, Sec& sec) {
    dfs<V, true>(kp, G, g, S, arena, lds, premin, cnt, hs, nodes, donated, words, sec);
}

// Localise root x (rows/cols over N(x) x distance-two candidates) and
// process the root node. Returns 0 done, 1 DFS from frame 1 needed, -1 does
// not fit the arena (nothing has been reported then, so a later pass with
// a larger arena can redo the root).
//
// Arena: [degree prefix][candidate ranks][the problem (unless it all fits
// LDS)] ... [kept-Q list][rows of every candidate] at the tail. The rows are
// built first at the tail; the root node is decided on them; the Q
// candidates are reduced (reduce_q); then the problem is laid out with the
// root's P (ranks below x, positions 0..kx-1) followed by the kept Q.
template <int V>
__device__ int build_root(const KParams& kp, const DevGraph& G, uint32_t x, u64* arena, u64* lds,
                          u64* mark, uint32_t* cpos, Local& S, u64& cnt, u64& hs, u64& nodes,
                          u64& build) {
    S.x = x;
    const uint32_t s0 = G.offR[x];
    S.d = G.offR[x + 1] - s0;
    S.Nx = G.adjR + s0;
    if (S.d == 0) return 0;
    size_t off = kArenaHdr;
    S.pre = (uint32_t*)(arena + off);
    off += (S.d + 2) / 2;
    S.cand = (uint32_t*)(arena + off);
    const size_t off_cand = off;
    if (off + S.d > kp.arena_words) return -1;

    // degree prefix over N(x)
    uint32_t carry = 0;
    for (uint32_t j0 = 0; j0 < S.d; j0 += 64) {
        const uint32_t j = j0 + lane();
        uint32_t dg = 0;
        if (j < S.d) {
            const uint32_t u = S.Nx[j];
            dg = G.offL[u + 1] - G.offL[u];
        }
        const uint32_t inc = scan_add32(dg);
        if (j < S.d) S.pre[j] = carry + inc - dg;
        carry += rdlane32(inc, 63);
    }
    if (lane() == 0) S.pre[S.d] = carry;
    const uint32_t T = carry;
    build += T;
    wsync();

    // discover the distance-two candidates (bitmap over R ranks)
    for (uint32_t f0 = 0; f0 < T; f0 += 64) {
        const uint32_t f = f0 + lane();
        if (f < T) {
            const uint32_t j = find_j(S.pre, S.d, f);
            const uint32_t u = S.Nx[j];
            const uint32_t v = G.adjL[G.offL[u] + f - S.pre[j]];
            if (v != x) atomicOr(&mark[v >> 6], 1ull << (v & 63));
        }
    }
    // this wave's atomics, then its own loads: an acquire-only fence orders
    // them (nothing here is published to other waves)
    sync_mem();
    const uint32_t nwords = (G.nR + 63) >> 6;
    uint32_t ncl = 0, kxl = 0;
    for (uint32_t w = lane(); w < nwords; w += 64) {
        const u64 m = mark[w];
        ncl += __popcll(m);
        if (w < (x >> 6)) kxl += __popcll(m);
        else if (w == (x >> 6)) kxl += __popcll(m & ((1ull << (x & 63)) - 1));
    }
    const uint32_t nc = red_add32(ncl);
    const uint32_t kx = red_add32(kxl);
    S.kx = kx;
    const uint32_t W = (S.d + 63) >> 6;
    off = off_cand + (nc + 1) / 2;
    if (off > kp.arena_words) {
        for (uint32_t w = lane(); w < nwords; w += 64) mark[w] = 0;
        sync_mem();
        return -1;
    }

    // assign local indices in rank order; clear the bitmap
    uint32_t k0 = 0;
    for (uint32_t w0 = 0; w0 < nwords; w0 += 64) {
        const uint32_t w = w0 + lane();
        u64 m = w < nwords ? mark[w] : 0;
        const uint32_t c = __popcll(m);
        const uint32_t inc = scan_add32(c);
        uint32_t k = k0 + inc - c;
        if (m) mark[w] = 0;
        while (m) {
            const uint32_t b = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t v = w * 64 + b;
            S.cand[k] = v;
            cpos[v] = k;
            k++;
        }
        k0 += rdlane32(inc, 63);
    }
    // list mode when the bit matrices would dwarf the local adjacency and
    // the root's tree is small (at most 2^kSparseMaxP nodes): a list-mode
    // problem is searched by its own worker
    if (kp.sparse_ratio && kx <= kSparseMaxP &&
        (size_t)nc * W + (size_t)S.d * ((nc + 63) >> 6) > (size_t)kp.sparse_ratio * T) {
        sync_mem();
        return build_root_sparse(kp, G, x, arena, lds, cpos, S, T, nc, kx, off, cnt, hs, nodes);
    }
    const size_t trow_words = (size_t)nc * W;
    const size_t tail_words = trow_words + (nc + 1) / 2 + 1;
    if (off + max(tail_words, (size_t)S.d) > kp.arena_words) return -1;
    u64* trow = arena + (kp.arena_words - trow_words);
    uint32_t* kidx = (uint32_t*)(arena + (kp.arena_words - tail_words));
    for (size_t i = lane(); i < trow_words; i += 64) trow[i] = 0;
    sync_mem();
    // the candidates' rows over N(x)
    for (uint32_t f0 = 0; f0 < T; f0 += 64) {
        const uint32_t f = f0 + lane();
        if (f < T) {
            const uint32_t j = find_j(S.pre, S.d, f);
            const uint32_t u = S.Nx[j];
            const uint32_t v = G.adjL[G.offL[u] + f - S.pre[j]];
            if (v != x) atomicOr(&trow[(size_t)cpos[v] * W + (j >> 6)], 1ull << (j & 63));
        }
    }
    sync_mem();

    // ---- root node: L' = N(x); a full Q row makes it non-maximal, full P
    // rows are absorbed
    bool nonmax = false;
    uint32_t ab = 0;
    u64 abh = 0;
    for (uint32_t c0 = 0; c0 < nc && !nonmax; c0 += 64) {
        const uint32_t k = c0 + lane();
        bool full = false;
        if (k < nc) {
            uint32_t c = 0;
            for (uint32_t w = 0; w < W; w++) c += __popcll(trow[(size_t)k * W + w]);
            full = c == S.d;
        }
        if (__ballot(full && k >= kx)) nonmax = true;
        if (full && k < kx) {
            ab++;
            abh += G.hR[S.cand[k]];
        }
    }
    if (nonmax) {
        nodes++;
        return 0;
    }
    ab = red_add32(ab);
    abh = red_add64(abh);
    const u64 sr = G.hR[x] + abh, nr = 1 + ab;
    const bool descend = kx != ab;  // P' nonempty
    uint32_t nq = nc - kx;
    if (descend) {
        // reduce Q (its scratch goes where the problem will be laid out)
        if (nq >= kReduceMinQ &&
            off + (nq + 63) / 64 + reduce_scratch_words(S.d, nq) + tail_words <= kp.arena_words) {
            u64* keep = arena + off;
            for (uint32_t w = lane(); w < (nq + 63) / 64; w += 64) keep[w] = 0;
            sync_mem();
            nq = reduce_q(trow + (size_t)kx * W, W, S.d, nullptr, 0, nq, keep,
                          keep + (nq + 63) / 64);
            // kept positions, ascending
            uint32_t at = 0;
            for (uint32_t w = 0; w < ((nc - kx) + 63) / 64; w++) {
                const u64 kw = uni64(keep[w]);
                if ((kw >> lane()) & 1) kidx[at + __popcll(kw & lanemask_lt())] = kx + w * 64 + lane();
                at += __popcll(kw);
            }
        } else {
            for (uint32_t i = lane(); i < nq; i += 64) kidx[i] = kx + i;
        }
        sync_mem();
        S.nc = kx + nq;
        if (!layout(kp, S, arena, lds, off, kLdsHdr) || S.end + tail_words > kp.arena_words) return -1;
    }
    // the root biclique (N(x), {x} u absorbed)
    nodes++;
    u64 sl = 0;
    for (uint32_t j = lane(); j < S.d; j += 64) sl += G.hL[S.Nx[j]];
    sl = red_add64(sl);
    record(G, sl, sr, S.d, nr, cnt, hs);
    if (kp.emit_pos) {
        // the L' terms are the hash terms of N(x) in order
        u64* hl = arena + off;
        if (descend) hl = S.hl;
        for (uint32_t j = lane(); j < S.d; j += 64) hl[j] = G.hL[S.Nx[j]];
        sync_mem();
        emit_biclique(kp.emit_pos, kp.emit, kp.emit_cap, hl, nullptr, S.d, nr, sr);
    }
    if (!descend) return 0;

    // the problem: hash terms, rows (P then kept Q), columns, frame 1
    for (uint32_t j = lane(); j < S.d; j += 64) S.hl[j] = G.hL[S.Nx[j]];
    for (uint32_t k = lane(); k < S.nc; k += 64) {
        const uint32_t o = k < kx ? k : kidx[k - kx];
        S.hc[k] = G.hR[S.cand[o]];
        for (uint32_t w = 0; w < W; w++) S.rows[(size_t)k * W + w] = trow[(size_t)o * W + w];
    }
    sync_mem();
    build_cols(S);
    u64* F1 = S.frame(1);
    u64* L1 = F1 + kFrameHdr;
    u64* P1 = L1 + S.W;
    u64* Q1 = P1 + S.Wc;
    u64* X1 = Q1 + S.Wc;
    fill_full_l(S, L1);
    for (uint32_t c0 = 0; c0 < S.nc; c0 += 64) {
        const uint32_t k = c0 + lane();
        bool full = false;
        if (k < kx) {
            uint32_t c = 0;
            for (uint32_t w = 0; w < W; w++) c += __popcll(S.rows[(size_t)k * W + w]);
            full = c == S.d;
        }
        const u64 bp = __ballot(k < kx && !full), bq = __ballot(k >= kx && k < S.nc);
        if (lane() == 0) {
            P1[c0 >> 6] = bp;
            Q1[c0 >> 6] = bq;
            X1[c0 >> 6] = 0;
        }
    }
    if (lane() == 0) {
        F1[1] = sr;
        F1[2] = nr;
        F1[3] = pq(kx - ab, S.nc - kx);
    }
    wsync();
    return 1;
}

// ---------------------------------------------------------------- serving
constexpr uint64_t kPollTicks = 200;   // mailbox polls at most every 2 us (100 MHz clock)
constexpr uint32_t kLoadBatch = 64;    // records one loader turn takes

// Wave w prefers the large claim lane when w % kLargeShare == 0, so large
// requests keep a bounded share of the waves while small ones go first.
constexpr uint32_t kLargeShare = 4;

// Claim a root (lane 0 decides). Each claim lane lists its graphs in
// arrival order; a wave takes the next root of the oldest graph of its
// preferred lane with roots left, else of the other lane: one fetch-and-add
// on that graph's root cursor, which never has to be retried under
// contention. A cursor that passes the graph's root count marks the graph
// fully claimed, and the lane's first entry moves past it. The pending count
// is raised only while a graph is open: idle waves poll this every
// iteration, and a count raised on every poll would almost never read zero,
// so the exit test would never pass.
__device__ bool serve_claim(const KParams& kp, uint32_t wid, uint32_t& g, uint32_t& x) {
    u64 got = 0, gg = 0, xx = 0;
    if (lane() == 0) {
        ServeCtl* sc = kp.serve;
        unsigned long long* pending = kp.dq + 3;
        const uint32_t pref = wid % kLargeShare == 0 ? 1 : 0;
        for (uint32_t i = 0; i < 2 && !got; i++) {
            const uint32_t ln = pref ^ i;
            for (;;) {
                // the loader published each graph's descriptor, cursor and
                // lane entry before the lane's count (release)
                const u64 la = __hip_atomic_load(&sc->lane_avail[ln], __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT);
                const u64 f = ld_rlx(&sc->lane_first[ln]);
                if (f >= la) break;
                const uint32_t gl = kp.lane_list[(size_t)ln * kp.mail_cap + f];
                const u64 n = kp.graphs[gl].n_roots;
                atomicAdd(pending, 1ull);
                const u64 c = atomicAdd(&kp.root_cursor[gl], 1ull);
                if (c < n) {
                    got = 1;
                    gg = gl;
                    xx = c;
                    break;
                }
                atomicAdd(pending, ~0ull);
                atomicMax(&sc->lane_first[ln], f + 1);
            }
        }
    }
    g = (uint32_t)uni64(gg);
    x = (uint32_t)uni64(xx);
    got = uni64(got);
    if (got) fence_acq();  // the loader published the graph's arrays with a release
    return got != 0;
}

// Loader: one wave at a time (lock), at most every kPollTicks, takes newly
// published mailbox records: copies each packed CSR from pinned staging into
// the device data buffer, appends its descriptor to the graph table and its
// roots to the task space, then publishes the new counts (release) and
// reports the records taken to the host. Returns true if it loaded any.
__device__ bool serve_load(const KParams& kp) {
    ServeCtl* sc = kp.serve;
    u64 pub = 0, taken = 0;
    if (lane() == 0) {
        const u64 now = (u64)wall_clock64();
        if (now >= ld_rlx(&sc->next_poll) && atomicCAS(&sc->lock, 0ull, 1ull) == 0ull) {
            fence_acq();  // the previous loader's counters
            st_rlx(&sc->next_poll, now + kPollTicks);
            // the host's flags, mirrored in device memory for the idle waves:
            // close first, then the final published count, then abort
            const u64 cl =
                __hip_atomic_load(kp.close, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            pub = __hip_atomic_load(kp.mail_pub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            const u64 ab = __hip_atomic_load(kp.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            st_rlx(&sc->pub_seen, pub);
            if (ab) st_rlx(&sc->abort_seen, 1ull);
            if (cl) st_rel(&sc->close_seen, 1ull);
            taken = ld_rlx(&sc->taken);
            if (pub <= taken) {
                st_rlx(&sc->lock, 0ull);
                pub = 0;
            }
        }
    }
    pub = uni64(pub);
    if (!pub) return false;
    taken = uni64(taken);
    const u64 end = min(pub, taken + kLoadBatch);
    u64 g = uni64(lane() == 0 ? ld_rlx(&sc->graphs_avail) : 0);
    u64 tb = uni64(lane() == 0 ? ld_rlx(&sc->tasks_avail) : 0);
    u64 la0 = uni64(lane() == 0 ? ld_rlx(&sc->lane_avail[0]) : 0);
    u64 la1 = uni64(lane() == 0 ? ld_rlx(&sc->lane_avail[1]) : 0);
    for (u64 i = taken; i < end; i++, g++) {
        const MailRec* rec = kp.mail + (i % kp.mail_cap);
        const u64 src = uni64(rec->src), dst = uni64(rec->dst), nw = uni64(rec->words);
        const u64* from = (const u64*)kp.staging + src;
        u64* to = (u64*)kp.data + dst;
        for (u64 w = lane(); w < nw; w += 256) {  // four host reads in flight per lane
            const u64 a = from[w];
            const u64 b = w + 64 < nw ? from[w + 64] : 0;
            const u64 c = w + 128 < nw ? from[w + 128] : 0;
            const u64 d = w + 192 < nw ? from[w + 192] : 0;
            to[w] = a;
            if (w + 64 < nw) to[w + 64] = b;
            if (w + 128 < nw) to[w + 128] = c;
            if (w + 192 < nw) to[w + 192] = d;
        }
        if (lane() == 0) {
            DevGraph G;
            G.offR = (const uint32_t*)(to + rec->off[0]);
            G.adjR = (const uint32_t*)(to + rec->off[1]);
            G.offL = (const uint32_t*)(to + rec->off[2]);
            G.adjL = (const uint32_t*)(to + rec->off[3]);
            G.hR = (const uint64_t*)(to + rec->off[4]);
            G.hL = (const uint64_t*)(to + rec->off[5]);
            G.nR = rec->nR;
            G.nL = rec->nL;
            G.n_roots = rec->n_roots;
            G.r_is_u = rec->r_is_u;
            G.task_base = tb;
            ((DevGraph*)kp.graphs)[g] = G;
            kp.remaining[2 * g] = G.n_roots;
            kp.root_cursor[g] = 0;
            if (rec->lane)
                kp.lane_list[kp.mail_cap + la1++] = (uint32_t)g;
            else
                kp.lane_list[la0++] = (uint32_t)g;
            kp.remaining[2 * g + 1] = 0;
            kp.results[2 * g] = 0;
            kp.results[2 * g + 1] = 0;
            if (G.n_roots == 0) {  // nothing to enumerate: complete now
                unsigned long long* d = kp.done + 4 * g;
                __hip_atomic_store(d, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(d + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(d + 3, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(d + 2, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            tb += G.n_roots;
        }
        tb = uni64(tb);
        la0 = uni64(la0);
        la1 = uni64(la1);
    }
    wsync();
    if (lane() == 0) {
        __threadfence();
        st_rel(&sc->graphs_avail, g);
        st_rel(&sc->tasks_avail, tb);
        st_rel(&sc->lane_avail[0], la0);
        st_rel(&sc->lane_avail[1], la1);
        st_rel(&sc->taken, end);
        __hip_atomic_store(kp.mail_taken, end, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        st_rel(&sc->lock, 0ull);
    }
    wsync();
    return true;
}

// An idle serving wave leaves when nothing is pending and either the
// watchdog fired or the host closed the mailbox and every published record
// has been loaded and every task claimed.
__device__ bool serve_exit(const KParams& kp, u64 pend, bool stop) {
    if (pend != 0) return false;
    u64 ex = 0;
    if (lane() == 0) {
        ServeCtl* sc = kp.serve;
        if (stop) {
            ex = 1;
        } else if (ld_rlx(&sc->close_seen)) {  // device mirror of the host's close
            fence_acq();
            // the loader counts a record as taken after appending it to its
            // lane (release)
            const u64 tk = __hip_atomic_load(&sc->taken, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            ex = tk >= ld_rlx(&sc->pub_seen) &&
                 ld_rlx(&sc->lane_first[0]) >= ld_rlx(&sc->lane_avail[0]) &&
                 ld_rlx(&sc->lane_first[1]) >= ld_rlx(&sc->lane_avail[1]);
        }
    }
    return uni64(ex) != 0;
}

// Per-graph completion (serving): one task of graph g -- a root or a stolen
// subtree -- has finished after adding its results. The wave that closes the
// graph's last open task copies the graph's count, hash and error bits to
// its host-mapped completion record and then raises the record's flag.
// Donations add open tasks before they are published (try_donate), and the
// donor's own task stays open meanwhile, so the count cannot reach zero
// early. Roots deferred for a larger arena stay open until their pass.
__device__ void finish_task(const KParams& kp, uint32_t g, bool failed) {
    if (!kp.remaining || lane() != 0) return;
    unsigned long long* rem = kp.remaining + 2 * (size_t)g;
    if (failed) atomicOr(rem + 1, 1ull);
    // The task's results were added with device-scope atomics, which are
    // performed at the memory side in issue order once acknowledged: wait
    // for them, then decrement (relaxed). The wave that takes the count to
    // zero reads the sums with device-scope loads. No L2 write-back per task.
    __builtin_amdgcn_s_waitcnt(0);
    const u64 left = atomicAdd(rem, ~0ull);
    if (left != 1 || !kp.done) return;
    unsigned long long* d = kp.done + 4 * (size_t)g;
    const u64 c = __hip_atomic_load(&kp.results[2 * g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const u64 h = __hip_atomic_load(&kp.results[2 * g + 1], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    const u64 e = __hip_atomic_load(rem + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(d + 1, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(d + 3, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(d + 2, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int V>
__global__ __launch_bounds__(64) void mbe_persistent(KParams kp) {
    extern __shared__ u64 lds_mem[];
    const uint32_t wid = blockIdx.x;
    u64* arena = (u64*)kp.arena + (size_t)wid * kp.arena_words;
    u64* mark = (u64*)kp.mark + (size_t)wid * kp.mark_words;
    uint32_t* cpos = kp.cpos + (size_t)wid * kp.cpos_len;
    uint32_t* premin = (uint32_t*)(lds_mem + kLdsPreMin);
    for (uint32_t i = lane(); i < kPreMinLevels; i += 64) premin[i] = 1;  // src/main.cu:86
    u64 nodes = 0, roots = 0, build = 0, ovf = 0, donated = 0, stolen = 0, dropped = 0, words = 0;
    bool roots_left = true, idle = false;
    uint32_t backoff = 1;
    unsigned long long* pending = kp.dq + 3;
    Sec sec;
    sec.init(kp.sections != nullptr, lds_mem);
    for (;;) {
        wsync();
        // served kernels poll the device mirror of the host's abort flag
        const bool stop =
            kp.serve ? uni64(lane() == 0 ? ld_rlx(&kp.serve->abort_seen) : 0) != 0 : aborted(kp);
        if (stop) roots_left = false;
        if (roots_left) {
            u64 t = 0;
            uint32_t g = 0, x = 0;
            bool got;
            if (kp.serve) {
                got = serve_claim(kp, wid, g, x);
            } else {
                if (lane() == 0) {
                    atomicAdd(pending, 1ull);
                    t = atomicAdd(kp.cursor, 1ull);
                    if (t >= kp.n_tasks) atomicAdd(pending, ~0ull);
                }
                t = uni64(t);
                got = t < kp.n_tasks;
            }
            if (got) {
                if (idle && lane() == 0) atomicAdd(kp.dq + 2, ~0ull);
                idle = false;
                backoff = 1;
                uint64_t task;
                if (!kp.serve) {
                    task = kp.task_list ? kp.task_list[t]
                                        : kp.task_first + t * (kp.task_step ? kp.task_step : 1);
                    g = find_graph(kp, kp.n_graphs, task);
                }
                const DevGraph G = kp.graphs[g];
                if (kp.serve)
                    task = G.task_base + x;
                else
                    x = (uint32_t)(task - G.task_base);
                u64 cnt = 0, hs = 0;
                Local S;
                sec.mark(kSecOther);
                const int st = build_root<V>(kp, G, x, arena, lds_mem, mark, cpos, S, cnt, hs,
                                             nodes, build);
                sec.mark(kSecBuild);
                if (st > 0 && S.sparse)
                    dfs_sparse<V>(kp, G, g, S, arena, lds_mem, premin, cnt, hs, nodes, donated, words,
                                  sec);
                else if (st > 0)
                    dfs<V, false>(kp, G, g, S, arena, lds_mem, premin, cnt, hs, nodes, donated, words,
                                  sec);
                if (st < 0) {
                    ovf++;
                    if (lane() == 0) {
                        u64 idx = atomicAdd(kp.overflow_count, 1ull);
                        if (idx < kp.overflow_cap) kp.overflow_list[idx] = task;
                    }
                    // served requests have no later pass: report the failure
                    if (kp.serve) finish_task(kp, g, true);
                } else {
                    roots++;
                    if (lane() == 0 && cnt) {
                        atomicAdd(&kp.results[2 * g], cnt);
                        atomicAdd(&kp.results[2 * g + 1], hs);
                    }
                    if (kp.remaining) finish_task(kp, g, aborted(kp));
                }
                if (lane() == 0) {
                    atomicAdd(pending, ~0ull);
                    if (kp.progress) {
                        __hip_atomic_store(&kp.progress[2 * wid], roots + ovf, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                        __hip_atomic_store(&kp.progress[2 * wid + 1], nodes, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                    }
                }
                continue;
            }
            if (!kp.serve) roots_left = false;  // served roots keep arriving
        }
        if (kp.serve && !stop && serve_load(kp)) {
            backoff = 1;
            continue;
        }
        if (kp.n_slots) {
            const int64_t pos = dq_claim_pop(kp);
            if (pos >= 0) {
                if (idle && lane() == 0) atomicAdd(kp.dq + 2, ~0ull);
                idle = false;
                backoff = 1;
                const u64* slot =
                    (const u64*)kp.slots + (size_t)((u64)pos % kp.n_slots) * kp.slot_words;
                const uint32_t g = (uint32_t)slot[0];
                Local S;
                sec.mark(kSecWait);
                const bool ok = !stop && slot[1] != 0 && load_snapshot(kp, S, arena, lds_mem, slot);
                dq_release(kp, (u64)pos);
                sec.mark(kSecFetch);
                if (stop) {
                    // watchdog: discard the subtree
                    finish_task(kp, g, true);
                } else if (ok) {
                    const DevGraph G = kp.graphs[g];
                    u64 cnt = 0, hs = 0;
                    stolen++;
                    dfs<V, false>(kp, G, g, S, arena, lds_mem, premin, cnt, hs, nodes, donated, words,
                                  sec);
                    if (lane() == 0 && cnt) {
                        atomicAdd(&kp.results[2 * g], cnt);
                        atomicAdd(&kp.results[2 * g + 1], hs);
                    }
                    if (kp.remaining) finish_task(kp, g, aborted(kp));
                } else {
                    dropped++;
                    finish_task(kp, g, true);
                }
                if (lane() == 0) {
                    atomicAdd(pending, ~0ull);
                    if (kp.progress)
                        __hip_atomic_store(&kp.progress[2 * wid + 1], nodes, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                }
                continue;
            }
        }
        if (!idle) {
            if (lane() == 0) atomicAdd(kp.dq + 2, 1ull);
            idle = true;
        }
        u64 pend = 0;
        if (lane() == 0) pend = ld_rlx(pending);
        pend = uni64(pend);
        if (kp.serve ? serve_exit(kp, pend, stop) : pend == 0) break;
        // exponential backoff keeps idle waves off the deque's cache lines
        for (uint32_t i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(32);
        backoff = min(backoff * 2, kp.serve ? 64u : 32u);
        sec.mark(kSecWait);
    }
    sec.mark(kSecWait);
    if (idle && lane() == 0) atomicAdd(kp.dq + 2, ~0ull);
    if (lane() == 0) {
        atomicAdd(&kp.stats[kStatNodes], nodes);
        atomicAdd(&kp.stats[kStatRoots], roots);
        atomicAdd(&kp.stats[kStatDonated], donated);
        atomicAdd(&kp.stats[kStatStolen], stolen);
        atomicAdd(&kp.stats[kStatOverflow], ovf);
        atomicAdd(&kp.stats[kStatBuildWords], build);
        atomicAdd(&kp.stats[kStatDropped], dropped);
        atomicAdd(&kp.stats[kStatWords], words);
        if (kp.sections)
            for (int i = 0; i < kSecCount; i++) atomicAdd(&kp.sections[i], sec.t[i]);
        if (kp.worker_busy) {  // per-worker busy ticks (all sections but the wait)
            u64 busy = 0;
            for (int i = 0; i < kSecCount; i++)
                if (i != kSecWait) busy += sec.t[i];
            kp.worker_busy[wid] = busy;
        }
    }
}

// Fill n words with v (clears control words inside captured graphs, where a
// memset node is not used).
__global__ void mbe_fill(unsigned long long* p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Store one graph descriptor (a kernel argument) into device memory.
__global__ void mbe_put_graph(DevGraph* dst, DevGraph g) {
    if (threadIdx.x == 0) *dst = g;
}

// Reset the deque control block: head, tail, idle, pending, slot sequences.
__global__ void mbe_dq_init(unsigned long long* dq, uint32_t n_slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4) dq[i] = 0;
    if (i < n_slots) dq[4 + i] = i;
}

}  // namespace

#ifndef MBE_EMULATE
// The allocation-free solve's scratch (mbe_solve_noalloc). Host code never
// uses this name as an address: it takes the device address once per device
// from noalloc_scratch_address() and passes that to the kernels.
static __device__ __align__(256) unsigned long long scratch[kNoAllocWords];

hipError_t noalloc_scratch_address(void** ptr) {
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

hipError_t launch_put_graph(DevGraph* dst, const DevGraph& g, hipStream_t stream) {
    hipLaunchKernelGGL(mbe_put_graph, dim3(1), dim3(64), 0, stream, dst, g);
    return hipGetLastError();
}

// Host-side launcher: one persistent launch of `workers` single-wave
// workgroups on `stream`.
hipError_t launch_mbe(const KParams& kp, uint32_t workers, hipStream_t stream) {
    const size_t lds_bytes = (size_t)kp.lds_words * sizeof(u64);
    switch (kp.variant) {
        case kNoRS:
            hipLaunchKernelGGL(mbe_persistent<kNoRS>, dim3(workers), dim3(64), lds_bytes, stream, kp);
            break;
        case kNoES:
            hipLaunchKernelGGL(mbe_persistent<kNoES>, dim3(workers), dim3(64), lds_bytes, stream, kp);
            break;
        case kNoWS:
            hipLaunchKernelGGL(mbe_persistent<kNoWS>, dim3(workers), dim3(64), lds_bytes, stream, kp);
            break;
        default:
            hipLaunchKernelGGL(mbe_persistent<kCuMBE>, dim3(workers), dim3(64), lds_bytes, stream, kp);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_dq_init(unsigned long long* dq, uint32_t n_slots, hipStream_t stream) {
    const uint32_t n = n_slots > 4 ? n_slots : 4;
    hipLaunchKernelGGL(mbe_dq_init, dim3((n + 255) / 256), dim3(256), 0, stream, dq, n_slots);
    return hipGetLastError();
}

hipError_t launch_fill(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(mbe_fill, dim3((uint32_t)(blocks < 1024 ? blocks : 1024)), dim3(256), 0,
                       stream, p, n, v);
    return hipGetLastError();
}

hipError_t set_lds_limit(uint32_t lds_words) {
    const int bytes = (int)(lds_words * sizeof(u64));
    hipError_t e = hipFuncSetAttribute((const void*)mbe_persistent<kCuMBE>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)mbe_persistent<kNoRS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)mbe_persistent<kNoES>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)mbe_persistent<kNoWS>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}
#endif  // MBE_EMULATE

}  // namespace mbe
