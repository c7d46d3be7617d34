// gfx950 kernels of the MBE engine.
//
// Execution model: a persistent grid of single-wave workgroups; each wave64
// is one worker (the reference's "block", src/mbe_noWS.cuh:101-615). A
// worker claims root tasks from a global cursor (the reference's P_ptr,
// src/mbe_noWS.cuh:123-133), localises the root's subproblem into bitmaps and
// walks its set-enumeration subtree with an explicit stack of frames that live
// in LDS and spill to the worker's HBM arena.
//
// Local subproblem of root x (rank i): the universe of L is N(x) (d bits,
// W = ceil(d/64) words); the candidates are the R vertices at distance two
// (nc of them, Wc = ceil(nc/64) words), numbered in rank order, so the
// candidates below position kx are the root's P (ranks < i) and the rest its
// Q (ranks > i) -- the root-level sets of src/mbe_noWS.cuh:135-154.
//   rows[k]  = N(cand k)  n N(x)   as a d-bit set      (forward view)
//   cols[j]  = N(u_j) n cands      as an nc-bit set    (reverse view)
// A frame at level k holds L_k (W words), P_k and Q_k (Wc words each) plus
// x, the running R hash sum and |R|.
//
// Per node (src/mbe_cuMBE.cuh:345-553):
//   candidate selection  x = argmin_{v in P} |row(v) & L|, reverse scan with
//                        early stop at the previous minimum (ES);
//   L' = row(x) & L;
//   maximality + expansion, reverse scan (RS): A = AND_{j in L'} cols[j],
//                        O = OR_{j in L'} cols[j]; non-maximal iff A & Q;
//                        absorbed = A & P; P' = (O & P) \ A; Q' = O & Q;
//   noRS: the same sets from forward row tests over P and Q.
// Q' drops candidates with no neighbour in L' (they can never cover a
// descendant's L), which removes work but no tree node.
#ifndef MBE_EMULATE
#include <hip/hip_runtime.h>
#endif

#include "engine.h"
#include "hash.h"

namespace mbe {

using u64 = unsigned long long;

namespace {

constexpr uint32_t kInf = 1u << 30;              // src/header.cuh:8
constexpr uint32_t kPreMinLevels = 4096;         // per-level ES memory
constexpr uint32_t kPreMinWords = kPreMinLevels / 2;
constexpr uint32_t kFrameHdr = 4;                // x, SR, |R|, spare

enum Variant : int { kCuMBE = 0, kNoRS = 1, kNoES = 2, kNoWS = 3 };

__device__ __forceinline__ uint32_t lane() { return threadIdx.x; }

__device__ __forceinline__ u64 wsum64(u64 v) {
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ uint32_t wsum32(uint32_t v) {
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ uint32_t wmin32(uint32_t v) {
    for (int o = 32; o; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o));
    return v;
}
// exclusive prefix sum across the wave
__device__ __forceinline__ uint32_t wscan_excl(uint32_t v) {
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if ((int)lane() >= o) x += y;
    }
    return x - v;
}
// value known to be equal in all lanes -> scalar register
__device__ __forceinline__ u64 uni64(u64 v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t uni32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t find_graph(const KParams& kp, uint64_t t) {
    uint32_t lo = 0, hi = kp.n_graphs;
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

struct Local {
    uint32_t x, d, W, nc, Wc, kx, fs, nlds, maxdepth;
    const uint32_t* Nx;
    uint32_t* pre;
    u64* hl;
    uint32_t* cand;
    u64* hc;
    u64* rows;
    u64* cols;
    uint32_t* list;
    u64* hstack;
    u64* lds;

    __device__ __forceinline__ u64* frame(uint32_t k) const {
        uint32_t i = k - 1;
        return i < nlds ? lds + (size_t)i * fs : hstack + (size_t)(i - nlds) * fs;
    }
};

// sum of hash terms over the set bits of L (W words), all lanes get it
__device__ __forceinline__ u64 sum_l(const Local& S, const u64* L) {
    u64 s = 0;
    for (uint32_t lw = 0; lw < S.W; lw++) {
        u64 word = uni64(L[lw]);
        if ((word >> lane()) & 1) s += S.hl[lw * 64 + lane()];
    }
    return wsum64(s);
}

__device__ __forceinline__ void record(const DevGraph& G, u64 sl, u64 sr, u64 nl, u64 nr, u64& cnt,
                                       u64& hs) {
    u64 h = G.r_is_u ? biclique_hash(sr, sl, nr, nl) : biclique_hash(sl, sr, nl, nr);
    cnt++;
    hs += h;
}

// Candidate selection at a level >= 1 (src/mbe_cuMBE.cuh:345-390): scan P in
// descending candidate order, keep the first strict minimum of |row(v) & L|;
// with ES stop as soon as the running minimum equals the level's previous
// minimum. Removes the chosen candidate from P and returns it.
template <int V>
__device__ uint32_t select_x(const Local& S, u64* F, uint32_t* premin_lvl) {
    u64* L = F + kFrameHdr;
    u64* P = L + S.W;
    // compact P's set bits into S.list in ascending order
    uint32_t n = 0;
    for (uint32_t w0 = 0; w0 < S.Wc; w0 += 64) {
        uint32_t w = w0 + lane();
        u64 pw = w < S.Wc ? P[w] : 0;
        uint32_t c = __popcll(pw);
        uint32_t idx = n + wscan_excl(c);
        while (pw) {
            uint32_t b = __builtin_ctzll(pw);
            pw &= pw - 1;
            S.list[idx++] = w * 64 + b;
        }
        n += wsum32(c);
    }
    __syncthreads();
    const uint32_t pm = *premin_lvl;
    uint32_t best = kInf, arg = 0;
    bool nostop = (V == kNoES);
    bool done = false;
    for (int top = (int)n - 1; top >= 0 && !done; top -= 64) {
        int i = top - (int)lane();
        bool valid = i >= 0;
        uint32_t v = valid ? S.list[i] : 0;
        uint32_t c = kInf;
        if (valid) {
            c = 0;
            const u64* r = S.rows + (size_t)v * S.W;
            for (uint32_t w = 0; w < S.W; w++) c += __popcll(r[w] & L[w]);
        }
        if (!nostop) {
            u64 m = __ballot(valid && c <= pm);
            if (m) {
                int s = __builtin_ctzll(m);
                uint32_t cs = __shfl(c, s);
                if (cs == pm) {
                    best = cs;
                    arg = __shfl(v, s);
                    done = true;
                    break;
                }
                nostop = true;
            }
        }
        uint32_t cm = wmin32(c);
        if (cm < best) {
            int l = __builtin_ctzll(__ballot(c == cm));
            best = cm;
            arg = __shfl(v, l);
        }
    }
    arg = uni32(arg);
    if (lane() == 0) {
        *premin_lvl = best;
        P[arg >> 6] &= ~(1ull << (arg & 63));
    }
    __syncthreads();
    return arg;
}

struct NodeOut {
    bool maximal;
    bool descend;
    uint32_t nl;    // |L'|
    uint32_t ab;    // absorbed count
    u64 abh;        // absorbed hash sum
};

// Build L' = row(x) & L into the child frame; return |L'|.
__device__ __forceinline__ uint32_t make_lprime(const Local& S, const u64* L, uint32_t x, u64* Lc) {
    uint32_t c = 0;
    const u64* r = S.rows + (size_t)x * S.W;
    for (uint32_t w = lane(); w < S.W; w += 64) {
        u64 lw = r[w] & L[w];
        Lc[w] = lw;
        c += __popcll(lw);
    }
    c = wsum32(c);
    __syncthreads();
    return c;
}

// Reverse-scan maximality check and expansion (src/mbe_noWS.cuh:215-289).
__device__ NodeOut process_rs(const Local& S, u64* F, u64* Fc, uint32_t nl) {
    const u64* P = F + kFrameHdr + S.W;
    const u64* Q = P + S.Wc;
    const u64* Lc = Fc + kFrameHdr;
    u64* Pc = Fc + kFrameHdr + S.W;
    u64* Qc = Pc + S.Wc;
    NodeOut o{true, false, nl, 0, 0};
    bool nonmax = false, anyp = false;
    uint32_t ab = 0;
    u64 abh = 0;
    for (uint32_t w0 = 0; w0 < S.Wc; w0 += 64) {
        uint32_t w = w0 + lane();
        bool valid = w < S.Wc;
        u64 a = ~0ull, ob = 0;
        for (uint32_t lw = 0; lw < S.W; lw++) {
            u64 word = uni64(Lc[lw]);
            while (word) {
                uint32_t j = lw * 64 + __builtin_ctzll(word);
                word &= word - 1;
                if (valid) {
                    u64 c = S.cols[(size_t)j * S.Wc + w];
                    a &= c;
                    ob |= c;
                }
            }
        }
        u64 q = valid ? Q[w] : 0, p = valid ? P[w] : 0;
        nonmax |= (a & q) != 0;
        u64 absd = a & p;
        u64 pn = ob & p & ~a;
        u64 qn = ob & q;
        if (valid) {
            Pc[w] = pn;
            Qc[w] = qn;
        }
        anyp |= pn != 0;
        ab += __popcll(absd);
        while (absd) {
            uint32_t b = __builtin_ctzll(absd);
            absd &= absd - 1;
            abh += S.hc[w * 64 + b];
        }
    }
    o.maximal = __ballot(nonmax) == 0;
    o.descend = o.maximal && __ballot(anyp) != 0;
    o.ab = wsum32(ab);
    o.abh = wsum64(abh);
    __syncthreads();
    return o;
}

// Forward-scan variant (noRS, src/mbe_noRS.cuh:216-287): every Q and P
// candidate's row is tested against L' directly.
__device__ NodeOut process_fwd(const Local& S, u64* F, u64* Fc, uint32_t nl) {
    const u64* P = F + kFrameHdr + S.W;
    const u64* Q = P + S.Wc;
    const u64* Lc = Fc + kFrameHdr;
    u64* Pc = Fc + kFrameHdr + S.W;
    u64* Qc = Pc + S.Wc;
    NodeOut o{true, false, nl, 0, 0};
    bool nonmax = false;
    for (uint32_t w0 = 0; w0 < S.Wc; w0 += 64) {
        uint32_t w = w0 + lane();
        u64 qw = w < S.Wc ? Q[w] : 0, qn = 0;
        while (qw) {
            uint32_t b = __builtin_ctzll(qw);
            qw &= qw - 1;
            const u64* r = S.rows + (size_t)(w * 64 + b) * S.W;
            bool cov = true, touch = false;
            for (uint32_t lw = 0; lw < S.W; lw++) {
                u64 l = Lc[lw], x = r[lw] & l;
                cov &= x == l;
                touch |= x != 0;
            }
            nonmax |= cov;
            if (touch) qn |= 1ull << b;
        }
        if (w < S.Wc) Qc[w] = qn;
    }
    o.maximal = __ballot(nonmax) == 0;
    if (!o.maximal) {
        __syncthreads();
        return o;
    }
    bool anyp = false;
    uint32_t ab = 0;
    u64 abh = 0;
    for (uint32_t w0 = 0; w0 < S.Wc; w0 += 64) {
        uint32_t w = w0 + lane();
        u64 pw = w < S.Wc ? P[w] : 0, pn = 0;
        while (pw) {
            uint32_t b = __builtin_ctzll(pw);
            pw &= pw - 1;
            uint32_t k = w * 64 + b;
            const u64* r = S.rows + (size_t)k * S.W;
            uint32_t c = 0;
            for (uint32_t lw = 0; lw < S.W; lw++) c += __popcll(r[lw] & Lc[lw]);
            if (c == nl) {
                ab++;
                abh += S.hc[k];
            } else if (c) {
                pn |= 1ull << b;
            }
        }
        if (w < S.Wc) Pc[w] = pn;
        anyp |= pn != 0;
    }
    o.descend = __ballot(anyp) != 0;
    o.ab = wsum32(ab);
    o.abh = wsum64(abh);
    __syncthreads();
    return o;
}

// Localise root x and enumerate its subtree. Returns false when the local
// problem does not fit the arena (the root is then deferred).
template <int V>
__device__ bool solve_root(const KParams& kp, const DevGraph& G, uint32_t x, u64* arena, u64* lds,
                           u64* mark, uint32_t* cpos, uint32_t* premin, u64& cnt, u64& hs,
                           u64& nodes, u64& build) {
    Local S;
    S.x = x;
    const uint32_t s0 = G.offR[x];
    S.d = G.offR[x + 1] - s0;
    S.Nx = G.adjR + s0;
    S.W = (S.d + 63) >> 6;
    S.lds = lds;
    if (S.d == 0) return true;
    size_t off = kPreMinWords;
    S.pre = (uint32_t*)(arena + off);
    off += (S.d + 2) / 2;
    S.hl = arena + off;
    off += S.d;
    if (off > kp.arena_words) return false;

    // degree prefix over N(x) and the L hash terms
    uint32_t carry = 0;
    for (uint32_t j0 = 0; j0 < S.d; j0 += 64) {
        uint32_t j = j0 + lane(), dg = 0;
        if (j < S.d) {
            uint32_t u = S.Nx[j];
            dg = G.offL[u + 1] - G.offL[u];
            S.hl[j] = G.hL[u];
        }
        uint32_t ex = wscan_excl(dg);
        if (j < S.d) S.pre[j] = carry + ex;
        carry += wsum32(dg);
    }
    if (lane() == 0) S.pre[S.d] = carry;
    const uint32_t T = carry;
    build += T;
    __syncthreads();

    // discover the distance-two candidates (bitmap over R ranks)
    for (uint32_t f0 = 0; f0 < T; f0 += 64) {
        uint32_t f = f0 + lane();
        if (f < T) {
            uint32_t j = find_j(S.pre, S.d, f);
            uint32_t u = S.Nx[j];
            uint32_t v = G.adjL[G.offL[u] + f - S.pre[j]];
            if (v != x) atomicOr(&mark[v >> 6], 1ull << (v & 63));
        }
    }
    __threadfence();
    __syncthreads();
    const uint32_t nwords = (G.nR + 63) >> 6;
    uint32_t ncl = 0, kxl = 0;
    for (uint32_t w = lane(); w < nwords; w += 64) {
        u64 m = mark[w];
        ncl += __popcll(m);
        if (w < (x >> 6)) kxl += __popcll(m);
        else if (w == (x >> 6)) kxl += __popcll(m & ((1ull << (x & 63)) - 1));
    }
    S.nc = wsum32(ncl);
    S.kx = wsum32(kxl);
    S.Wc = (S.nc + 63) >> 6;
    S.maxdepth = min(S.d, S.nc) + 2;
    S.fs = kFrameHdr + S.W + 2 * S.Wc;
    S.nlds = kp.lds_words / S.fs;
    size_t o_cand = off;
    off += (S.nc + 1) / 2;
    size_t o_hc = off;
    off += S.nc;
    size_t o_rows = off;
    off += (size_t)S.nc * S.W;
    size_t o_cols = off;
    off += (size_t)S.d * S.Wc;
    size_t o_list = off;
    off += (S.nc + 1) / 2 + 1;
    size_t o_stack = off;
    if (S.maxdepth > S.nlds) off += (size_t)(S.maxdepth - S.nlds) * S.fs;
    if (off > kp.arena_words) {
        for (uint32_t w = lane(); w < nwords; w += 64) mark[w] = 0;
        __threadfence();
        __syncthreads();
        return false;
    }
    S.cand = (uint32_t*)(arena + o_cand);
    S.hc = arena + o_hc;
    S.rows = arena + o_rows;
    S.cols = arena + o_cols;
    S.list = (uint32_t*)(arena + o_list);
    S.hstack = arena + o_stack;

    // assign local indices in rank order; clear the bitmap
    uint32_t k0 = 0;
    for (uint32_t w0 = 0; w0 < nwords; w0 += 64) {
        uint32_t w = w0 + lane();
        u64 m = w < nwords ? mark[w] : 0;
        uint32_t c = __popcll(m);
        uint32_t k = k0 + wscan_excl(c);
        if (m) mark[w] = 0;
        while (m) {
            uint32_t b = __builtin_ctzll(m);
            m &= m - 1;
            uint32_t v = w * 64 + b;
            S.cand[k] = v;
            cpos[v] = k;
            S.hc[k] = G.hR[v];
            k++;
        }
        k0 += wsum32(c);
    }
    const size_t nz = (size_t)S.nc * S.W + (size_t)S.d * S.Wc;
    for (size_t i = lane(); i < nz; i += 64) S.rows[i] = 0;
    __threadfence();
    __syncthreads();
    // rows and cols of the local bit matrix
    for (uint32_t f0 = 0; f0 < T; f0 += 64) {
        uint32_t f = f0 + lane();
        if (f < T) {
            uint32_t j = find_j(S.pre, S.d, f);
            uint32_t u = S.Nx[j];
            uint32_t v = G.adjL[G.offL[u] + f - S.pre[j]];
            if (v != x) {
                uint32_t k = cpos[v];
                atomicOr(&S.rows[(size_t)k * S.W + (j >> 6)], 1ull << (j & 63));
                atomicOr(&S.cols[(size_t)j * S.Wc + (k >> 6)], 1ull << (k & 63));
            }
        }
    }
    __threadfence();
    __syncthreads();

    // ---- root node: L' = N(x), P = cands below kx, Q = the rest
    u64* F1 = S.frame(1);
    u64* L1 = F1 + kFrameHdr;
    u64* P1 = L1 + S.W;
    u64* Q1 = P1 + S.Wc;
    for (uint32_t w = lane(); w < S.W; w += 64)
        L1[w] = (w == S.W - 1 && (S.d & 63)) ? ((1ull << (S.d & 63)) - 1) : ~0ull;
    bool nonmax = false;
    uint32_t ab = 0;
    u64 abh = 0;
    for (uint32_t c0 = 0; c0 < S.nc && !nonmax; c0 += 64) {
        uint32_t k = c0 + lane();
        bool valid = k < S.nc, full = false;
        if (valid) {
            uint32_t pc = 0;
            const u64* r = S.rows + (size_t)k * S.W;
            for (uint32_t w = 0; w < S.W; w++) pc += __popcll(r[w]);
            full = pc == S.d;
        }
        bool isP = valid && k < S.kx, isQ = valid && k >= S.kx;
        if (__ballot(full && isQ)) nonmax = true;
        u64 bp = __ballot(isP && !full), bq = __ballot(isQ);
        if (lane() == 0) {
            P1[c0 >> 6] = bp;
            Q1[c0 >> 6] = bq;
        }
        if (full && isP) {
            ab++;
            abh += S.hc[k];
        }
    }
    nodes++;
    if (nonmax) return true;
    ab = wsum32(ab);
    abh = wsum64(abh);
    {
        u64 sl = 0;
        for (uint32_t j = lane(); j < S.d; j += 64) sl += S.hl[j];
        sl = wsum64(sl);
        u64 sr = G.hR[x] + abh, nr = 1 + ab;
        record(G, sl, sr, S.d, nr, cnt, hs);
        if (S.kx == ab) return true;   // P' empty
        if (lane() == 0) {
            F1[1] = sr;
            F1[2] = nr;
        }
    }
    __syncthreads();

    // ---- iterative DFS over levels >= 1 (src/mbe_noWS.cuh:337-615)
    uint32_t k = 1;
    while (k >= 1) {
        u64* F = S.frame(k);
        u64* L = F + kFrameHdr;
        u64* P = L + S.W;
        u64* Q = P + S.Wc;
        bool anyP = false;
        for (uint32_t w = lane(); w < S.Wc; w += 64) anyP |= P[w] != 0;
        if (__ballot(anyP) == 0) {
            // level exhausted: ascend, Q[k-1] <- Q[k-1] u {x[k-1]}
            k--;
            if (k >= 1) {
                u64* Fp = S.frame(k);
                uint32_t xp = (uint32_t)Fp[0];
                u64* Qp = Fp + kFrameHdr + S.W + S.Wc;
                if (lane() == 0) Qp[xp >> 6] |= 1ull << (xp & 63);
                __syncthreads();
            }
            continue;
        }
        uint32_t* pm = premin + min(k, kPreMinLevels - 1);
        uint32_t xk = select_x<V>(S, F, pm);
        nodes++;
        if (lane() == 0) F[0] = xk;
        u64* Fc = S.frame(k + 1);
        uint32_t nl = make_lprime(S, L, xk, Fc + kFrameHdr);
        NodeOut o = (V == kNoRS) ? process_fwd(S, F, Fc, nl) : process_rs(S, F, Fc, nl);
        if (o.maximal) {
            u64 sl = sum_l(S, Fc + kFrameHdr);
            u64 sr = F[1] + S.hc[xk] + o.abh;
            u64 nr = F[2] + 1 + o.ab;
            record(G, sl, sr, nl, nr, cnt, hs);
            if (o.descend) {
                if (lane() == 0) {
                    Fc[1] = sr;
                    Fc[2] = nr;
                }
                __syncthreads();
                k++;
                continue;
            }
        }
        if (lane() == 0) Q[xk >> 6] |= 1ull << (xk & 63);
        __syncthreads();
    }
    return true;
}

template <int V>
__global__ __launch_bounds__(64) void mbe_persistent(KParams kp) {
    extern __shared__ u64 lds_mem[];
    const uint32_t wid = blockIdx.x;
    u64* arena = (u64*)kp.arena + (size_t)wid * kp.arena_words;
    u64* mark = (u64*)kp.mark + (size_t)wid * kp.mark_words;
    uint32_t* cpos = kp.cpos + (size_t)wid * kp.cpos_len;
    uint32_t* premin = (uint32_t*)arena;
    for (uint32_t i = lane(); i < kPreMinLevels; i += 64) premin[i] = 1;  // src/main.cu:86
    u64 nodes = 0, roots = 0, build = 0, ovf = 0;
    for (;;) {
        __syncthreads();
        u64 t = 0;
        if (lane() == 0) t = atomicAdd(kp.cursor, 1ull);
        t = uni64(__shfl(t, 0));
        if (t >= kp.n_tasks) break;
        const uint64_t task = kp.task_list ? kp.task_list[t] : t;
        const uint32_t g = find_graph(kp, task);
        const DevGraph G = kp.graphs[g];
        const uint32_t x = (uint32_t)(task - G.task_base);
        u64 cnt = 0, hs = 0;
        bool ok = solve_root<V>(kp, G, x, arena, lds_mem, mark, cpos, premin, cnt, hs, nodes, build);
        if (!ok) {
            ovf++;
            if (lane() == 0) {
                u64 idx = atomicAdd(kp.overflow_count, 1ull);
                if (idx < kp.overflow_cap) kp.overflow_list[idx] = task;
            }
        } else {
            roots++;
            if (lane() == 0 && cnt) {
                atomicAdd(&kp.results[2 * g], cnt);
                atomicAdd(&kp.results[2 * g + 1], hs);
            }
        }
    }
    if (lane() == 0) {
        atomicAdd(&kp.stats[kStatNodes], nodes);
        atomicAdd(&kp.stats[kStatRoots], roots);
        atomicAdd(&kp.stats[kStatOverflow], ovf);
        atomicAdd(&kp.stats[kStatBuildWords], build);
    }
}

}  // namespace

#ifndef MBE_EMULATE
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

hipError_t set_lds_limit(uint32_t lds_words) {
    const int bytes = (int)(lds_words * sizeof(u64));
    hipError_t e = hipSuccess;
    e = hipFuncSetAttribute((const void*)mbe_persistent<kCuMBE>,
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
