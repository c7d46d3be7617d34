/*
 * CPU oracle for maximal biclique enumeration -- TEST INFRASTRUCTURE ONLY.
 * Serial restatement of the reference algorithm; see mbe_oracle.h for the
 * line-by-line map to the reference kernels. Never linked into the product.
 *
 * Sets are prefixes of permutation arrays with inverse position maps, as in
 * the reference (src/mbe_noWS.cuh:10-36): u in L_lvl <=> u2L[u] < L_lp[lvl],
 * v in P_lvl <=> v2P[v] < P_lp[lvl], v in Q_lvl <=> v2Q[v] < Q_lp[lvl],
 * R_lvl = R[0 .. R_lp[lvl]). One "block" of the reference becomes one thread.
 */
#define _GNU_SOURCE
#include "mbe_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORACLE_INF 1073741824 /* src/header.cuh:8 */

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

static inline uint64_t combine(uint64_t su, uint64_t sv, uint64_t na, uint64_t nb) {
    return mix64(su ^ rotl64(sv, 29) ^ (na * MBE_HASH_K1 + nb * MBE_HASH_K2));
}

uint64_t mbe_oracle_biclique_hash(const uint32_t* a, uint32_t na, const uint32_t* b,
                                  uint32_t nb) {
    uint64_t su = 0, sv = 0;
    for (uint32_t i = 0; i < na; i++) su += mix64((uint64_t)a[i] ^ MBE_HASH_SALT_U);
    for (uint32_t i = 0; i < nb; i++) sv += mix64((uint64_t)b[i] ^ MBE_HASH_SALT_V);
    return combine(su, sv, na, nb);
}

/* ---------------------------------------------------------------- input -- */

typedef struct {
    uint32_t n;          /* vertices on this side */
    uint32_t* off;       /* n + 1 */
    uint32_t* adj;       /* sorted, deduplicated */
} side_csr;

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

/* Copy the U-side CSR, sorting and deduplicating each list; build the V side
 * by a counting transpose, which leaves every V list sorted ascending
 * (the property src/transpose.cuh:48-52 produces). */
static int build_sides(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                       const uint32_t* adj_u, side_csr* U, side_csr* V) {
    uint64_t m = off_u[n_u];
    U->n = n_u; V->n = n_v;
    U->off = (uint32_t*)calloc((size_t)n_u + 1, sizeof(uint32_t));
    U->adj = (uint32_t*)malloc((m ? m : 1) * sizeof(uint32_t));
    V->off = (uint32_t*)calloc((size_t)n_v + 1, sizeof(uint32_t));
    if (!U->off || !U->adj || !V->off) return -2;
    uint64_t w = 0;
    for (uint32_t u = 0; u < n_u; u++) {
        uint32_t s = off_u[u], e = off_u[u + 1];
        if (e < s) return -1;
        uint64_t base = w;
        for (uint32_t k = s; k < e; k++) {
            if (adj_u[k] >= n_v) return -1;
            U->adj[w++] = adj_u[k];
        }
        qsort(U->adj + base, w - base, sizeof(uint32_t), cmp_u32);
        uint64_t o = base;
        for (uint64_t k = base; k < w; k++)
            if (k == base || U->adj[k] != U->adj[k - 1]) U->adj[o++] = U->adj[k];
        w = o;
        U->off[u + 1] = (uint32_t)w;
    }
    V->adj = (uint32_t*)malloc((w ? w : 1) * sizeof(uint32_t));
    uint32_t* cur = (uint32_t*)calloc((size_t)n_v + 1, sizeof(uint32_t));
    if (!V->adj || !cur) { free(cur); return -2; }
    for (uint64_t k = 0; k < w; k++) V->off[U->adj[k] + 1]++;
    for (uint32_t v = 0; v < n_v; v++) V->off[v + 1] += V->off[v];
    memcpy(cur, V->off, sizeof(uint32_t) * n_v);
    for (uint32_t u = 0; u < n_u; u++)
        for (uint32_t k = U->off[u]; k < U->off[u + 1]; k++) V->adj[cur[U->adj[k]]++] = u;
    free(cur);
    return 0;
}

static void free_side(side_csr* s) { free(s->off); free(s->adj); }

/* -------------------------------------------------------------- problem -- */

typedef struct {
    uint32_t nR, nL;
    const uint32_t *offR, *adjR;   /* enumerated side: lists of L ids */
    const uint32_t *offL, *adjL;   /* other side: lists of R ids */
    uint32_t* ori_P;               /* root order, src/header.cuh:55-63 */
    int64_t P_ptr0;                /* last index with nonzero degree */
    uint64_t *hR, *hL;             /* mix64(file id ^ salt of its side) */
    int r_is_u;                    /* 1 if the enumerated side is U (rows) */
    int rs, es;
    int prune;                     /* twin removal (not the reference's tree) */
    mbe_oracle_emit_fn emit;
    void* user;
    atomic_llong cursor;           /* the reference's P_ptr, src/header.cuh:35 */
    int64_t sub_begin, sub_end, sub_stride;  /* root subset by rank */
} problem;

typedef struct {
    const problem* pb;
    int32_t *u2L, *L, *L_buf;
    int32_t *v2P, *P, *v2Q, *Q, *R, *num_N_u;
    int32_t *x, *L_lp, *R_lp, *P_lp, *Q_lp, *pre_min;
    uint32_t *ea, *eb;             /* emit scratch */
    uint64_t count, hash, work, nodes;
    int err;
} worker;

static inline uint32_t degR(const problem* pb, uint32_t v) { return pb->offR[v + 1] - pb->offR[v]; }

typedef struct { uint32_t v, d; } vd_pair;
static int cmp_root(const void* a, const void* b) {
    const vd_pair *x = (const vd_pair*)a, *y = (const vd_pair*)b;
    if (x->d != y->d) return x->d > y->d ? -1 : 1;     /* degree descending */
    return (x->v > y->v) - (x->v < y->v);              /* ties: id ascending */
}

static int worker_init(worker* w, const problem* pb) {
    memset(w, 0, sizeof(*w));
    w->pb = pb;
    size_t nL = pb->nL, nR = pb->nR, nLev = (size_t)pb->nR + 2;
    size_t nBuf = nL > nR ? nL : nR;
    w->u2L = malloc(sizeof(int32_t) * (nL + 1));
    w->L = malloc(sizeof(int32_t) * (nL + 1));
    w->L_buf = malloc(sizeof(int32_t) * (nBuf + 1));
    w->v2P = malloc(sizeof(int32_t) * (nR + 1));
    w->P = malloc(sizeof(int32_t) * (nR + 1));
    w->v2Q = malloc(sizeof(int32_t) * (nR + 1));
    w->Q = malloc(sizeof(int32_t) * (nR + 1));
    w->R = malloc(sizeof(int32_t) * (nR + 1));
    w->num_N_u = calloc(nR + 1, sizeof(int32_t));
    w->x = malloc(sizeof(int32_t) * nLev);
    w->L_lp = malloc(sizeof(int32_t) * nLev);
    w->R_lp = malloc(sizeof(int32_t) * nLev);
    w->P_lp = malloc(sizeof(int32_t) * nLev);
    w->Q_lp = malloc(sizeof(int32_t) * nLev);
    w->pre_min = malloc(sizeof(int32_t) * nLev);
    w->ea = malloc(sizeof(uint32_t) * (nBuf + 1));
    w->eb = malloc(sizeof(uint32_t) * (nBuf + 1));
    if (!w->u2L || !w->L || !w->L_buf || !w->v2P || !w->P || !w->v2Q || !w->Q || !w->R ||
        !w->num_N_u || !w->x || !w->L_lp || !w->R_lp || !w->P_lp || !w->Q_lp ||
        !w->pre_min || !w->ea || !w->eb)
        return -2;
    /* initial state, src/main.cu:72-86 */
    for (size_t i = 0; i < nL; i++) w->u2L[i] = w->L[i] = (int32_t)i;
    for (size_t i = 0; i < nR; i++) w->v2P[i] = w->P[i] = w->v2Q[i] = w->Q[i] = w->R[i] = (int32_t)i;
    for (size_t i = 0; i < nLev; i++) {
        w->x[i] = -1; w->L_lp[i] = (int32_t)nL; w->R_lp[i] = 0;
        w->P_lp[i] = (int32_t)nR; w->Q_lp[i] = 0; w->pre_min[i] = 1;
    }
    return 0;
}

static void worker_free(worker* w) {
    free(w->u2L); free(w->L); free(w->L_buf); free(w->v2P); free(w->P); free(w->v2Q);
    free(w->Q); free(w->R); free(w->num_N_u); free(w->x); free(w->L_lp); free(w->R_lp);
    free(w->P_lp); free(w->Q_lp); free(w->pre_min); free(w->ea); free(w->eb);
}

/* Q <- Q u {v} at the given prefix length (src/mbe_noWS.cuh:141-147). */
static inline void q_push(worker* w, int32_t* Q_lp, int32_t v) {
    int32_t q = w->v2Q[v];
    int32_t tmp = w->Q[*Q_lp];
    w->Q[*Q_lp] = v;
    w->Q[q] = tmp;
    w->v2Q[v] = (*Q_lp)++;
    w->v2Q[tmp] = q;
}

static void report(worker* w, int lvl) {
    const problem* pb = w->pb;
    int32_t nl = w->L_lp[lvl + 1], nr = w->R_lp[lvl + 1];
    uint64_t sl = 0, sr = 0;
    for (int32_t i = 0; i < nl; i++) sl += pb->hL[w->L[i]];
    for (int32_t i = 0; i < nr; i++) sr += pb->hR[w->R[i]];
    uint64_t h = pb->r_is_u ? combine(sr, sl, (uint64_t)nr, (uint64_t)nl)
                            : combine(sl, sr, (uint64_t)nl, (uint64_t)nr);
    w->count++;
    w->hash += h;
    if (pb->emit) {
        for (int32_t i = 0; i < nl; i++) w->ea[i] = (uint32_t)w->L[i];
        for (int32_t i = 0; i < nr; i++) w->eb[i] = (uint32_t)w->R[i];
        if (pb->r_is_u) pb->emit(pb->user, w->eb, (uint32_t)nr, w->ea, (uint32_t)nl);
        else pb->emit(pb->user, w->ea, (uint32_t)nl, w->eb, (uint32_t)nr);
    }
}

/* PROCESS(x) at level lvl: src/mbe_noWS.cuh:160-316 (level 0) and :396-553
 * (level >= 1), with the noRS forward scans of src/mbe_noRS.cuh:216-287.
 * Returns 1 when the node is maximal and P' is nonempty (descend). */
static int process(worker* w, int lvl) {
    const problem* pb = w->pb;
    const int32_t xv = w->x[lvl];
    int32_t* L = w->L; int32_t* u2L = w->u2L; int32_t* L_buf = w->L_buf;
    int32_t* num_N_u = w->num_N_u;

    /* R' <- R u {x}; P' <- {}; Q' <- Q; L' <- {} */
    w->R_lp[lvl + 1] = w->R_lp[lvl];
    w->R[w->R_lp[lvl + 1]++] = xv;
    w->P_lp[lvl + 1] = 0;
    w->Q_lp[lvl + 1] = w->Q_lp[lvl];
    w->L_lp[lvl + 1] = 0;

    /* L' <- {u in L | (u, x) in E}: gather, then swap into the L prefix */
    int32_t nlp = 0;
    for (uint32_t e = pb->offR[xv]; e < pb->offR[xv + 1]; e++) {
        int32_t u = (int32_t)pb->adjR[e];
        w->work++;
        if (u2L[u] < w->L_lp[lvl]) L_buf[nlp++] = u;
    }
    for (int32_t i = 0; i < nlp; i++) {
        int32_t u = L_buf[i];
        int32_t l = u2L[u];
        int32_t pos = w->L_lp[lvl + 1]++;
        int32_t tmp = L[pos];
        L[pos] = L[l];
        L[l] = tmp;
        int32_t t2 = u2L[L[l]];
        u2L[L[l]] = u2L[u];
        u2L[u] = t2;
    }
    const int32_t lp_next = w->L_lp[lvl + 1];

    /* maximality check against Q (reverse scan gather) */
    int32_t nt = 0;
    for (int32_t i = 0; i < lp_next; i++) {
        int32_t u = L[i];
        for (uint32_t e = pb->offL[u]; e < pb->offL[u + 1]; e++) {
            int32_t v = (int32_t)pb->adjL[e];
            w->work++;
            if (w->v2Q[v] < w->Q_lp[lvl] && num_N_u[v]++ == 0) L_buf[nt++] = v;
        }
    }
    int maximal = 1;
    if (pb->rs) {
        for (int32_t i = 0; i < nt; i++) {
            int32_t v = L_buf[i];
            if (num_N_u[v] == nlp) maximal = 0;
            num_N_u[v] = 0;
        }
    } else {
        /* noRS: recount every gathered v over N(v) (src/mbe_noRS.cuh:231-262) */
        for (int32_t i = 0; i < nt; i++) num_N_u[L_buf[i]] = 0;
        for (int32_t i = 0; i < nt; i++) {
            int32_t v = L_buf[i];
            for (uint32_t e = pb->offR[v]; e < pb->offR[v + 1]; e++) {
                w->work++;
                if (u2L[pb->adjR[e]] < lp_next) num_N_u[v]++;
            }
        }
        for (int32_t i = 0; i < nt; i++) {
            int32_t v = L_buf[i];
            if (num_N_u[v] == nlp) maximal = 0;
            num_N_u[v] = 0;
        }
    }
    if (!maximal) return 0;

    /* expansion: classify v in P by c = |N(v) n L'| */
    nt = 0;
    if (pb->rs) {
        for (int32_t i = 0; i < lp_next; i++) {
            int32_t u = L[i];
            for (uint32_t e = pb->offL[u]; e < pb->offL[u + 1]; e++) {
                int32_t v = (int32_t)pb->adjL[e];
                w->work++;
                if (w->v2P[v] < w->P_lp[lvl] && num_N_u[v]++ == 0) L_buf[nt++] = v;
            }
        }
    } else {
        /* noRS: scan N(v) for every v in P (src/mbe_noRS.cuh:274-287) */
        for (int32_t i = 0; i < w->P_lp[lvl]; i++) {
            int32_t v = w->P[i];
            for (uint32_t e = pb->offR[v]; e < pb->offR[v + 1]; e++) {
                w->work++;
                if (u2L[pb->adjR[e]] < lp_next && num_N_u[v]++ == 0) L_buf[nt++] = v;
            }
        }
    }
    for (int32_t i = 0; i < nt; i++) {
        int32_t v = L_buf[i];
        if (num_N_u[v] == nlp) w->R[w->R_lp[lvl + 1]++] = v;
    }
    for (int32_t i = 0; i < nt; i++) {
        int32_t v = L_buf[i];
        int32_t p = w->v2P[v];
        if (num_N_u[v] != nlp) {
            int32_t pos = w->P_lp[lvl + 1]++;
            int32_t tmp = w->P[pos];
            w->P[pos] = v;
            w->P[p] = tmp;
            w->v2P[v] = pos;
            w->v2P[tmp] = p;
        }
        num_N_u[v] = 0;
    }
    report(w, lvl);
    return w->P_lp[lvl + 1] != 0;
}

/* candidate selection at level >= 1: src/mbe_cuMBE.cuh:345-390 (ES) or
 * src/mbe_noES.cuh:352,365-366 (full scan); the chosen v moves to the end of
 * the P prefix. */
static void select_candidate(worker* w, int lvl) {
    const problem* pb = w->pb;
    int32_t best = ORACLE_INF, imin = w->P_lp[lvl] - 1;
    for (int32_t i = w->P_lp[lvl]; i-- > 0;) {
        if (pb->es && best == w->pre_min[lvl]) break;
        int32_t v = w->P[i];
        int32_t c = 0;
        for (uint32_t e = pb->offR[v]; e < pb->offR[v + 1]; e++) {
            w->work++;
            if (w->u2L[pb->adjR[e]] < w->L_lp[lvl]) {
                int32_t old = c++;
                if (pb->es && old == best) break;
            }
        }
        if (c < best) { imin = i; best = c; }
    }
    w->pre_min[lvl] = best;
    int32_t last = w->P_lp[lvl] - 1;
    int32_t tmp = w->P[last];
    w->P[last] = w->P[imin];
    w->P[imin] = tmp;
    w->v2P[w->P[last]] = last;
    w->v2P[tmp] = imin;
}

/* Twin removal of iMBEA (Zhang et al., BMC Bioinformatics 2014), which the
 * reference leaves out; used only to check the GPU engine's default pruning.
 * After node x at lvl, every v left in P with N(v) n L = L' = N(x) n L leaves
 * P: selected later it would repeat L', which x (then in Q) covers.
 * prune == 2 applies a plausible wrong rule instead -- dropping every v whose
 * row merely contains L' -- so that the tests can show they catch it. */
static void drop_twins(worker* w, int lvl) {
    const problem* pb = w->pb;
    const int32_t Llp = w->L_lp[lvl], nl = w->L_lp[lvl + 1];
    for (int32_t i = w->P_lp[lvl]; i-- > 0;) {
        const int32_t v = w->P[i];
        int32_t in_l = 0, in_lp = 0;
        for (uint32_t e = pb->offR[v]; e < pb->offR[v + 1]; e++) {
            const int32_t l = w->u2L[pb->adjR[e]];
            if (l < Llp) { in_l++; if (l < nl) in_lp++; }
        }
        if (in_lp == nl && (in_l == nl || pb->prune == 2)) {  /* swap v out of the P prefix */
            const int32_t last = --w->P_lp[lvl];
            const int32_t tmp = w->P[last];
            w->P[last] = v; w->P[i] = tmp;
            w->v2P[v] = last; w->v2P[tmp] = i;
        }
    }
}

/* Claim the next root (decreasing index). Roots outside the requested
 * subset are skipped; the Q sweep of the next claim moves them into Q
 * exactly as roots claimed by other blocks are (src/mbe_noWS.cuh:135-150). */
static int64_t claim_root(problem* pb) {
    for (;;) {
        long long i = atomic_fetch_sub(&pb->cursor, 1);
        if (i < 0) return -1;
        if (i >= pb->sub_begin && i < pb->sub_end && (i - pb->sub_begin) % pb->sub_stride == 0)
            return (int64_t)i;
    }
}

/* one reference block's phase-1 loop, serially: src/mbe_noWS.cuh:101-615 */
static void run_worker(worker* w, problem* pb) {
    int lvl = 0, is_up = 1;
    w->P_lp[0] = (int32_t)pb->nR;
    while (lvl >= 0) {
        int is_recursive = 0;
        if (lvl == 0) {
            for (;;) {
                int32_t before = w->P_lp[0];
                int64_t i = claim_root(pb);
                w->P_lp[0] = (int32_t)i;
                if (i < 0) break;
                /* Q <- Q u {ori_P[j] : i < j < before} (src/mbe_noWS.cuh:135-150) */
                for (int32_t j = (int32_t)i + 1; j < before; j++) q_push(w, &w->Q_lp[0], (int32_t)pb->ori_P[j]);
                /* reset P to root order (src/mbe_noWS.cuh:152-154) */
                for (int32_t j = 0; j < before; j++) {
                    w->P[j] = (int32_t)pb->ori_P[j];
                    w->v2P[pb->ori_P[j]] = j;
                }
                w->x[0] = (int32_t)pb->ori_P[i];
                if (process(w, 0)) { is_recursive = 1; break; }
                q_push(w, &w->Q_lp[0], w->x[0]);
            }
        } else {
            while (w->P_lp[lvl] != 0) {
                select_candidate(w, lvl);
                w->x[lvl] = w->P[--w->P_lp[lvl]];
                w->nodes++;
                int descend = process(w, lvl);
                if (pb->prune) drop_twins(w, lvl);
                if (descend) { is_recursive = 1; break; }
                q_push(w, &w->Q_lp[lvl], w->x[lvl]);
                is_up = w->P_lp[lvl] == 0;
            }
        }
        if (is_recursive) {
            lvl++;
        } else if (is_up && lvl--) {
            /* ascend: Q[lvl] <- Q[lvl] u {x[lvl]} (src/mbe_noWS.cuh:590-601) */
            q_push(w, &w->Q_lp[lvl], w->x[lvl]);
        }
    }
}

static void* thread_main(void* arg) {
    worker* w = (worker*)arg;
    run_worker(w, (problem*)w->pb);
    return NULL;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static int solve_impl(uint32_t n_u, uint32_t n_v, const uint32_t* off_u, const uint32_t* adj_u,
                      int rs, int es, int side, int threads, mbe_oracle_emit_fn emit,
                      void* user, int64_t rb, int64_t re, int64_t rstride,
                      mbe_oracle_result* out, int prune, uint64_t* nodes) {
    if (!out || (!off_u && n_u) || threads < 1) return -1;
    memset(out, 0, sizeof(*out));
    if (emit && threads != 1) return -1;
    side_csr U = {0}, V = {0};
    int rc = build_sides(n_u, n_v, off_u, adj_u, &U, &V);
    if (rc) { free_side(&U); free_side(&V); return rc; }

    problem pb;
    memset(&pb, 0, sizeof(pb));
    /* enumerate the smaller side; rows on a tie (src/main.cu:65-70) */
    int r_is_u = side < 0 ? (n_u <= n_v) : (side == 0);
    const side_csr* Rs = r_is_u ? &U : &V;
    const side_csr* Ls = r_is_u ? &V : &U;
    pb.nR = Rs->n; pb.nL = Ls->n;
    pb.offR = Rs->off; pb.adjR = Rs->adj;
    pb.offL = Ls->off; pb.adjL = Ls->adj;
    pb.r_is_u = r_is_u;
    pb.rs = rs; pb.es = es; pb.emit = emit; pb.user = user; pb.prune = prune;
    pb.sub_begin = rb;
    pb.sub_end = re > 0 ? re : (int64_t)1 << 62;
    pb.sub_stride = rstride > 0 ? rstride : 1;
    pb.hR = malloc(sizeof(uint64_t) * ((size_t)pb.nR + 1));
    pb.hL = malloc(sizeof(uint64_t) * ((size_t)pb.nL + 1));
    pb.ori_P = malloc(sizeof(uint32_t) * ((size_t)pb.nR + 1));
    vd_pair* vd = malloc(sizeof(vd_pair) * ((size_t)pb.nR + 1));
    if (!pb.hR || !pb.hL || !pb.ori_P || !vd) { rc = -2; goto done; }
    uint64_t saltR = r_is_u ? MBE_HASH_SALT_U : MBE_HASH_SALT_V;
    uint64_t saltL = r_is_u ? MBE_HASH_SALT_V : MBE_HASH_SALT_U;
    for (uint32_t v = 0; v < pb.nR; v++) pb.hR[v] = mix64((uint64_t)v ^ saltR);
    for (uint32_t u = 0; u < pb.nL; u++) pb.hL[u] = mix64((uint64_t)u ^ saltL);
    /* root order: degree descending (src/header.cuh:55-63), ties by id */
    for (uint32_t v = 0; v < pb.nR; v++) { vd[v].v = v; vd[v].d = degR(&pb, v); }
    qsort(vd, pb.nR, sizeof(vd_pair), cmp_root);
    for (uint32_t i = 0; i < pb.nR; i++) pb.ori_P[i] = vd[i].v;
    /* P_ptr <- last index with nonzero degree (src/mbe_noWS.cuh:57-62) */
    pb.P_ptr0 = -1;
    for (int64_t i = (int64_t)pb.nR - 1; i >= 0; i--)
        if (degR(&pb, pb.ori_P[i])) { pb.P_ptr0 = i; break; }
    atomic_store(&pb.cursor, (long long)pb.P_ptr0);

    worker* ws = calloc((size_t)threads, sizeof(worker));
    if (!ws) { rc = -2; goto done; }
    for (int t = 0; t < threads; t++)
        if (worker_init(&ws[t], &pb)) { rc = -2; }
    if (!rc) {
        double t0 = now_s();
        if (threads == 1) {
            run_worker(&ws[0], &pb);
        } else {
            pthread_t* th = malloc(sizeof(pthread_t) * threads);
            for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, thread_main, &ws[t]);
            for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
            free(th);
        }
        out->seconds = now_s() - t0;
        for (int t = 0; t < threads; t++) {
            out->count += ws[t].count;
            out->hash += ws[t].hash;
            out->work += ws[t].work;
            if (nodes) *nodes += ws[t].nodes;
        }
    }
    for (int t = 0; t < threads; t++) worker_free(&ws[t]);
    free(ws);
done:
    free(vd); free(pb.hR); free(pb.hL); free(pb.ori_P);
    free_side(&U); free_side(&V);
    return rc;
}

int mbe_oracle_solve(uint32_t n_u, uint32_t n_v, const uint32_t* off_u, const uint32_t* adj_u,
                     int rs, int es, int side, int threads, mbe_oracle_result* out) {
    return solve_impl(n_u, n_v, off_u, adj_u, rs, es, side, threads, NULL, NULL, 0, 0, 1, out,
                      0, NULL);
}

int mbe_oracle_solve_twins(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                           const uint32_t* adj_u, int prune, mbe_oracle_result* out,
                           uint64_t* nodes) {
    if (nodes) *nodes = 0;
    return solve_impl(n_u, n_v, off_u, adj_u, 1, 1, -1, 1, NULL, NULL, 0, 0, 1, out, prune,
                      nodes);
}

int mbe_oracle_solve_roots(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                           const uint32_t* adj_u, int rs, int es, int side, int threads,
                           int64_t root_begin, int64_t root_end, int64_t root_stride,
                           mbe_oracle_result* out) {
    if (root_begin < 0 || root_stride < 0) return -1;
    return solve_impl(n_u, n_v, off_u, adj_u, rs, es, side, threads, NULL, NULL, root_begin,
                      root_end, root_stride, out, 0, NULL);
}

int mbe_oracle_solve_emit(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                          const uint32_t* adj_u, int rs, int es, int side,
                          mbe_oracle_emit_fn emit, void* user, mbe_oracle_result* out) {
    return solve_impl(n_u, n_v, off_u, adj_u, rs, es, side, 1, emit, user, 0, 0, 1, out, 0,
                      NULL);
}

/* ------------------------------------------------------------ brute force */

int mbe_oracle_brute(uint32_t n_u, uint32_t n_v, const uint32_t* off_u, const uint32_t* adj_u,
                     mbe_oracle_result* out) {
    if (!out) return -1;
    memset(out, 0, sizeof(*out));
    side_csr U = {0}, V = {0};
    int rc = build_sides(n_u, n_v, off_u, adj_u, &U, &V);
    if (rc) { free_side(&U); free_side(&V); return rc; }
    int s_is_u = n_u <= n_v;
    const side_csr* S = s_is_u ? &U : &V;   /* subsets are taken on this side */
    uint32_t ns = S->n, no = s_is_u ? n_v : n_u;
    if (ns > 24) { free_side(&U); free_side(&V); return -1; }
    uint32_t words = (no + 63) / 64;
    if (!words) words = 1;
    uint64_t* nb = calloc((size_t)ns * words + 1, sizeof(uint64_t));
    uint64_t* A = malloc(sizeof(uint64_t) * words);
    uint32_t* la = malloc(sizeof(uint32_t) * (no + 1));
    uint32_t* lb = malloc(sizeof(uint32_t) * (ns + 1));
    if (!nb || !A || !la || !lb) { rc = -2; goto out; }
    for (uint32_t s = 0; s < ns; s++)
        for (uint32_t e = S->off[s]; e < S->off[s + 1]; e++)
            nb[(size_t)s * words + S->adj[e] / 64] |= 1ULL << (S->adj[e] % 64);
    double t0 = now_s();
    for (uint64_t mask = 1; mask < (1ULL << ns); mask++) {
        for (uint32_t k = 0; k < words; k++) A[k] = ~0ULL;
        if (no % 64) A[words - 1] = (1ULL << (no % 64)) - 1;
        if (!no) A[0] = 0;
        uint32_t nbsz = 0;
        for (uint32_t s = 0; s < ns; s++)
            if (mask >> s & 1) {
                lb[nbsz++] = s;
                for (uint32_t k = 0; k < words; k++) A[k] &= nb[(size_t)s * words + k];
            }
        uint32_t na = 0;
        for (uint32_t k = 0; k < words; k++) na += (uint32_t)__builtin_popcountll(A[k]);
        if (!na) continue;
        /* closure: no s outside B is adjacent to all of A */
        int closed = 1;
        for (uint32_t s = 0; s < ns && closed; s++) {
            if (mask >> s & 1) continue;
            int covers = 1;
            for (uint32_t k = 0; k < words; k++)
                if ((A[k] & nb[(size_t)s * words + k]) != A[k]) { covers = 0; break; }
            if (covers) closed = 0;
        }
        if (!closed) continue;
        na = 0;
        for (uint32_t k = 0; k < words; k++)
            for (uint64_t bits = A[k]; bits; bits &= bits - 1)
                la[na++] = k * 64 + (uint32_t)__builtin_ctzll(bits);
        out->count++;
        out->hash += s_is_u ? mbe_oracle_biclique_hash(lb, nbsz, la, na)
                            : mbe_oracle_biclique_hash(la, na, lb, nbsz);
    }
    out->seconds = now_s() - t0;
out:
    free(nb); free(A); free(la); free(lb);
    free_side(&U); free_side(&V);
    return rc;
}
