/*
 * CPU oracle for maximal biclique enumeration (MBE) -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a serial (optionally root-parallel) restatement of the reference
 * algorithm (MBEA with iMBEA-style candidate ordering, without absorbed-vertex
 * removal), following the reference kernels line by line at the level of sets:
 *   - side choice and initial state ...... src/main.cu:65-86
 *   - CSR transpose (sorted output) ....... src/transpose.cuh:1-52
 *   - root order (degree descending) ...... src/header.cuh:55-63
 *   - root claim / P reset / Q sweep ...... src/mbe_noWS.cuh:117-181
 *   - L' construction ..................... src/mbe_noWS.cuh:187-209
 *   - maximality check, reverse scan ...... src/mbe_noWS.cuh:215-238
 *   - expansion, reverse scan ............. src/mbe_noWS.cuh:243-316
 *   - candidate selection with early stop . src/mbe_cuMBE.cuh:345-390
 *   - noES selection ...................... src/mbe_noES.cuh:352,365-366
 *   - noRS forward scans .................. src/mbe_noRS.cuh:216-287
 *   - backtrack / ascend .................. src/mbe_noWS.cuh:318-333,584-611
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load this library; the product (libmbe.so) never links it.
 *
 * Result = (count, hash, work):
 *   count  number of maximal bicliques (A subset of U, B subset of V, both
 *          nonempty, A x B subset of E, maximal);
 *   hash   wrapping sum over bicliques of h(A, B) where, with
 *          mix64 = SplitMix64 finaliser,
 *            S_U = sum_{a in A} mix64(a ^ SALT_U),  S_V = sum_{b in B} mix64(b ^ SALT_V)
 *            h   = mix64(S_U ^ rotl64(S_V, 29) ^ (|A| * K1 + |B| * K2))
 *          (ids are file coordinates: U = rows of the .bi file, V = columns);
 *   work   adjacency entries scanned (the roofline's W), meaningful for the
 *          canonical configuration rs=1, es=1, threads=1.
 */
#ifndef MBE_ORACLE_H
#define MBE_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBE_HASH_SALT_U 0x9E3779B97F4A7C15ULL
#define MBE_HASH_SALT_V 0xC2B2AE3D27D4EB4FULL
#define MBE_HASH_K1     0x165667B19E3779F9ULL
#define MBE_HASH_K2     0xD6E8FEB86659FD93ULL

typedef struct {
    uint64_t count;
    uint64_t hash;
    uint64_t work;
    double   seconds;   /* wall time of the enumeration only (no input prep) */
} mbe_oracle_result;

/* Optional per-biclique callback (small graphs): A in U ids, B in V ids,
 * each unsorted. Called from the enumerating thread; threads must be 1. */
typedef void (*mbe_oracle_emit_fn)(void* user, const uint32_t* a, uint32_t na,
                                   const uint32_t* b, uint32_t nb);

/*
 * Solve the graph given as a CSR over U (rows): off_u[n_u+1], adj_u[off_u[n_u]]
 * with V ids < n_v. Lists need not be sorted; duplicates are removed.
 *   rs: 1 = reverse scans from L' (reference default), 0 = forward (noRS)
 *   es: 1 = early-stop candidate selection (default), 0 = full scan (noES)
 *   side: -1 = reference choice (enumerate the smaller side, rows on a tie),
 *          0 = enumerate U, 1 = enumerate V
 *   threads: >=1; >1 claims roots from a shared atomic cursor (count and
 *            hash are exact in any schedule; work is then schedule-dependent)
 * Returns 0 on success, negative on invalid input.
 */
int mbe_oracle_solve(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                     const uint32_t* adj_u, int rs, int es, int side, int threads,
                     mbe_oracle_result* out);

/* Same, restricted to the roots of rank r (root order: degree descending,
 * id ascending) with root_begin <= r < root_end (0 = no bound) and
 * (r - root_begin) % root_stride == 0. The results of a partition of the
 * roots add up to the whole graph's. */
int mbe_oracle_solve_roots(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                           const uint32_t* adj_u, int rs, int es, int side, int threads,
                           int64_t root_begin, int64_t root_end, int64_t root_stride,
                           mbe_oracle_result* out);

/* RS + ES with the reference's side choice, one thread; prune = 1 adds
 * iMBEA's twin removal (not the reference's tree), which the GPU engine
 * uses by default; prune = 2 applies a wrong rule on purpose (drop every
 * candidate whose row contains L'), which the tests show they detect.
 * *nodes receives the tree nodes below the roots. */
int mbe_oracle_solve_twins(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                           const uint32_t* adj_u, int prune, mbe_oracle_result* out,
                           uint64_t* nodes);

int mbe_oracle_solve_emit(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                          const uint32_t* adj_u, int rs, int es, int side,
                          mbe_oracle_emit_fn emit, void* user,
                          mbe_oracle_result* out);

/* Hash of one biclique (A in U ids, B in V ids), exposed for tests. */
uint64_t mbe_oracle_biclique_hash(const uint32_t* a, uint32_t na,
                                  const uint32_t* b, uint32_t nb);

/*
 * Brute-force closure enumeration for tiny graphs: every nonempty B subset of
 * the smaller side with A = N(B) nonempty and N(A) = B is one maximal
 * biclique. Independent of the MBEA restatement above. Smaller side <= 24.
 */
int mbe_oracle_brute(uint32_t n_u, uint32_t n_v, const uint32_t* off_u,
                     const uint32_t* adj_u, mbe_oracle_result* out);

#ifdef __cplusplus
}
#endif
#endif
