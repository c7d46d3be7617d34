// Host-side graph formats and preprocessing for the MBE engine.
//
// File coordinates: U = the rows of a .bi file (the reference's node_r side as
// read, src/main.cu:40-47), V = its columns. Every CSR kept here is sorted and
// deduplicated, which the reference's reader assumes but never checks.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mbe {

struct Graph {
    uint32_t n_u = 0, n_v = 0;
    std::vector<uint32_t> off;  // n_u + 1
    std::vector<uint32_t> adj;  // V ids, each list sorted ascending, unique
    uint64_t edges() const { return off.empty() ? 0 : off.back(); }
};

// Validate and canonicalise a CSR over U: bounds-checks ids, sorts and
// deduplicates each list. Returns false on an out-of-range id or bad offsets.
bool graph_from_csr(uint32_t n_u, uint32_t n_v, const uint32_t* off, const uint32_t* adj,
                    Graph& out, std::string* err);

// .bi text format (written by the reference's data/gen_bi.cpp:72-79, read by
// src/main.cu:38-48): "NUM_R NUM_L NUM_EDGES", then NUM_R lines "start length",
// then NUM_EDGES entries "col weight"; the weight is ignored.
bool read_bi(const std::string& path, Graph& out, std::string* err);
bool write_bi(const std::string& path, const Graph& g, std::string* err);

// Edge list -> Graph with first-seen relabelling of both columns, duplicates
// dropped (the conversion of data/gen_bi.cpp:34-51). skip_words header words
// are skipped, and skip_per_edge words after each pair.
bool edge_list_to_graph(const std::string& path, int skip_words, int skip_per_edge, Graph& out,
                        std::string* err);

// The enumeration view: R = enumerated side (the smaller one, rows on a tie,
// src/main.cu:65-70), relabelled by root order (degree descending, id
// ascending: src/header.cuh:55-63), so that the root task of index i has
// P = {r < i} and Q = {r > i}. L keeps its file ids.
struct Prepared {
    uint32_t nR = 0, nL = 0;
    bool r_is_u = true;               // enumerated side is U (file rows)
    uint32_t n_roots = 0;             // ranks with nonzero degree: [0, n_roots)
    std::vector<uint32_t> offR, adjR; // by rank: sorted L ids
    std::vector<uint32_t> offL, adjL; // by L id: sorted ranks
    std::vector<uint32_t> rank2id;    // rank -> file id on the R side
    std::vector<uint64_t> hR, hL;     // per-vertex hash terms (see hash.h)
    uint64_t edges() const { return adjR.size(); }
};

// side: -1 reference choice, 0 force R = U, 1 force R = V.
void prepare(const Graph& g, int side, Prepared& p);

// ---------------------------------------------------------------- generators
// All generators use SplitMix64-seeded xoshiro256** and return sorted,
// deduplicated CSR graphs over U.

// Chung-Lu bipartite graph: expected-degree weights w_i ~ (i+1)^(-1/(gamma-1))
// on each side, expected degrees capped at max_deg (0 = the other side's
// size); endpoint pairs are drawn from the two weight distributions and
// duplicates rejected until m unique edges exist (or 64m + 10^6 draws).
Graph gen_chung_lu(uint32_t n_u, uint32_t n_v, uint64_t m, double gamma_u, double gamma_v,
                   uint32_t max_deg, uint64_t seed);

// Erdos-Renyi background G(n_u, n_v, p) by geometric skipping, plus
// n_planted complete bicliques with side sizes uniform in [lo, hi] whose
// vertices come from the top top_frac of each side by background degree.
Graph gen_er_planted(uint32_t n_u, uint32_t n_v, double p, uint32_t n_planted, uint32_t lo,
                     uint32_t hi, double top_frac, uint64_t seed);

// Parameters of one graph of the batched workload for a given seed:
// |U| log-uniform in [50, 5000], |V| log-uniform in [20, 2000], edges
// log-uniform in [200, 50000] capped at |U||V|/4, exponent uniform in [1.8, 2.5].
struct BatchParams {
    uint32_t n_u, n_v;
    uint64_t m;
    double gamma;
};
BatchParams batch_params(uint64_t seed);
Graph gen_batch_graph(uint64_t seed);

}  // namespace mbe
