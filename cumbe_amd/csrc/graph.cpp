// Host-side graph formats, preprocessing and seeded generators.
#include "graph.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <unordered_set>

#include "hash.h"

namespace mbe {

namespace {

void set_err(std::string* err, const std::string& msg) {
    if (err) *err = msg;
}

// SplitMix64 (used to seed) and xoshiro256** (Blackman & Vigna, 2018).
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9E3779B97F4A7C15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ULL;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBULL;
            w = t ^ (t >> 31);
        }
    }
    uint64_t next() {
        const uint64_t result = rotl64(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl64(s[3], 45);
        return result;
    }
    double uniform() { return (next() >> 11) * 0x1.0p-53; }  // [0, 1)
    uint64_t below(uint64_t n) { return n ? (uint64_t)(uniform() * (double)n) % n : 0; }
    uint32_t range(uint32_t lo, uint32_t hi) { return lo + (uint32_t)below((uint64_t)hi - lo + 1); }
};

void finish_lists(Graph& g, std::vector<std::vector<uint32_t>>& lists) {
    g.off.assign((size_t)g.n_u + 1, 0);
    g.adj.clear();
    for (uint32_t u = 0; u < g.n_u; u++) {
        auto& l = lists[u];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        g.adj.insert(g.adj.end(), l.begin(), l.end());
        g.off[u + 1] = (uint32_t)g.adj.size();
    }
}

// Chung-Lu weights w_i = (i+1)^(-1/(gamma-1)), with expected degree
// m * w_i / sum(w) capped at cap: the cap is applied as min(w_i, c) with c
// found by bisection so that the largest expected degree equals cap.
std::vector<double> cl_weights(uint32_t n, double gamma, uint64_t m, double cap) {
    std::vector<double> w(n);
    const double ex = -1.0 / (gamma - 1.0);
    for (uint32_t i = 0; i < n; i++) w[i] = std::pow((double)(i + 1), ex);
    auto top_deg = [&](double c) {
        double s = 0;
        for (double x : w) s += std::min(x, c);
        return (double)m * std::min(w[0], c) / s;
    };
    if (n && top_deg(w[0]) > cap) {
        double lo = 0, hi = w[0];
        for (int it = 0; it < 200; it++) {
            double mid = 0.5 * (lo + hi);
            if (top_deg(mid) > cap) hi = mid; else lo = mid;
        }
        for (double& x : w) x = std::min(x, lo);
    }
    return w;
}

uint32_t draw(const std::vector<double>& cdf, Rng& r) {
    double t = r.uniform() * cdf.back();
    size_t k = std::upper_bound(cdf.begin(), cdf.end(), t) - cdf.begin();
    return (uint32_t)std::min(k, cdf.size() - 1);
}

}  // namespace

bool graph_from_csr(uint32_t n_u, uint32_t n_v, const uint32_t* off, const uint32_t* adj,
                    Graph& out, std::string* err) {
    if (n_u && !off) { set_err(err, "null offsets"); return false; }
    if (off && off[0] != 0) { set_err(err, "offsets must start at 0"); return false; }
    std::vector<std::vector<uint32_t>> lists(n_u);
    for (uint32_t u = 0; u < n_u; u++) {
        if (off[u + 1] < off[u]) { set_err(err, "offsets not monotone"); return false; }
        for (uint32_t k = off[u]; k < off[u + 1]; k++) {
            if (adj[k] >= n_v) { set_err(err, "neighbour id out of range"); return false; }
            lists[u].push_back(adj[k]);
        }
    }
    out.n_u = n_u;
    out.n_v = n_v;
    finish_lists(out, lists);
    return true;
}

bool read_bi(const std::string& path, Graph& out, std::string* err) {
    std::ifstream in(path);
    if (!in) { set_err(err, "cannot open " + path); return false; }
    long long nr, nl, ne;
    if (!(in >> nr >> nl >> ne) || nr < 0 || nl < 0 || ne < 0 || nr > 0xFFFFFFF0LL ||
        nl > 0xFFFFFFF0LL || ne > 0xFFFFFFF0LL) {
        set_err(err, "bad .bi header");
        return false;
    }
    std::vector<long long> start(nr), len(nr);
    for (long long i = 0; i < nr; i++)
        if (!(in >> start[i] >> len[i]) || start[i] < 0 || len[i] < 0 || start[i] + len[i] > ne) {
            set_err(err, "bad .bi offset line " + std::to_string(i));
            return false;
        }
    std::vector<uint32_t> cols(ne);
    for (long long e = 0; e < ne; e++) {
        long long c, w;
        if (!(in >> c >> w) || c < 0 || c >= nl) {
            set_err(err, "bad .bi edge entry " + std::to_string(e));
            return false;
        }
        cols[e] = (uint32_t)c;
    }
    std::vector<std::vector<uint32_t>> lists(nr);
    for (long long i = 0; i < nr; i++)
        lists[i].assign(cols.begin() + start[i], cols.begin() + start[i] + len[i]);
    out.n_u = (uint32_t)nr;
    out.n_v = (uint32_t)nl;
    finish_lists(out, lists);
    return true;
}

bool write_bi(const std::string& path, const Graph& g, std::string* err) {
    std::ofstream o(path);
    if (!o) { set_err(err, "cannot write " + path); return false; }
    o << g.n_u << ' ' << g.n_v << ' ' << g.edges() << "\n\n";
    for (uint32_t u = 0; u < g.n_u; u++) o << g.off[u] << ' ' << (g.off[u + 1] - g.off[u]) << '\n';
    for (uint64_t e = 0; e < g.edges(); e++) o << '\n' << g.adj[e] << " 1";
    o << '\n';
    if (!o) { set_err(err, "write failed: " + path); return false; }
    return true;
}

bool edge_list_to_graph(const std::string& path, int skip_words, int skip_per_edge, Graph& out,
                        std::string* err) {
    std::ifstream in(path);
    if (!in) { set_err(err, "cannot open " + path); return false; }
    std::string tok;
    for (int i = 0; i < skip_words; i++) in >> tok;
    std::map<long long, uint32_t> rid, cid;
    std::vector<std::vector<uint32_t>> lists;
    long long r, c;
    while (in >> r >> c) {
        for (int i = 0; i < skip_per_edge; i++) in >> tok;
        auto ir = rid.emplace(r, (uint32_t)rid.size());
        if (ir.second) lists.emplace_back();
        auto ic = cid.emplace(c, (uint32_t)cid.size());
        lists[ir.first->second].push_back(ic.first->second);
    }
    out.n_u = (uint32_t)rid.size();
    out.n_v = (uint32_t)cid.size();
    finish_lists(out, lists);
    return true;
}

void prepare(const Graph& g, int side, Prepared& p) {
    // V-side lists by counting transpose (sorted, since U is scanned in order)
    std::vector<uint32_t> offV((size_t)g.n_v + 1, 0), adjV(g.edges());
    for (uint64_t e = 0; e < g.edges(); e++) offV[g.adj[e] + 1]++;
    for (uint32_t v = 0; v < g.n_v; v++) offV[v + 1] += offV[v];
    {
        std::vector<uint32_t> cur(offV.begin(), offV.end() - 1);
        for (uint32_t u = 0; u < g.n_u; u++)
            for (uint32_t k = g.off[u]; k < g.off[u + 1]; k++) adjV[cur[g.adj[k]]++] = u;
    }
    p.r_is_u = side < 0 ? (g.n_u <= g.n_v) : (side == 0);
    const std::vector<uint32_t>& oR = p.r_is_u ? g.off : offV;
    const std::vector<uint32_t>& aR = p.r_is_u ? g.adj : adjV;
    p.nR = p.r_is_u ? g.n_u : g.n_v;
    p.nL = p.r_is_u ? g.n_v : g.n_u;

    // root order: degree descending, id ascending
    p.rank2id.resize(p.nR);
    std::iota(p.rank2id.begin(), p.rank2id.end(), 0u);
    std::sort(p.rank2id.begin(), p.rank2id.end(), [&](uint32_t a, uint32_t b) {
        uint32_t da = oR[a + 1] - oR[a], db = oR[b + 1] - oR[b];
        return da != db ? da > db : a < b;
    });
    std::vector<uint32_t> id2rank(p.nR);
    for (uint32_t r = 0; r < p.nR; r++) id2rank[p.rank2id[r]] = r;
    p.n_roots = 0;
    p.offR.assign((size_t)p.nR + 1, 0);
    p.adjR.clear();
    p.adjR.reserve(g.edges());
    for (uint32_t r = 0; r < p.nR; r++) {
        uint32_t v = p.rank2id[r];
        p.adjR.insert(p.adjR.end(), aR.begin() + oR[v], aR.begin() + oR[v + 1]);
        p.offR[r + 1] = (uint32_t)p.adjR.size();
        if (oR[v + 1] > oR[v]) p.n_roots = r + 1;
    }
    // L-side lists of ranks, sorted (scan ranks in order)
    p.offL.assign((size_t)p.nL + 1, 0);
    p.adjL.assign(p.adjR.size(), 0);
    for (uint64_t e = 0; e < p.adjR.size(); e++) p.offL[p.adjR[e] + 1]++;
    for (uint32_t u = 0; u < p.nL; u++) p.offL[u + 1] += p.offL[u];
    {
        std::vector<uint32_t> cur(p.offL.begin(), p.offL.end() - 1);
        for (uint32_t r = 0; r < p.nR; r++)
            for (uint32_t k = p.offR[r]; k < p.offR[r + 1]; k++) p.adjL[cur[p.adjR[k]]++] = r;
    }
    const uint64_t saltR = p.r_is_u ? kSaltU : kSaltV;
    const uint64_t saltL = p.r_is_u ? kSaltV : kSaltU;
    p.hR.resize(p.nR);
    p.hL.resize(p.nL);
    for (uint32_t r = 0; r < p.nR; r++) p.hR[r] = mix64((uint64_t)p.rank2id[r] ^ saltR);
    for (uint32_t u = 0; u < p.nL; u++) p.hL[u] = mix64((uint64_t)u ^ saltL);
}

Graph gen_chung_lu(uint32_t n_u, uint32_t n_v, uint64_t m, double gamma_u, double gamma_v,
                   uint32_t max_deg, uint64_t seed) {
    Graph g;
    g.n_u = n_u;
    g.n_v = n_v;
    std::vector<std::vector<uint32_t>> lists(n_u);
    if (n_u && n_v && m) {
        const uint64_t cells = (uint64_t)n_u * n_v;
        if (m > cells) m = cells;
        double cap_u = max_deg ? std::min<double>(max_deg, n_v) : n_v;
        double cap_v = max_deg ? std::min<double>(max_deg, n_u) : n_u;
        auto wu = cl_weights(n_u, gamma_u, m, cap_u);
        auto wv = cl_weights(n_v, gamma_v, m, cap_v);
        std::vector<double> cu(n_u), cv(n_v);
        std::partial_sum(wu.begin(), wu.end(), cu.begin());
        std::partial_sum(wv.begin(), wv.end(), cv.begin());
        Rng rng(seed);
        std::unordered_set<uint64_t> seen;
        seen.reserve(m * 2);
        const uint64_t max_draws = 64 * m + 1000000;
        for (uint64_t d = 0; d < max_draws && seen.size() < m; d++) {
            uint32_t u = draw(cu, rng), v = draw(cv, rng);
            if (seen.insert((uint64_t)u * n_v + v).second) lists[u].push_back(v);
        }
    }
    finish_lists(g, lists);
    return g;
}

Graph gen_er_planted(uint32_t n_u, uint32_t n_v, double p, uint32_t n_planted, uint32_t lo,
                     uint32_t hi, double top_frac, uint64_t seed) {
    Graph g;
    g.n_u = n_u;
    g.n_v = n_v;
    std::vector<std::vector<uint32_t>> lists(n_u);
    Rng rng(seed);
    const uint64_t cells = (uint64_t)n_u * n_v;
    if (p > 0 && cells) {
        // geometric skipping (Batagelj & Brandes, 2005)
        const double lq = std::log1p(-std::min(p, 1.0 - 1e-12));
        uint64_t pos = 0;
        bool first = true;
        for (;;) {
            double r = rng.uniform();
            uint64_t skip = (uint64_t)std::floor(std::log1p(-r) / lq);
            pos += first ? skip : skip + 1;
            first = false;
            if (pos >= cells) break;
            lists[pos / n_v].push_back((uint32_t)(pos % n_v));
        }
    }
    std::vector<uint32_t> du(n_u, 0), dv(n_v, 0);
    for (uint32_t u = 0; u < n_u; u++)
        for (uint32_t v : lists[u]) du[u]++, dv[v]++;
    auto top = [](const std::vector<uint32_t>& deg, double frac) {
        std::vector<uint32_t> ids(deg.size());
        std::iota(ids.begin(), ids.end(), 0u);
        std::sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) {
            return deg[a] != deg[b] ? deg[a] > deg[b] : a < b;
        });
        size_t k = (size_t)std::ceil(frac * (double)deg.size());
        ids.resize(std::min(k, ids.size()));
        return ids;
    };
    std::vector<uint32_t> tu = top(du, top_frac), tv = top(dv, top_frac);
    for (uint32_t b = 0; b < n_planted && !tu.empty() && !tv.empty(); b++) {
        uint32_t a_sz = std::min<uint32_t>(rng.range(lo, hi), (uint32_t)tu.size());
        uint32_t b_sz = std::min<uint32_t>(rng.range(lo, hi), (uint32_t)tv.size());
        // partial Fisher-Yates on both pools
        for (uint32_t i = 0; i < a_sz; i++) std::swap(tu[i], tu[i + rng.below(tu.size() - i)]);
        for (uint32_t i = 0; i < b_sz; i++) std::swap(tv[i], tv[i + rng.below(tv.size() - i)]);
        for (uint32_t i = 0; i < a_sz; i++)
            for (uint32_t j = 0; j < b_sz; j++) lists[tu[i]].push_back(tv[j]);
    }
    finish_lists(g, lists);
    return g;
}

BatchParams batch_params(uint64_t seed) {
    Rng rng(seed ^ 0xB47C4ED5EEDULL);
    auto logu = [&](double lo, double hi) {
        return std::exp(std::log(lo) + rng.uniform() * (std::log(hi) - std::log(lo)));
    };
    BatchParams bp;
    bp.n_u = (uint32_t)std::llround(logu(50, 5000));
    bp.n_v = (uint32_t)std::llround(logu(20, 2000));
    bp.m = (uint64_t)std::llround(logu(200, 50000));
    uint64_t cap = (uint64_t)bp.n_u * bp.n_v / 4;
    if (bp.m > cap) bp.m = cap;
    bp.gamma = 1.8 + rng.uniform() * 0.7;
    return bp;
}

Graph gen_batch_graph(uint64_t seed) {
    BatchParams bp = batch_params(seed);
    return gen_chung_lu(bp.n_u, bp.n_v, bp.m, bp.gamma, bp.gamma, 0, seed);
}

}  // namespace mbe
