// `mbe` command-line tool: the reference's entry point
//   mbe <dataset.bi> [cuMBE|noRS|noES|noWS]
// with its stdout lines and result/<dataset>_<algo> file (src/main.cu:121-234),
// plus a "hash:" line, and two data tools:
//   mbe convert <edge-list> <out.bi> <skip_words> <skip_words_per_edge>
//       (the conversion of data/gen_bi.cpp)
//   mbe gen chung-lu <out.bi> n_u n_v m gamma_u gamma_v max_deg seed
//   mbe gen er-planted <out.bi> n_u n_v p n_planted lo hi top_frac seed
//   mbe gen batch <out.bi> seed
// and one debugging tool:
//   mbe list <dataset.bi> <out.txt> [cuMBE|noRS|noES|noWS]
//       writes every maximal biclique, one "A: rows | B: columns" line each
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "mbe.h"

static int die(mbe_status s, const char* what) {
    std::cerr << what << ": " << mbe_status_str(s) << " (" << mbe_last_error() << ")\n";
    return 1;
}

static int cmd_convert(int argc, char** argv) {
    if (argc < 6) {
        std::cerr << "usage: mbe convert <edge-list> <out.bi> <skip_words> <skip_words_per_edge>\n";
        return 2;
    }
    mbe_graph* g = nullptr;
    mbe_status s = mbe_graph_from_edge_list(argv[2], atoi(argv[4]), atoi(argv[5]), &g);
    if (s != MBE_OK) return die(s, "convert");
    s = mbe_graph_write_bi(g, argv[3]);
    mbe_graph_free(g);
    return s == MBE_OK ? 0 : die(s, "write");
}

static int cmd_gen(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "usage: mbe gen <chung-lu|er-planted|batch> <out.bi> args...\n";
        return 2;
    }
    std::string kind = argv[2];
    mbe_graph* g = nullptr;
    mbe_status s = MBE_EINVAL;
    if (kind == "chung-lu" && argc >= 11) {
        s = mbe_gen_chung_lu(strtoul(argv[4], 0, 10), strtoul(argv[5], 0, 10), strtoull(argv[6], 0, 10),
                             atof(argv[7]), atof(argv[8]), strtoul(argv[9], 0, 10),
                             strtoull(argv[10], 0, 10), &g);
    } else if (kind == "er-planted" && argc >= 12) {
        s = mbe_gen_er_planted(strtoul(argv[4], 0, 10), strtoul(argv[5], 0, 10), atof(argv[6]),
                               strtoul(argv[7], 0, 10), strtoul(argv[8], 0, 10),
                               strtoul(argv[9], 0, 10), atof(argv[10]), strtoull(argv[11], 0, 10), &g);
    } else if (kind == "batch" && argc >= 5) {
        s = mbe_gen_batch_graph(strtoull(argv[4], 0, 10), &g);
    } else {
        std::cerr << "bad generator arguments\n";
        return 2;
    }
    if (s != MBE_OK) return die(s, "gen");
    s = mbe_graph_write_bi(g, argv[3]);
    mbe_graph_free(g);
    return s == MBE_OK ? 0 : die(s, "write");
}

static int cmd_list(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "usage: mbe list <dataset.bi> <out.txt> [cuMBE|noRS|noES|noWS]\n";
        return 2;
    }
    std::string algo = argc > 4 ? argv[4] : "cuMBE";
    mbe_variant var = algo == "noRS" ? MBE_NORS : algo == "noES" ? MBE_NOES : algo == "noWS" ? MBE_NOWS : MBE_CUMBE;
    mbe_graph* g = nullptr;
    mbe_status s = mbe_graph_load_bi(argv[2], &g);
    if (s != MBE_OK) return die(s, "load");
    uint64_t words = 0;
    mbe_result r{};
    s = mbe_solve_list(g, var, nullptr, nullptr, 0, &words, &r);  // size query
    std::vector<uint32_t> buf(words);
    if (s == MBE_OK) s = mbe_solve_list(g, var, nullptr, buf.data(), words, &words, &r);
    mbe_graph_free(g);
    if (s != MBE_OK) return die(s, "list");
    std::ofstream out(argv[3]);
    for (uint64_t i = 0; i < words;) {
        const uint32_t na = buf[i], nb = buf[i + 1];
        out << "A:";
        for (uint32_t k = 0; k < na; k++) out << ' ' << buf[i + 2 + k];
        out << " | B:";
        for (uint32_t k = 0; k < nb; k++) out << ' ' << buf[i + 2 + na + k];
        out << '\n';
        i += 2 + (uint64_t)na + nb;
    }
    if (!out) {
        std::cerr << "list: cannot write " << argv[3] << "\n";
        return 1;
    }
    std::cout << "maximal bicliques: " << r.count << "\nhash: 0x" << std::hex << std::setw(16)
              << std::setfill('0') << r.hash << std::dec << "\n";
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: mbe <dataset.bi> [cuMBE|noRS|noES|noWS]\n";
        return 2;
    }
    std::string a1 = argv[1];
    if (a1 == "convert") return cmd_convert(argc, argv);
    if (a1 == "gen") return cmd_gen(argc, argv);
    if (a1 == "list") return cmd_list(argc, argv);

    time_t now = time(nullptr);
    std::string dataset = a1.substr(a1.rfind('/') + 1);
    std::string algo = argc > 2 ? argv[2] : "cuMBE";
    if (algo != "noRS" && algo != "noES" && algo != "noWS") algo = "cuMBE";  // src/main.cu:122
    mbe_variant var = algo == "noRS" ? MBE_NORS : algo == "noES" ? MBE_NOES : algo == "noWS" ? MBE_NOWS : MBE_CUMBE;

    mbe_graph* g = nullptr;
    mbe_status s = mbe_graph_load_bi(argv[1], &g);
    if (s != MBE_OK) return die(s, "load");
    uint32_t nu = 0, nv = 0;
    uint64_t ne = 0;
    mbe_graph_info(g, &nu, &nv, &ne);
    uint32_t nr = nu <= nv ? nu : nv, nl = nu <= nv ? nv : nu;  // src/main.cu:65-70

    // MBE_SECTIONS=1, or running as mbe_desection, adds the per-section time
    // split, like the reference's DESECTION build (src/header.cuh:25-32,
    // src/main.cu:208-216; its harness runs ./bin/mbe_desection,
    // utils/run.py:45-57)
    const std::string self = std::string(argv[0]).substr(std::string(argv[0]).rfind('/') + 1);
    const char* sec_env = getenv("MBE_SECTIONS");
    const bool sections = (sec_env && atoi(sec_env) != 0) || self == "mbe_desection";
    // MBE_VARIANCE=1, or running as mbe_devariance, prints each worker's busy
    // ticks between the header and the result lines, like the reference's
    // DEVARIANCE build (src/mbe_noES.cuh:49-55, read by utils/plot_variance.py)
    const char* var_env = getenv("MBE_VARIANCE");
    const bool variance = (var_env && atoi(var_env) != 0) || self == "mbe_devariance";
    mbe_tuning tun{};
    tun.side = -1;
    tun.flags = (sections ? MBE_FLAG_SECTIONS : 0) | (variance ? MBE_FLAG_VARIANCE : 0);
    mbe_result r{};
    s = mbe_solve(g, var, &tun, &r);
    mbe_graph_free(g);
    if (s != MBE_OK) return die(s, "solve");

    std::string filename = dataset.substr(0, dataset.rfind('.')) + "_" + algo;
    mkdir("result", 0755);
    std::ofstream fout("result/" + filename);
    std::ostringstream os;
    os << "date/time: " << ctime(&now);
    os << "algorithm: " << algo << "\n";
    os << "dataset: " << dataset << "\n";
    os << "|R|: " << nr << ", |L|: " << nl << ", |E|: " << ne << "\n";
    os << "grid_size: " << r.workers << ", block_size: " << 64 << "\n";
    if (variance) {
        std::vector<uint64_t> busy(mbe_last_worker_busy(nullptr, 0));
        mbe_last_worker_busy(busy.data(), (uint32_t)busy.size());
        for (uint64_t b : busy) os << b << "\n";
    }
    os << "maximal bicliques: " << r.count << "\n";
    if (sections) {
        double sum = 0;
        for (int i = 0; i < 11; i++) sum += (double)r.sections[i];
        os << "time percentage:" << std::fixed << std::setprecision(4);
        for (int i = 0; i < 11; i++) os << " " << (sum > 0 ? r.sections[i] * 100.0 / sum : 0.0);
        os << std::setprecision(6) << "\n";
        os.unsetf(std::ios::floatfield);
    }
    os << "runtime (s): " << r.kernel_ms / 1000.0 << "\n";
    if (!variance)  // the variance parser takes the last two lines as results
        os << "hash: 0x" << std::hex << std::setw(16) << std::setfill('0') << r.hash << std::dec
           << "\n";
    std::cout << os.str();
    fout << os.str();
    return 0;
}
