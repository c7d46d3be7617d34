/*
 * Golden-fixture scan for the batched workload (test infrastructure).
 * For every seed in [first, last): generate the graph with libmbe.so's
 * generator, run the CPU oracle (liboracle.so, serial, RS+ES) in a forked
 * child under a wall-clock limit, and print one JSON line per seed:
 *   {"seed": s, "n_u": .., "n_v": .., "edges": .., "count": .., "hash": "0x..",
 *    "work": .., "seconds": ..}   or   {"seed": s, ..., "timeout": limit}
 * Usage: golden_scan <first> <last> <limit_s> <jobs>
 * Build: see tests/golden/Makefile.
 */
#define _GNU_SOURCE
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/wait.h>
#include <unistd.h>

#include "../../include/mbe.h"
#include "../../oracle/mbe_oracle.h"

static void run_child(uint64_t seed, int fd, unsigned limit) {
    alarm(limit);
    mbe_graph* g = NULL;
    char buf[512];
    if (mbe_gen_batch_graph(seed, &g) != MBE_OK) _exit(2);
    uint32_t nu, nv;
    uint64_t ne;
    const uint32_t *off, *adj;
    mbe_graph_info(g, &nu, &nv, &ne);
    mbe_graph_csr(g, &off, &adj);
    int n = snprintf(buf, sizeof buf, "\"n_u\": %u, \"n_v\": %u, \"edges\": %llu", nu, nv,
                     (unsigned long long)ne);
    if (write(fd, buf, n) != n) _exit(3);
    mbe_oracle_result r;
    if (mbe_oracle_solve(nu, nv, off, adj, 1, 1, -1, 1, &r) != 0) _exit(4);
    n = snprintf(buf, sizeof buf,
                 ", \"count\": %llu, \"hash\": \"0x%016llx\", \"work\": %llu, \"seconds\": %.4f",
                 (unsigned long long)r.count, (unsigned long long)r.hash,
                 (unsigned long long)r.work, r.seconds);
    if (write(fd, buf, n) != n) _exit(3);
    _exit(0);
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s first last limit_s jobs\n", argv[0]);
        return 2;
    }
    uint64_t first = strtoull(argv[1], 0, 10), last = strtoull(argv[2], 0, 10);
    unsigned limit = (unsigned)atoi(argv[3]);
    int jobs = atoi(argv[4]);
    struct { pid_t pid; uint64_t seed; int fd; } run[256];
    int nrun = 0;
    uint64_t next = first;
    while (next < last || nrun > 0) {
        while (nrun < jobs && next < last) {
            int p[2];
            if (pipe(p)) return 1;
            pid_t pid = fork();
            if (pid == 0) {
                close(p[0]);
                run_child(next, p[1], limit);
            }
            close(p[1]);
            run[nrun].pid = pid;
            run[nrun].seed = next++;
            run[nrun].fd = p[0];
            nrun++;
        }
        int status;
        pid_t done = wait(&status);
        for (int i = 0; i < nrun; i++) {
            if (run[i].pid != done) continue;
            char out[1024];
            ssize_t k = read(run[i].fd, out, sizeof out - 1);
            out[k > 0 ? k : 0] = 0;
            close(run[i].fd);
            if (WIFEXITED(status) && WEXITSTATUS(status) == 0)
                printf("{\"seed\": %llu, %s}\n", (unsigned long long)run[i].seed, out);
            else
                printf("{\"seed\": %llu, %s%s\"timeout\": %u}\n", (unsigned long long)run[i].seed,
                       out, k > 0 ? ", " : "", limit);
            fflush(stdout);
            run[i] = run[--nrun];
            break;
        }
    }
    return 0;
}
