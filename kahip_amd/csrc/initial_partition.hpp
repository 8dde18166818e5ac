// CPU initial partitioning of the coarsest graph (stays on the host, as in the reference:
// lib/partition/initial_partitioning/initial_partitioning.cpp:39-134).
#pragma once

#include <cstdint>
#include <vector>

namespace kahip {

struct HostGraph {
    int n = 0;
    std::vector<int> xadj, adjncy, adjwgt, vwgt;
    long long total_weight() const;
};

struct IpParams {
    int k = 2;
    long long block_bound = 0;  // U: every final block must weigh at most this
    int repetitions = 2;        // independent recursive-bisection runs; the lowest cut wins
    int bisection_tries = 32;   // BFS-grown starts per bisection, on its coarsest (<= stop_nodes) graph
    int coarsen_factor = 12;    // cluster bound of the bisection coarsening: ceil(max side / 12)
    int stop_nodes = 400;       // bisection coarsening stops below this many nodes
    int fm_stall = 500;         // FM stops after this many moves without improvement, unless:
    int fm_pct_coarsest = 0;    // > 0: limit = max(15, pct % of the boundary nodes), on the
    int fm_pct_levels = 0;      //   coarsest graph / on the uncoarsening levels (two_way_fm.cpp:69-71)
    uint64_t seed = 0;
    int threads = 0;            // 0: hardware concurrency (capped at 16)
};

// Writes part[v] in [0, k). Deterministic for a given seed, independent of the thread count.
void initial_partition(const HostGraph& g, const IpParams& p, std::vector<int>& part);

long long host_edge_cut(const HostGraph& g, const std::vector<int>& part);

}  // namespace kahip
