// Exercises the library's host C++ (initial partitioner, pairwise FM, perfectly balanced
// post-pass) and the oracle's C++ on random graphs, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitize_host.py. Exits non-zero on a failed
// validity check; the sanitizers abort on a memory or UB error.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <utility>
#include <vector>

#include "fm_refine.hpp"
#include "initial_partition.hpp"
#include "kahip_oracle.h"
#include "rebalance.hpp"

using kahip::HostGraph;

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            std::exit(1);                                               \
        }                                                               \
    } while (0)

// Random graph: G(n, p) plus `hubs` nodes joined to a random third of the others, some isolated
// nodes, optional node and edge weights; symmetric, no self-loops, rows sorted.
static HostGraph random_graph(int n, double p, int hubs, bool weighted, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(0, 1);
    std::set<std::pair<int, int>> edges;
    for (int u = 0; u < n; ++u)
        for (int v = u + 1; v < n; ++v)
            if (U(rng) < p) edges.insert({u, v});
    for (int h = 0; h < hubs && h < n; ++h)
        for (int v = 0; v < n; ++v)
            if (v != h && U(rng) < 0.33) edges.insert({std::min(h, v), std::max(h, v)});
    std::vector<std::vector<std::pair<int, int>>> adj(n);
    for (auto& e : edges) {
        if (e.first % 17 == 5 || e.second % 17 == 5) continue;  // isolated nodes
        const int w = weighted ? 1 + (int)(rng() % 7) : 1;
        adj[e.first].push_back({e.second, w});
        adj[e.second].push_back({e.first, w});
    }
    HostGraph g;
    g.n = n;
    g.xadj.push_back(0);
    for (int v = 0; v < n; ++v) {
        std::sort(adj[v].begin(), adj[v].end());
        for (auto& [t, w] : adj[v]) {
            g.adjncy.push_back(t);
            g.adjwgt.push_back(w);
        }
        g.xadj.push_back((int)g.adjncy.size());
        g.vwgt.push_back(weighted ? 1 + (int)(rng() % 4) : 1);
    }
    return g;
}

int main() {
    int cases = 0;
    for (unsigned seed = 0; seed < 6; ++seed) {
        for (int n : {1, 2, 37, 300, 1200}) {
            const bool weighted = seed % 2;
            HostGraph g = random_graph(n, n > 300 ? 0.004 : 0.03, n > 30 ? 2 : 0, weighted, seed * 7 + n);
            const long long W = g.total_weight();
            const int m2 = g.xadj[n];
            for (int k : {2, 3, 8}) {
                // initial partitioner
                kahip::IpParams ip;
                ip.k = k;
                ip.block_bound = (long long)(1.03 * ((W + k - 1) / k)) + 4;
                ip.repetitions = 2;
                ip.bisection_tries = 4;
                ip.seed = seed;
                ip.threads = 2;
                std::vector<int> part;
                kahip::initial_partition(g, ip, part);
                CHECK((int)part.size() == n);
                for (int v = 0; v < n; ++v) CHECK(part[v] >= 0 && part[v] < k);
                // pairwise FM, then the perfectly balanced post-pass
                kahip::FmRefineParams fp;
                fp.k = k;
                fp.block_bound = ip.block_bound;
                fp.rounds = 3;
                fp.seed = seed;
                kahip::fm_refine_pairs(g, fp, part);
                const long long tight = (W + k - 1) / k + 4;
                const long long cut = kahip::balance_postpass(n, g.xadj.data(), g.adjncy.data(),
                                                              g.vwgt.data(), g.adjwgt.data(), k,
                                                              tight, seed, part.data());
                CHECK(cut == orc_edge_cut(n, g.xadj.data(), g.adjncy.data(), g.adjwgt.data(),
                                          part.data()));
                // oracle: synchronous and asynchronous LP, contraction, refinement, metrics
                const unsigned B = (unsigned)std::max<long long>(2, W / (4 * k));
                std::vector<int> cl(n), cl2(n), moves(10);
                const int nc = orc_lp_cluster_sync(n, g.xadj.data(), g.adjncy.data(), g.vwgt.data(),
                                                   g.adjwgt.data(), B, 10, 4, seed, cl.data(),
                                                   moves.data());
                const int nc2 = orc_lp_cluster_async(n, g.xadj.data(), g.adjncy.data(),
                                                     g.vwgt.data(), g.adjwgt.data(), B, 10, seed,
                                                     cl2.data());
                CHECK(nc >= (n ? 1 : 0) && nc <= n && nc2 <= n);
                std::vector<int> cx(nc + 1), ca(std::max(m2, 1)), cw(std::max(m2, 1)), cv(std::max(nc, 1));
                orc_contract(n, g.xadj.data(), g.adjncy.data(), g.vwgt.data(), g.adjwgt.data(),
                             cl.data(), nc, cx.data(), ca.data(), cw.data(), cv.data());
                CHECK(cx[nc] <= m2);
                std::vector<int> rp = part, rm(25);
                orc_lp_refine_sync(n, g.xadj.data(), g.adjncy.data(), g.vwgt.data(), g.adjwgt.data(),
                                   k, (unsigned)ip.block_bound, 25, 4, seed, rp.data(), rm.data());
                std::vector<int> ra = part;
                orc_lp_refine_async(n, g.xadj.data(), g.adjncy.data(), g.vwgt.data(),
                                    g.adjwgt.data(), k, (unsigned)ip.block_bound, 25, (int)seed,
                                    ra.data());
                std::vector<int64_t> bw(k);
                orc_block_weights(n, g.vwgt.data(), k, rp.data(), bw.data());
                std::vector<int> lb(n, (int)B);
                std::vector<uint8_t> inc(n, 1);
                if (n > 1) inc[1] = 0;
                orc_lp_cluster_sync_ex(n, g.xadj.data(), g.adjncy.data(), g.vwgt.data(),
                                       g.adjwgt.data(), lb.data(), inc.data(), 5, 2, seed,
                                       cl.data(), moves.data());
                ++cases;
            }
        }
    }
    std::printf("host_check ok: %d cases\n", cases);
    return 0;
}
