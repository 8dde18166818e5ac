// CPU post-pass of the perfectly balanced modes: imbalance 0 in kaffpa / kaffpa_balance_NE and
// perfectly_balance in kaffpa_balance (interface/kaHIP_interface.cpp:123-132 runs
// cycle_refinement there, which is outside this path).

#include "rebalance.hpp"

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "fm_refine.hpp"

namespace kahip {

// CPU rebalancing to a block bound, the post-pass of the perfectly balanced modes. The reference
// runs cycle_refinement there (interface/kaHIP_interface.cpp:123-132, out of this path's scope);
// this stand-in moves single nodes out of overloaded blocks in order of cut change, with the
// gains kept current: a max-heap holds, for nodes of overloaded blocks, (gain, node, target) with
// target the adjacent block of highest connection that has room (else the lightest block with
// room) and gain = conn(target) - conn(own); a popped entry whose node moved, whose block is no
// longer overloaded or whose target filled up is recomputed and pushed again. After a move the
// moved node's neighbours in overloaded blocks are re-evaluated, so the moves grow from the
// boundary inwards instead of cutting isolated nodes out. Ties: higher gain, then lower node
// id, then lower target. Stops when no block exceeds the bound or no node of an overloaded block
// has a target with room. Returns the cut of the result.
long long rebalance(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, long long bound, int* part) {
    std::vector<long long> W(k, 0);
    auto vw = [&](int v) -> long long { return vwgt ? vwgt[v] : 1; };
    auto ew = [&](long long e) -> long long { return adjwgt ? adjwgt[e] : 1; };
    for (int v = 0; v < n; ++v) W[part[v]] += vw(v);
    std::vector<long long> conn(k, 0);
    std::vector<int> touched;
    // lightest block, kept with a lazy min-heap of (weight, block)
    using WB = std::pair<long long, int>;
    std::priority_queue<WB, std::vector<WB>, std::greater<WB>> light;
    for (int b = 0; b < k; ++b) light.push({W[b], b});
    auto lightest = [&]() {
        while (light.top().first != W[light.top().second]) light.pop();
        return light.top().second;
    };
    // best (gain, target) of v, or target -1
    auto best_move = [&](int v, long long& gain) -> int {
        const int own = part[v];
        for (long long e = xadj[v]; e < xadj[v + 1]; ++e) {
            const int b = part[adjncy[e]];
            if (conn[b] == 0) touched.push_back(b);
            conn[b] += ew(e);
        }
        int t = -1;
        long long tc = -1;
        for (int b : touched)
            if (b != own && W[b] + vw(v) <= bound && (conn[b] > tc || (conn[b] == tc && b < t))) {
                t = b;
                tc = conn[b];
            }
        if (t < 0) {
            const int l = lightest();
            if (l != own && W[l] + vw(v) <= bound) {
                t = l;
                tc = conn[l];
            }
        }
        gain = t >= 0 ? tc - conn[own] : 0;
        for (int b : touched) conn[b] = 0;
        touched.clear();
        return t;
    };
    struct Entry {
        long long gain;
        int v, t;
        bool operator<(const Entry& o) const {  // max-heap order
            if (gain != o.gain) return gain < o.gain;
            if (v != o.v) return v > o.v;
            return t > o.t;
        }
    };
    std::priority_queue<Entry> heap;
    std::vector<long long> stamp(n, -1);  // gain of v's current heap entry (-1: none valid)
    std::vector<int> stamp_t(n, -1);
    auto consider = [&](int v) {
        if (W[part[v]] <= bound) return;
        long long g = 0;
        const int t = best_move(v, g);
        if (t < 0) return;
        if (stamp_t[v] == t && stamp[v] == g) return;  // the same entry is already queued
        stamp[v] = g;
        stamp_t[v] = t;
        heap.push(Entry{g, v, t});
    };
    for (int v = 0; v < n; ++v) consider(v);
    while (!heap.empty()) {
        const Entry top = heap.top();
        heap.pop();
        const int v = top.v;
        if (stamp_t[v] != top.t || stamp[v] != top.gain) continue;  // superseded
        stamp_t[v] = -1;
        const int own = part[v];
        if (W[own] <= bound) continue;
        long long g = 0;
        const int t = best_move(v, g);
        if (t < 0) continue;
        if (t != top.t || g != top.gain) {  // stale: queue the current move instead
            stamp[v] = g;
            stamp_t[v] = t;
            heap.push(Entry{g, v, t});
            continue;
        }
        W[own] -= vw(v);
        W[t] += vw(v);
        light.push({W[own], own});
        light.push({W[t], t});
        part[v] = t;
        for (long long e = xadj[v]; e < xadj[v + 1]; ++e) consider(adjncy[e]);
    }
    long long cut = 0;
    for (int v = 0; v < n; ++v)
        for (long long e = xadj[v]; e < xadj[v + 1]; ++e)
            if (part[v] != part[adjncy[e]]) cut += ew(e);
    return cut / 2;
}


// The perfectly balanced post-pass: rebalance() to `bound`, then 2-way FM over the adjacent
// block pairs under the same bound, allowed to exceed it by one node weight within a pass so
// that balanced pairs can exchange nodes (only prefixes as balanced as the pass start are kept).
// Returns the cut.
long long balance_postpass(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, int k, long long bound, uint64_t seed, int* part) {
    long long cut = rebalance(n, xadj, adjncy, vwgt, adjwgt, k, bound, part);
    if (n == 0 || k < 2) return cut;
    HostGraph g;
    g.n = n;
    g.xadj.assign(xadj, xadj + n + 1);
    g.adjncy.assign(adjncy, adjncy + xadj[n]);
    g.adjwgt.resize(xadj[n]);
    g.vwgt.resize(n);
    long long wmax = 1;
    for (int e = 0; e < xadj[n]; ++e) g.adjwgt[e] = adjwgt ? adjwgt[e] : 1;
    for (int v = 0; v < n; ++v) {
        g.vwgt[v] = vwgt ? vwgt[v] : 1;
        wmax = std::max<long long>(wmax, g.vwgt[v]);
    }
    std::vector<int> pv(part, part + n);
    FmRefineParams fp;
    fp.k = k;
    fp.block_bound = bound;
    fp.rounds = 10;
    fp.seed = seed;
    fp.pass_slack = wmax;
    cut -= fm_refine_pairs(g, fp, pv);
    std::copy(pv.begin(), pv.end(), part);
    return cut;
}

}  // namespace kahip
