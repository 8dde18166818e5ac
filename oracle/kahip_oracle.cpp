// The oracle: CPU restatement of KaHIP's LP coarsening/refinement path (R1) and the synchronous
// schedule the GPU runs (R2). TEST INFRASTRUCTURE ONLY: loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker; never by the shipped
// library. Every R1 function cites the reference lines it restates.
//
// Parity status (how each function is pinned, see tests/test_oracle*.py):
//   orc_lp_cluster_async      bit-exact vs the reference build on examples/ graphs, many seeds
//   orc_remap_first_appearance closed form + brute force + reference build
//   orc_contract              bit-exact vs the reference build; brute force; invariants
//   orc_lp_refine_async       bit-exact vs the reference build
//   orc_edge_cut / orc_balance / orc_block_weights  vs the reference build; brute force
//   orc_upper_bound / orc_cluster_bound  vs the reference build; table of the documented configs
//   orc_stop_rule_continue    closed form (stop_rules.h:108-111)
//   orc_project               closed form
//   orc_order_nodes_degree    indirectly, through orc_lp_cluster_async's bit-exactness
//   orc_lp_cluster_sync / orc_lp_refine_sync  (R2, no reference counterpart) pinned by
//       invariants and special cases: every cluster weight < B, labels valid, edgeless graph
//       keeps singletons, B <= min weight + 1 forbids merges, complete graph collapses,
//       dumbbell splits into its two cliques, refinement never raises the cut or breaks U.

#include "kahip_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <random>
#include <vector>

namespace {

inline int wv(const int* vwgt, int v) { return vwgt ? vwgt[v] : 1; }
inline int we(const int* adjwgt, int e) { return adjwgt ? adjwgt[e] : 1; }

// lib/tools/random_functions.h:152-155 — nextBool = uniform_int_distribution<unsigned>(0,1) on the
// process-global std::mt19937 seeded with setSeed(seed) (:173-177). The oracle holds its own
// engine so calls are reentrant; the stream of draws is the same.
struct RefRng {
    std::mt19937 mt;
    explicit RefRng(int seed) { mt.seed(seed); }
    bool next_bool() {
        std::uniform_int_distribution<unsigned int> d(0, 1);
        return (bool)d(mt);
    }
};

}  // namespace

extern "C" {

// lib/partition/coarsening/clustering/node_ordering.h:43-48: identity permutation sorted by
// ascending degree with the (unstable) std::sort of libstdc++.
void orc_order_nodes_degree(int n, const int* xadj, int* perm) {
    std::vector<unsigned> p(n);
    for (int i = 0; i < n; ++i) p[i] = i;
    std::sort(p.begin(), p.end(), [&](const unsigned& l, const unsigned& r) -> bool {
        int dl = xadj[l + 1] - xadj[l];
        int dr = xadj[r + 1] - xadj[r];
        return dl < dr;
    });
    for (int i = 0; i < n; ++i) perm[i] = (int)p[i];
}

// size_constraint_label_propagation.cpp:207-232: clusters renumbered 0..C-1 in order of first
// appearance scanning v = 0..n-1. Returns C.
int orc_remap_first_appearance(int n, int* cluster) {
    std::vector<int> remap(n > 0 ? n : 1, -1);
    int next = 0;
    for (int v = 0; v < n; ++v) {
        int c = cluster[v];
        if (remap[c] < 0) remap[c] = next++;
        cluster[v] = remap[c];
    }
    return next;
}

// size_constraint_label_propagation.cpp:134-194 (sweep loop) followed by remap :207-232.
int orc_lp_cluster_async(int n, const int* xadj, const int* adjncy, const int* vwgt,
                         const int* adjwgt, unsigned bound, int iterations, int seed,
                         int* cluster_out) {
    RefRng rng(seed);
    std::vector<unsigned> hash_map(n, 0), cid(n), size(n);
    std::vector<int> perm(n);
    for (int v = 0; v < n; ++v) {
        size[v] = (unsigned)wv(vwgt, v);
        cid[v] = v;
    }
    orc_order_nodes_degree(n, xadj, perm.data());
    for (int j = 0; j < iterations; ++j) {
        for (int i = 0; i < n; ++i) {
            int node = perm[i];
            for (int e = xadj[node]; e < xadj[node + 1]; ++e)
                hash_map[cid[adjncy[e]]] += (unsigned)we(adjwgt, e);
            unsigned max_block = cid[node], my_block = cid[node], max_value = 0;
            unsigned w = (unsigned)wv(vwgt, node);
            for (int e = xadj[node]; e < xadj[node + 1]; ++e) {
                unsigned cur_block = cid[adjncy[e]];
                unsigned cur_value = hash_map[cur_block];
                // :174-178 — the coin is drawn only when the values tie (short circuit), and
                // before the size test.
                if ((cur_value > max_value || (cur_value == max_value && rng.next_bool())) &&
                    (size[cur_block] + w < bound || cur_block == my_block)) {
                    max_value = cur_value;
                    max_block = cur_block;
                }
                hash_map[cur_block] = 0;
            }
            size[cid[node]] -= w;
            size[max_block] += w;
            cid[node] = max_block;
        }
    }
    for (int v = 0; v < n; ++v) cluster_out[v] = (int)cid[v];
    return orc_remap_first_appearance(n, cluster_out);
}

// complete_boundary.h:511-562 as called by contraction.cpp:94-133. Coarse node p's weight is the
// sum of its members' weights; its edges are emitted in order of first appearance when scanning
// members in ascending id and their edges in CSR order; intra-cluster edges are dropped; weights
// of parallel edges are summed. Returns the number of coarse nodes (nc).
int orc_contract(int n, const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                 const int* cluster, int nc, int* cxadj, int* cadjncy, int* cadjwgt, int* cvwgt) {
    std::vector<std::vector<int>> members(nc);
    std::vector<int64_t> bw(nc, 0);
    for (int v = 0; v < n; ++v) {
        members[cluster[v]].push_back(v);
        bw[cluster[v]] += wv(vwgt, v);
    }
    std::vector<int64_t> tw(nc, 0);
    std::vector<int> list;
    int pos = 0;
    for (int p = 0; p < nc; ++p) {
        cxadj[p] = pos;
        cvwgt[p] = (int)bw[p];
        for (int v : members[p]) {
            for (int e = xadj[v]; e < xadj[v + 1]; ++e) {
                int q = cluster[adjncy[e]];
                if (q == p) continue;
                if (tw[q] == 0) list.push_back(q);
                tw[q] += we(adjwgt, e);
            }
        }
        for (int q : list) {
            cadjncy[pos] = q;
            cadjwgt[pos] = (int)tw[q];
            tw[q] = 0;
            ++pos;
        }
        list.clear();
    }
    cxadj[nc] = pos;
    return nc;
}

// label_propagation_refinement.cpp:21-155: block weights, FIFO queue seeded in degree order,
// up to `rounds` rounds, argmax with coin-flip ties under the block bound, neighbours of moved
// nodes queued once for the next round.
void orc_lp_refine_async(int n, const int* xadj, const int* adjncy, const int* vwgt,
                         const int* adjwgt, int k, unsigned bound, int rounds, int seed,
                         int* part) {
    RefRng rng(seed);
    std::vector<unsigned> hash_map(k, 0), size(k, 0);
    std::vector<int> perm(n);
    orc_order_nodes_degree(n, xadj, perm.data());
    std::queue<int> q, nq;
    std::vector<char> in_q(n, 0), in_nq(n, 0);
    for (int v = 0; v < n; ++v) {
        size[part[v]] += (unsigned)wv(vwgt, v);
        q.push(perm[v]);
    }
    for (int j = 0; j < rounds; ++j) {
        while (!q.empty()) {
            int node = q.front();
            q.pop();
            in_q[node] = 0;
            for (int e = xadj[node]; e < xadj[node + 1]; ++e)
                hash_map[part[adjncy[e]]] += (unsigned)we(adjwgt, e);
            unsigned max_block = part[node], my_block = part[node], max_value = 0;
            unsigned w = (unsigned)wv(vwgt, node);
            for (int e = xadj[node]; e < xadj[node + 1]; ++e) {
                unsigned cur_block = part[adjncy[e]];
                unsigned cur_value = hash_map[cur_block];
                if ((cur_value > max_value || (cur_value == max_value && rng.next_bool())) &&
                    (size[cur_block] + w < bound ||
                     (cur_block == my_block && size[my_block] <= bound))) {
                    max_value = cur_value;
                    max_block = cur_block;
                }
                hash_map[cur_block] = 0;
            }
            size[part[node]] -= w;
            size[max_block] += w;
            bool changed = (unsigned)part[node] != max_block;
            part[node] = (int)max_block;
            if (changed) {
                for (int e = xadj[node]; e < xadj[node + 1]; ++e) {
                    int t = adjncy[e];
                    if (!in_nq[t]) {
                        nq.push(t);
                        in_nq[t] = 1;
                    }
                }
            }
        }
        std::swap(q, nq);
        std::swap(in_q, in_nq);
    }
}

// quality_metrics.cpp:23-37: sum of weights of directed edges whose endpoints lie in different
// blocks, halved. Accumulated in 64 bits (the reference's int accumulator agrees whenever it does
// not overflow).
int64_t orc_edge_cut(int n, const int* xadj, const int* adjncy, const int* adjwgt,
                     const int* part) {
    int64_t cut = 0;
    for (int v = 0; v < n; ++v)
        for (int e = xadj[v]; e < xadj[v + 1]; ++e)
            if (part[v] != part[adjncy[e]]) cut += we(adjwgt, e);
    return cut / 2;
}

void orc_block_weights(int n, const int* vwgt, int k, const int* part, int64_t* weights) {
    for (int b = 0; b < k; ++b) weights[b] = 0;
    for (int v = 0; v < n; ++v) weights[part[v]] += wv(vwgt, v);
}

// quality_metrics.cpp:265-288: max block weight / ceil(W / k), in double.
double orc_balance(int n, const int* vwgt, int k, const int* part) {
    std::vector<int64_t> w(k);
    orc_block_weights(n, vwgt, k, part, w.data());
    double total = 0;
    for (int v = 0; v < n; ++v) total += wv(vwgt, v);
    double avg = std::ceil(total / (double)k);
    double mx = -1;
    for (int b = 0; b < k; ++b)
        if ((double)w[b] > mx) mx = (double)w[b];
    return mx / avg;
}

// app/balance_configuration.h:40-46 (imbalance != 0, balance_edges off):
// U = (1 + eps) * ceil(W / k), truncated to the unsigned NodeWeight. The interface passes the
// fraction through `imbalance = 100 * fraction` then `epsilon = imbalance / 100.0`
// (interface/kaHIP_interface.cpp:115), which this reproduces in the same double operations.
unsigned orc_upper_bound(int64_t total_weight, int k, double imbalance) {
    // app/balance_configuration.h:39-46: imbalance 0 is the perfectly balanced mode, whose
    // partitioning bound has 1 % of slack
    double imb_percent = 100 * imbalance;
    double epsilon = imb_percent / 100.0;
    unsigned load = (unsigned)total_weight;
    if (imb_percent == 0) return (unsigned)((1 + epsilon + 0.01) * std::ceil(load / (double)k));
    return (unsigned)((1 + epsilon) * std::ceil(load / (double)k));
}

// size_constraint_label_propagation.cpp:57: B = ceil(U / cluster_coarsening_factor).
unsigned orc_cluster_bound(unsigned upper_bound, int factor) {
    return (unsigned)std::ceil(upper_bound / (double)factor);
}

// stop_rules.h:108-111 (multiple_k_stop_rule): keep coarsening while the contraction rate is at
// least 1.1 and the coarser graph still has at least stop_factor * k nodes.
int orc_stop_rule_continue(int n_finer, int n_coarser, int k, int stop_factor) {
    double rate = 1.0 * (unsigned)n_finer / (double)(unsigned)n_coarser;
    unsigned num_stop = (unsigned)stop_factor * (unsigned)k;
    return rate >= 1.1 && (unsigned)n_coarser >= num_stop;
}

// graph_hierarchy.cpp:54-58.
void orc_project(int n_fine, const int* mapping, const int* coarse_part, int* fine_part) {
    for (int v = 0; v < n_fine; ++v) fine_part[v] = coarse_part[mapping[v]];
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// R2: the synchronous schedule. Definitions shared with the GPU kernels:
//   mix32          Wellons' "lowbias32" integer finaliser.
//   hash3(s,a,b)   mix32(mix32(mix32(s ^ 0x85ebca6b) ^ a) ^ b)
//   subround(v)    hash3(seed ^ 0x68bc21eb, 0, v) % S          (fixed for one call)
//   tie(v,c,r)     hash3(seed ^ 0x2f6b7d93 ^ (r * 0x9e3779b9), v, c)
// Candidate order: c beats d when conn(c) > conn(d), or conn equal and tie(c) > tie(d), or both
// equal and c < d. Proposal priority: (gain desc, v asc).
// Admission into a label t within one sub-round: let inc(t) be the summed weight of the
// proposals into t. If size(t) + inc(t) < bound every proposal is admitted; otherwise only the
// highest-priority proposal is, and only if size(t) + w < bound.
// ---------------------------------------------------------------------------------------------

static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

extern "C" uint32_t orc_hash3(uint32_t s, uint32_t a, uint32_t b) {
    return mix32(mix32(mix32(s ^ 0x85ebca6bU) ^ a) ^ b);
}

namespace {

constexpr uint32_t kSubSalt = 0x68bc21ebU;
constexpr uint32_t kTieSalt = 0x2f6b7d93U;
constexpr uint32_t kGolden = 0x9e3779b9U;

struct Proposal {
    int v;
    int from;
    int to;
    int64_t gain;
    int w;
};

inline bool higher_priority(const Proposal& a, const Proposal& b) {
    return a.gain > b.gain || (a.gain == b.gain && a.v < b.v);
}

inline int subround_of(uint32_t seed, int v, int S) {
    return (int)(orc_hash3(seed ^ kSubSalt, 0u, (uint32_t)v) % (uint32_t)S);
}

// Best candidate over the distinct labels of v's neighbours other than own that pass
// `eligible`; -1 when none. conn_own receives the connection to the own label.
template <class Eligible>
int best_label(int v, const int* xadj, const int* adjncy, const int* adjwgt, const int* label,
               int own, uint32_t tie_seed, std::vector<int64_t>& acc, std::vector<int>& touched,
               int64_t* conn_own, int64_t* conn_best, Eligible eligible) {
    for (int e = xadj[v]; e < xadj[v + 1]; ++e) {
        int c = label[adjncy[e]];
        if (acc[c] == 0) touched.push_back(c);
        acc[c] += we(adjwgt, e);
    }
    *conn_own = acc[own];
    int best = -1;
    int64_t bconn = 0;
    uint32_t btie = 0;
    for (int c : touched) {
        if (c == own || !eligible(c)) continue;
        int64_t cc = acc[c];
        uint32_t t = orc_hash3(tie_seed, (uint32_t)v, (uint32_t)c);
        if (best < 0 || cc > bconn || (cc == bconn && (t > btie || (t == btie && c < best)))) {
            best = c;
            bconn = cc;
            btie = t;
        }
    }
    for (int c : touched) acc[c] = 0;
    touched.clear();
    *conn_best = bconn;
    return best;
}

// The admission rule above, applied to every target group of P; `size` is the snapshot and
// bound(t) the bound of label t.
template <class Bound>
void admit(const std::vector<Proposal>& P, const std::vector<int64_t>& size, Bound bound,
           std::vector<Proposal>& out, std::vector<int64_t>& inc, std::vector<int>& top) {
    for (size_t i = 0; i < P.size(); ++i) {
        const Proposal& p = P[i];
        inc[p.to] += p.w;
        if (top[p.to] < 0 || higher_priority(p, P[top[p.to]])) top[p.to] = (int)i;
    }
    for (size_t i = 0; i < P.size(); ++i) {
        const Proposal& p = P[i];
        const int64_t b = bound(p.to);
        if (size[p.to] + inc[p.to] < b || (top[p.to] == (int)i && size[p.to] + p.w < b))
            out.push_back(p);
    }
    for (const Proposal& p : P) {
        inc[p.to] = 0;
        top[p.to] = -1;
    }
}

}  // namespace

extern "C" {

// Synchronous size-constrained LP clustering (R2). `rounds` rounds of `subrounds` sub-rounds; in
// sub-round s the nodes with subround(v) == s propose against the labels and cluster weights as
// they stand at the start of the sub-round. v proposes its best neighbour label c with
// size(c) + w(v) < bound (the weight test of size_constraint_label_propagation.cpp:175) when
// conn(c) > conn(own), or conn(c) == conn(own) and size(c) >= size(own). Admitted moves are
// applied together; labels are then renumbered in first-appearance order. Returns the number of
// clusters.
//
// The _ex form takes the bound per label (lbound[c], labels being node ids; used by the batched
// mode, where each graph of a disjoint union has its own bound) and an optional inclusion mask:
// nodes with include[v] == 0 never propose (they keep their own label).
int orc_lp_cluster_sync_ex(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, const int* lbound, const uint8_t* include,
                           int rounds, int subrounds, uint32_t seed, int* cluster_out,
                           int* moves_per_round) {
    std::vector<int> label(n), sub(n), top(n, -1);
    std::vector<int64_t> size(n), acc(n, 0), inc(n, 0);
    std::vector<int> touched;
    for (int v = 0; v < n; ++v) {
        label[v] = v;
        size[v] = wv(vwgt, v);
        sub[v] = subround_of(seed, v, subrounds);
    }
    std::vector<Proposal> P, A;
    auto bound = [&](int c) -> int64_t { return (int64_t)lbound[c]; };
    for (int r = 0; r < rounds; ++r) {
        int moves = 0;
        uint32_t tie_seed = seed ^ kTieSalt ^ ((uint32_t)r * kGolden);
        for (int s = 0; s < subrounds; ++s) {
            P.clear();
            A.clear();
            for (int v = 0; v < n; ++v) {
                if (sub[v] != s || (include && !include[v])) continue;
                int own = label[v];
                int w = wv(vwgt, v);
                int64_t conn_own, conn_best;
                int best = best_label(v, xadj, adjncy, adjwgt, label.data(), own, tie_seed, acc,
                                      touched, &conn_own, &conn_best,
                                      [&](int c) { return size[c] + w < bound(c); });
                if (best >= 0 && (conn_best > conn_own ||
                                  (conn_best == conn_own && size[best] >= size[own])))
                    P.push_back({v, own, best, conn_best - conn_own, w});
            }
            admit(P, size, bound, A, inc, top);
            for (const Proposal& p : A) {
                size[p.from] -= p.w;
                size[p.to] += p.w;
                label[p.v] = p.to;
            }
            moves += (int)A.size();
        }
        if (moves_per_round) moves_per_round[r] = moves;
    }
    for (int v = 0; v < n; ++v) cluster_out[v] = label[v];
    return orc_remap_first_appearance(n, cluster_out);
}

int orc_lp_cluster_sync(int n, const int* xadj, const int* adjncy, const int* vwgt,
                        const int* adjwgt, unsigned bound, int rounds, int subrounds,
                        uint32_t seed, int* cluster_out, int* moves_per_round) {
    std::vector<int> lb(n, (int)std::min<unsigned>(bound, 0x7fffffffu));
    return orc_lp_cluster_sync_ex(n, xadj, adjncy, vwgt, adjwgt, lb.data(), nullptr, rounds,
                                  subrounds, seed, cluster_out, moves_per_round);
}

// Synchronous LP refinement (R2). Block weights W. Round 0 has every node active; later rounds
// the neighbours of the nodes moved in the previous round (label_propagation_refinement.cpp
// :84-92). In sub-round s an active node v with subround(v) == s proposes its best neighbouring
// block b with W(b) + w(v) < bound (:68) when conn(b) > conn(own), when conn(b) == conn(own) and
// W(b) + 2 w(v) <= W(own), or whenever its own block is overloaded (W(own) > bound, which makes
// own ineligible at :68). Departures from an overloaded block b all pass when W(b) minus their
// summed weight stays >= bound; otherwise only the highest-priority departure passes. The
// surviving proposals are admitted as in the clustering and applied together.
//
// The _ex form takes the bound per block (lbound[b], b < k) and an inclusion mask as above.
void orc_lp_refine_sync_ex(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, int k, const int* lbound, const uint8_t* include,
                           int rounds, int subrounds, uint32_t seed, int* part,
                           int* moves_per_round) {
    std::vector<int64_t> W(k, 0), acc(k, 0), inc(k, 0), out(k, 0);
    std::vector<int> touched, top(k, -1), dtop(k, -1), sub(n);
    for (int v = 0; v < n; ++v) {
        W[part[v]] += wv(vwgt, v);
        sub[v] = subround_of(seed, v, subrounds);
    }
    std::vector<char> active(n, 1), next(n, 0);
    std::vector<Proposal> P, D, A;
    auto U = [&](int b) -> int64_t { return (int64_t)lbound[b]; };
    for (int r = 0; r < rounds; ++r) {
        int moves = 0;
        uint32_t tie_seed = seed ^ kTieSalt ^ ((uint32_t)r * kGolden);
        std::fill(next.begin(), next.end(), 0);
        for (int s = 0; s < subrounds; ++s) {
            P.clear();
            D.clear();
            A.clear();
            for (int v = 0; v < n; ++v) {
                if (!active[v] || sub[v] != s || (include && !include[v])) continue;
                int own = part[v];
                int w = wv(vwgt, v);
                int64_t conn_own, conn_best;
                int best = best_label(v, xadj, adjncy, adjwgt, part, own, tie_seed, acc, touched,
                                      &conn_own, &conn_best,
                                      [&](int b) { return W[b] + w < U(b); });
                bool overloaded = W[own] > U(own);
                if (best >= 0 && (conn_best > conn_own || overloaded ||
                                  (conn_best == conn_own && W[best] + 2 * (int64_t)w <= W[own])))
                    P.push_back({v, own, best, conn_best - conn_own, w});
            }
            for (size_t i = 0; i < P.size(); ++i) {
                const Proposal& p = P[i];
                if (W[p.from] <= U(p.from)) continue;
                out[p.from] += p.w;
                if (dtop[p.from] < 0 || higher_priority(p, P[dtop[p.from]])) dtop[p.from] = (int)i;
            }
            for (size_t i = 0; i < P.size(); ++i) {
                const Proposal& p = P[i];
                if (W[p.from] <= U(p.from) || W[p.from] - out[p.from] >= U(p.from) ||
                    dtop[p.from] == (int)i)
                    D.push_back(p);
            }
            for (const Proposal& p : P) {
                out[p.from] = 0;
                dtop[p.from] = -1;
            }
            admit(D, W, U, A, inc, top);
            for (const Proposal& p : A) {
                W[p.from] -= p.w;
                W[p.to] += p.w;
                part[p.v] = p.to;
                for (int e = xadj[p.v]; e < xadj[p.v + 1]; ++e) next[adjncy[e]] = 1;
            }
            moves += (int)A.size();
        }
        if (moves_per_round) moves_per_round[r] = moves;
        std::swap(active, next);
    }
}

void orc_lp_refine_sync(int n, const int* xadj, const int* adjncy, const int* vwgt,
                        const int* adjwgt, int k, unsigned bound, int rounds, int subrounds,
                        uint32_t seed, int* part, int* moves_per_round) {
    std::vector<int> lb(k, (int)std::min<unsigned>(bound, 0x7fffffffu));
    orc_lp_refine_sync_ex(n, xadj, adjncy, vwgt, adjwgt, k, lb.data(), nullptr, rounds,
                          subrounds, seed, part, moves_per_round);
}

}  // extern "C"
