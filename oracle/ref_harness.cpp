// Test infrastructure only: plain-C bindings over the reference KaHIP classes, linked with the
// reference sources (compiled in place by oracle/Makefile) into oracle/_ref/libkahip_ref.so.
// Used by tests/ to pin the oracle restatement against the reference and by bench.py as the
// reference CPU baseline. Nothing in the shipped library links or loads this file.
//
// Every entry point sets up the same PartitionConfig the reference itself uses for the call it
// wraps and seeds the reference's process-global RNG (lib/tools/random_functions.h:173-177).

#include <chrono>
#include <cstring>
#include <vector>

#include "balance_configuration.h"
#include "configuration.h"
#include "data_structure/graph_access.h"
#include "kaHIP_interface.h"
#include "data_structure/graph_hierarchy.h"
#include "partition/coarsening/clustering/size_constraint_label_propagation.h"
#include "partition/coarsening/coarsening.h"
#include "partition/coarsening/contraction.h"
#include "partition/graph_partitioner.h"
#include "partition/initial_partitioning/initial_partitioning.h"
#include "partition/uncoarsening/refinement/label_propagation_refinement/label_propagation_refinement.h"
#include "partition/uncoarsening/refinement/quotient_graph_refinement/complete_boundary.h"
#include "partition/uncoarsening/uncoarsening.h"
#include "tools/quality_metrics.h"
#include "tools/random_functions.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

// interface/kaHIP_interface.cpp:66-93 (internal_build_graph) without the balance step.
void build_graph(graph_access& G, int n, const int* xadj, const int* adjncy, const int* vwgt,
                 const int* adjwgt) {
    G.build_from_metis(n, const_cast<int*>(xadj), const_cast<int*>(adjncy));
    if (vwgt) {
        forall_nodes(G, v) { G.setNodeWeight(v, vwgt[v]); } endfor
    }
    if (adjwgt) {
        forall_edges(G, e) { G.setEdgeWeight(e, adjwgt[e]); } endfor
    }
}

void fastsocial_config(PartitionConfig& cfg, int k) {
    configuration c;
    cfg.k = k;
    c.fastsocial(cfg);  // app/configuration.h:565-571
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

// size_constraint_label_propagation::label_propagation with an explicit cluster bound B
// (size_constraint_label_propagation.cpp:134-194 + remap :207-232), as app/label_propagation.cpp
// drives it. Returns the number of clusters.
REF_API int ref_lp_cluster(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, unsigned bound, int iterations, int seed,
                           int* cluster_out) {
    graph_access G;
    build_graph(G, n, xadj, adjncy, vwgt, adjwgt);
    PartitionConfig cfg;
    fastsocial_config(cfg, 1);
    cfg.label_iterations = iterations;
    cfg.graph_allready_partitioned = false;
    cfg.combine = false;
    srand(seed);
    random_functions::setSeed(seed);
    std::vector<NodeWeight> cid;
    NodeID blocks = 0;
    size_constraint_label_propagation sclp;
    NodeWeight b = bound;
    sclp.label_propagation(cfg, G, b, cid, blocks);
    for (int v = 0; v < n; ++v) cluster_out[v] = (int)cid[v];
    return (int)blocks;
}

// contraction::contract_clustering -> complete_boundary::fastComputeQuotientGraph
// (contraction.cpp:94-133, complete_boundary.h:511-562). Output arrays must hold n+1 offsets,
// xadj[n] edges and n node weights. Returns the number of coarse nodes.
REF_API int ref_contract(int n, const int* xadj, const int* adjncy, const int* vwgt,
                         const int* adjwgt, const int* cluster, int nc, int* cxadj, int* cadjncy,
                         int* cadjwgt, int* cvwgt) {
    graph_access G;
    build_graph(G, n, xadj, adjncy, vwgt, adjwgt);
    PartitionConfig cfg;
    fastsocial_config(cfg, 2);
    cfg.combine = false;
    CoarseMapping map(n);
    for (int v = 0; v < n; ++v) map[v] = cluster[v];
    Matching matching;
    NodePermutationMap perm;
    graph_access coarse;
    contraction c;
    NodeID ncv = nc;
    c.contract_clustering(cfg, G, coarse, matching, map, ncv, perm);
    int cn = coarse.number_of_nodes();
    int pos = 0;
    for (int p = 0; p < cn; ++p) {
        cxadj[p] = pos;
        cvwgt[p] = coarse.getNodeWeight(p);
        forall_out_edges(coarse, e, p) {
            cadjncy[pos] = coarse.getEdgeTarget(e);
            cadjwgt[pos] = coarse.getEdgeWeight(e);
            ++pos;
        } endfor
    }
    cxadj[cn] = pos;
    return cn;
}

// label_propagation_refinement::perform_refinement (label_propagation_refinement.cpp:21-155) with
// block bound U and the given number of rounds; part is read and overwritten.
REF_API void ref_lp_refine(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, int k, unsigned bound, int rounds, int seed,
                           int* part) {
    graph_access G;
    build_graph(G, n, xadj, adjncy, vwgt, adjwgt);
    G.set_partition_count(k);
    for (int v = 0; v < n; ++v) G.setPartitionIndex(v, part[v]);
    PartitionConfig cfg;
    fastsocial_config(cfg, k);
    cfg.upper_bound_partition = bound;
    cfg.label_iterations_refinement = rounds;
    srand(seed);
    random_functions::setSeed(seed);
    complete_boundary bnd(&G);
    label_propagation_refinement r;
    r.perform_refinement(cfg, G, bnd);
    for (int v = 0; v < n; ++v) part[v] = G.getPartitionIndex(v);
}

// quality_metrics::edge_cut (quality_metrics.cpp:23-37).
REF_API int ref_edge_cut(int n, const int* xadj, const int* adjncy, const int* adjwgt,
                         const int* part) {
    graph_access G;
    build_graph(G, n, xadj, adjncy, nullptr, adjwgt);
    quality_metrics qm;
    return qm.edge_cut(G, const_cast<int*>(part));
}

// quality_metrics::balance (quality_metrics.cpp:265-288).
REF_API double ref_balance(int n, const int* xadj, const int* adjncy, const int* vwgt, int k,
                           const int* part) {
    graph_access G;
    build_graph(G, n, xadj, adjncy, vwgt, nullptr);
    G.set_partition_count(k);
    for (int v = 0; v < n; ++v) G.setPartitionIndex(v, part[v]);
    quality_metrics qm;
    return qm.balance(G);
}

// The reference's own kaffpa() C entry point (interface/kaHIP_interface.cpp:146-166).
REF_API void ref_kaffpa(int n, const int* vwgt, const int* xadj, const int* adjwgt,
                        const int* adjncy, int k, double imbalance, int seed, int mode,
                        int* edgecut, int* part) {
    int nn = n, kk = k;
    double imb = imbalance;
    kaffpa(&nn, const_cast<int*>(vwgt), const_cast<int*>(xadj), const_cast<int*>(adjwgt),
           const_cast<int*>(adjncy), &kk, &imb, true, seed, mode, edgecut, part);
}

// kaffpa(FASTSOCIAL) unrolled into its phases with a wall timer around each: the same calls, in
// the same order, as kaffpa -> internal_kaffpa_call -> perform_partitioning -> single_run
// (interface/kaHIP_interface.cpp:95-166, graph_partitioner.cpp:240-273) for one V-cycle and one
// repetition. times[0..3] = coarsening, initial partitioning, uncoarsening (LP refinement),
// edge cut. levels[0] = number of levels, levels[1+i] = node count of level i (up to 62 levels).
REF_API void ref_kaffpa_fastsocial_phases(int n, const int* vwgt, const int* xadj,
                                          const int* adjwgt, const int* adjncy, int k,
                                          double imbalance, int seed, int* edgecut, int* part,
                                          double* times, int* levels) {
    PartitionConfig cfg;
    fastsocial_config(cfg, k);
    cfg.seed = seed;
    cfg.imbalance = 100 * imbalance;
    cfg.kaffpa_perfectly_balance = false;
    graph_access G;
    G.build_from_metis(n, const_cast<int*>(xadj), const_cast<int*>(adjncy));
    G.set_partition_count(cfg.k);
    srand(cfg.seed);
    random_functions::setSeed(cfg.seed);
    if (vwgt) {
        forall_nodes(G, v) { G.setNodeWeight(v, vwgt[v]); } endfor
    }
    if (adjwgt) {
        forall_edges(G, e) { G.setEdgeWeight(e, adjwgt[e]); } endfor
    }
    balance_configuration bc;
    bc.configurate_balance(cfg, G);

    coarsening coarsen;
    initial_partitioning init_part;
    uncoarsening uncoarsen;
    graph_hierarchy hierarchy;
    double t0 = now_s();
    coarsen.perform_coarsening(cfg, G, hierarchy);
    double t1 = now_s();
    if (levels) {
        // The hierarchy is a stack; walk it without disturbing it is not exposed, so record the
        // coarsest size and the number of levels (graph_hierarchy.h: size()).
        levels[0] = (int)hierarchy.size();
        levels[1] = (int)hierarchy.get_coarsest()->number_of_nodes();
    }
    init_part.perform_initial_partitioning(cfg, hierarchy);
    double t2 = now_s();
    uncoarsen.perform_uncoarsening(cfg, hierarchy);
    double t3 = now_s();
    for (int v = 0; v < n; ++v) part[v] = G.getPartitionIndex(v);
    quality_metrics qm;
    *edgecut = qm.edge_cut(G);
    double t4 = now_s();
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
        times[2] = t3 - t2;
        times[3] = t4 - t3;
    }
}

// One level of the reference coarsening loop body for CLUSTER_COARSENING
// (coarsening.cpp:58-80): LP clustering with B = ceil(U / factor) via match(), then contraction.
// Used to pin the oracle's level loop. Returns the number of coarse nodes; mapping receives the
// coarse mapping; output CSR arrays sized as for ref_contract.
REF_API int ref_coarsen_one_level(int n, const int* xadj, const int* adjncy, const int* vwgt,
                                  const int* adjwgt, int k, double imbalance, int seed,
                                  int* mapping, int* cxadj, int* cadjncy, int* cadjwgt,
                                  int* cvwgt) {
    PartitionConfig cfg;
    fastsocial_config(cfg, k);
    cfg.seed = seed;
    cfg.imbalance = 100 * imbalance;
    graph_access G;
    build_graph(G, n, xadj, adjncy, vwgt, adjwgt);
    G.set_partition_count(cfg.k);
    srand(seed);
    random_functions::setSeed(seed);
    balance_configuration bc;
    bc.configurate_balance(cfg, G);
    size_constraint_label_propagation sclp;
    Matching matching;
    CoarseMapping map;
    NodeID nc = 0;
    NodePermutationMap perm;
    sclp.match(cfg, G, matching, map, nc, perm);
    graph_access coarse;
    contraction c;
    c.contract(cfg, G, coarse, matching, map, nc, perm);
    for (int v = 0; v < n; ++v) mapping[v] = map[v];
    int cn = coarse.number_of_nodes();
    int pos = 0;
    for (int p = 0; p < cn; ++p) {
        cxadj[p] = pos;
        cvwgt[p] = coarse.getNodeWeight(p);
        forall_out_edges(coarse, e, p) {
            cadjncy[pos] = coarse.getEdgeTarget(e);
            cadjwgt[pos] = coarse.getEdgeWeight(e);
            ++pos;
        } endfor
    }
    cxadj[cn] = pos;
    return cn;
}

// The block bound U the reference derives for (W, k, imbalance) (app/balance_configuration.h:19-56).
REF_API unsigned ref_upper_bound(int n, const int* vwgt, int k, double imbalance) {
    PartitionConfig cfg;
    fastsocial_config(cfg, k);
    cfg.imbalance = 100 * imbalance;
    std::vector<int> xadj(n + 1, 0);
    graph_access G;
    G.build_from_metis(n, xadj.data(), nullptr);
    if (vwgt) {
        forall_nodes(G, v) { G.setNodeWeight(v, vwgt[v]); } endfor
    }
    balance_configuration bc;
    bc.configurate_balance(cfg, G);
    return cfg.upper_bound_partition;
}

// Times the reference's LP clustering (size_constraint_label_propagation::label_propagation,
// size_constraint_label_propagation.cpp:134-194, `iterations` sweeps, bound B) and the
// contraction of the resulting clustering (contraction.cpp:94-133) on one graph; graph
// construction is not timed. times[0] = LP seconds, times[1] = contraction seconds. Returns the
// number of clusters.
REF_API int ref_time_lp_level(int n, const int* xadj, const int* adjncy, const int* vwgt,
                              const int* adjwgt, unsigned bound, int iterations, int seed,
                              double* times) {
    graph_access G;
    build_graph(G, n, xadj, adjncy, vwgt, adjwgt);
    PartitionConfig cfg;
    fastsocial_config(cfg, 2);
    cfg.label_iterations = iterations;
    cfg.graph_allready_partitioned = false;
    cfg.combine = false;
    srand(seed);
    random_functions::setSeed(seed);
    std::vector<NodeWeight> cid;
    NodeID blocks = 0;
    size_constraint_label_propagation sclp;
    NodeWeight b = bound;
    double t0 = now_s();
    sclp.label_propagation(cfg, G, b, cid, blocks);
    double t1 = now_s();
    CoarseMapping map(n);
    for (int v = 0; v < n; ++v) map[v] = cid[v];
    Matching matching;
    NodePermutationMap perm;
    graph_access coarse;
    contraction c;
    NodeID ncv = blocks;
    double t2 = now_s();
    c.contract_clustering(cfg, G, coarse, matching, map, ncv, perm);
    double t3 = now_s();
    times[0] = t1 - t0;
    times[1] = t3 - t2;
    return (int)blocks;
}
