// Batched mode: many independent graphs partitioned together as one disjoint union in HBM.
//
// The union's CSR is the concatenation of the graphs (node ids offset by node_off[g]); every LP,
// contraction and cut launch covers all graphs at once, so thousands of small graphs share each
// launch. Per-graph quantities enter as per-label arrays: the cluster bound B_g of a node's graph
// (labels are node ids during clustering) and the block bound U_g of labels g*k + b during
// refinement. A graph leaves the level loop when its own stop rule says so; its nodes are then
// left out of later LP launches (identity clustering) and out of refinement above its coarsest
// level.
#pragma once

#include <vector>

#include "driver.hpp"

namespace kahip {

struct BatchStats {
    double ms_coarsen = 0, ms_initial = 0, ms_refine = 0, ms_cut = 0;
    int levels = 0;                  // contractions of the union
    long long edges_cluster = 0, edges_contract = 0, edges_refine = 0, edges_cut = 0;
    std::vector<long long> cut;      // per graph
    std::vector<long long> max_block_weight;
    std::vector<int> coarsest_level; // per graph
    std::vector<int> coarsest_n;
    std::vector<unsigned> upper_bound;
};

// g: the union (level 0), node_off: host array of G + 1 node offsets (every graph has >= 1
// node). part_out: host array of g.n local block ids (0..k-1 per graph). use_graph must be false
// (the captured round loops are disabled; an error otherwise).
void partition_batch(Engine& eng, const DGraph& g, int G, const int* node_off,
                     const PartitionParams& p, int* part_out, BatchStats* st,
                     bool use_graph = false);

// Concatenates G host graphs (local ids) into a device union.
void upload_batch(int G, const long long* node_off, const long long* edge_off, const int* xadj,
                  const int* adjncy, const int* vwgt, const int* adjwgt, OwnedGraph& out,
                  hipStream_t s);

// Generates the batched workload on the GPU: graph i is an RGG (even i) or an R-MAT graph with
// 16 n samples (odd i) with n = round(2^(log2_min + u (log2_max - log2_min))), u uniform from
// the graph's seed. Fills node_off (G + 1).
void generate_batch(int G, double log2_min, double log2_max, unsigned long long seed,
                    OwnedGraph& out, std::vector<int>& node_off);

// Times two LP clustering rounds per candidate configuration on g and returns the fastest, by
// coordinate descent: the wavefront bin's upper degree (kAutotuneCandidates values), then, with
// shapes, the group kernels' nodes per lane group (1, 2, 4) and the small wave kernel's degree
// limit (64, 128, 256). ms (optional) receives each evaluation's time in that order
// (kAutotuneCandidates, or kAutotuneEvals with shapes).
constexpr int kAutotuneCandidates = 6;
constexpr int kAutotuneEvals = kAutotuneCandidates + 6;
BinConfig autotune_bins(const DGraph& g, const PartitionParams& p, bool shapes, double* ms);

}  // namespace kahip
