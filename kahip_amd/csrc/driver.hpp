// Multilevel driver of the FASTSOCIAL path: GPU LP coarsening + contraction, CPU initial
// partitioning of the coarsest graph, GPU projection + LP refinement, GPU edge cut.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "contract.hpp"
#include "devbuf.hpp"
#include "initial_partition.hpp"
#include "lp.hpp"
#include "metrics.hpp"
#include "rating.hpp"
#include "reorder.hpp"

namespace kahip {

constexpr int kMaxLevels = 64;

struct PartitionParams {
    int k = 2;
    double imbalance = 0.03;
    int seed = 0;
    int cluster_rounds = 10;   // app/configuration.h:424 (label_iterations)
    int refine_rounds = 25;    // app/configuration.h:425 (label_iterations_refinement)
    int subrounds = 4;
    int cluster_factor = 18;   // app/configuration.h:422 (cluster_coarsening_factor)
    int stop_factor = 5000;    // app/configuration.h:497 (num_vert_stop_factor)
    int ip_repetitions = -1;   // -1: the reference's count (initial_partitioning.cpp:69-81)
    int ip_threads = 0;
    int ip_tries = 32;         // BFS-grown starts per bisection of the initial partitioner
    int fm_rounds = 0;         // > 0: CPU pairwise FM on every level after its LP refinement
    int graph_refine = 0;      // must be 0: hipGraph-captured refinement rounds are disabled
    BinConfig bins;
    int profile = 0;  // 0 off, 1 every launch, 2 phases and the clustering proposal kernels
    // Edge rating that replaces the edge weights in the LP clustering of levels < rating_levels
    // (None: the edge weights themselves, as in the FASTSOCIAL preset).
    RatingMode rating = RatingMode::None;
    int reorder = -1;  // degree renumbering of the input: -1 automatic (skewed), 0 off, 1 on
    // Ensemble clusterings (size_constraint_label_propagation.cpp:91-123): that many LP
    // clusterings per level, overlaid (two nodes share a coarse node iff every clustering puts
    // them together); 1 = a single clustering (FASTSOCIAL).
    int ensemble = 1;
    int rating_levels = kMaxLevels;
    const MlpWeights* mlp = nullptr;  // required for RatingMode::Learned
};

struct PartitionStats {
    long long cut = 0;
    double balance = 0;
    unsigned upper_bound = 0;
    unsigned cluster_bound = 0;
    int levels = 0;  // number of contractions
    int level_n[kMaxLevels + 1] = {};
    long long level_m2[kMaxLevels + 1] = {};
    double ms_coarsen = 0, ms_initial = 0, ms_refine = 0, ms_cut = 0;
    double ms_rating = 0;  // part of ms_coarsen
    bool reordered = false;
    double ms_fm = 0;      // CPU pairwise FM, all levels
    long long fm_gain = 0;
    long long edges_cluster = 0, edges_contract = 0, edges_refine = 0, edges_cut = 0;
    long long max_block_weight = 0;
};

// Per-level seeds of the synchronous schedule.
inline uint32_t cluster_seed(int seed, int level) {
    return hash3((uint32_t)seed, 0xC1u, (uint32_t)level);
}
inline uint32_t refine_seed(int seed, int level) {
    return hash3((uint32_t)seed, 0x4Eu, (uint32_t)level);
}
// The reference's clusterings number for the social presets (app/configuration.h:494-509).
int ensemble_size_for(int k);
// Cluster factor of ensemble run i >= 1: 10 + hash % 21, i.e. in [10, 30] as the reference's
// nextInt(10, 30) (size_constraint_label_propagation.cpp:119); run 0 uses the configured factor.
inline int ensemble_factor(int seed, int level, int run) {
    return 10 + (int)(hash3((uint32_t)seed, 0xE5u + (uint32_t)level, (uint32_t)run) % 21u);
}
inline uint32_t ensemble_seed(int seed, int level, int run) {
    return hash3((uint32_t)seed, 0xE6u + (uint32_t)level, (uint32_t)run);
}
inline uint32_t rating_seed(int seed, int level) {
    return hash3((uint32_t)seed, 0x5Au, (uint32_t)level);
}

// app/balance_configuration.h:40-46 with the interface's percent round trip
// (interface/kaHIP_interface.cpp:115): U = (1 + eps) * ceil(W / k) truncated to unsigned;
// imbalance == 0 selects the perfectly balanced bound (1 + 0.01) * ceil(W / k) (:41-43).
unsigned upper_bound_for(long long total_weight, int k, double imbalance);
// size_constraint_label_propagation.cpp:57: ceil(U / factor).
unsigned cluster_bound_for(unsigned upper, int factor);
// initial_partitioning.cpp:69-81 for the FASTSOCIAL preset (16 repetitions, minipreps 4).
int ip_repetitions_for(int k);
constexpr int kPresetFmRounds = 3;  // pairwise FM passes of the FM-based kaffpa presets
// seed of the pairwise FM of a partition with this seed (level l uses fm_seed(seed) + l):
// (uint32 seed) x 0x9E3779B97F4A7C15 ^ 0xF3F3F3F3
inline uint64_t fm_seed(int seed) {
    return ((uint64_t)(uint32_t)seed * 0x9E3779B97F4A7C15ULL) ^ 0xF3F3F3F3ULL;
}

class Engine {
public:
    Engine();
    ~Engine();
    // Partitions the device graph g (level 0). part_out: host array of g.n block ids.
    void partition(const DGraph& g, const PartitionParams& p, int* part_out,
                   PartitionStats* stats);
    Profiler prof;
    hipStream_t stream() const { return s_; }

    // building blocks used by the fine-grained C entry points
    LpWorkspace lp;
    ContractWorkspace cw;
    MetricsWorkspace mw;
    RatingWorkspace rw;

private:
    hipStream_t s_ = nullptr;
    // Device buffers of partition() (grown, never shrunk: a later call on a graph no larger
    // allocates nothing); every one is written before it is read within a call.
    struct Buffers {
        OwnedGraph reordered;
        DevBuf<int> inv;
        ReorderScratch rtmp;
        std::vector<std::unique_ptr<OwnedGraph>> coarse;  // coarse[l] = level l+1
        std::vector<DevBuf<int>> maps;                     // maps[l]: level l -> level l+1
        DevBuf<int> label, size, score, elabel, part_c, part_f, W;
        DevBuf<unsigned long long> okeys, okeys2;
        DevBuf<int> oids, oids2;
        DevBuf<char> otmp;
    } buf_;
};

// Uploads a host CSR (vwgt / adjwgt may be null: unit weights) into an owned device graph.
void upload_graph(int n, const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                  OwnedGraph& out, hipStream_t s);
void download_graph(const DGraph& g, HostGraph& out, hipStream_t s);

}  // namespace kahip
