/* Extended C interface of the MI355X LP partitioner (libkahip.so).
 *
 * Status-returning entry points: 0 on success, non-zero on failure with a message available from
 * kahip_last_error() (thread-local). All arrays are host arrays unless a kahip_graph_t is used,
 * whose CSR lives in GPU memory. Conventions as in kaHIP_interface.h.
 */
#ifndef KAHIP_AMD_H
#define KAHIP_AMD_H

#include "kaHIP_interface.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kahip_graph* kahip_graph_t;

/* Parameters of one partitioning call; kahip_default_params() fills the FASTSOCIAL values. */
typedef struct {
    int k;
    double imbalance;
    int seed;
    int cluster_rounds;  /* LP clustering rounds per level (reference label_iterations = 10) */
    int refine_rounds;   /* LP refinement rounds per level (label_iterations_refinement = 25) */
    int subrounds;       /* sub-rounds per synchronous round, 1..8 */
    int cluster_factor;  /* cluster bound B = ceil(U / cluster_factor) (reference 18) */
    int stop_factor;     /* coarsening continues while n_coarse >= stop_factor * k (5000) */
    int ip_repetitions;  /* initial-partitioning repetitions; -1 = the reference's count */
    int ip_threads;      /* CPU threads of the initial partitioner; 0 = automatic */
    int bins[6];         /* upper degree bounds of the LP kernel bins; bins[4] <= 512 */
    int profile;         /* record timings (kahip_profile_json): 1 every launch, 2 only the
                            phases and the clustering proposal kernels */
    int rating;          /* edge rating of the LP clustering: KAHIP_RATING_* */
    int rating_levels;   /* the rating applies to coarsening levels < rating_levels */
    const float* mlp_weights; /* KAHIP_MLP_PARAMS floats (host), for KAHIP_RATING_LEARNED */
    int reorder;         /* renumber the input by degree first: -1 automatic (maximum degree
                            above 32 x the average), 0 off, 1 on; the partition is returned in
                            the caller's numbering */
    int ensemble;        /* LP clusterings per level, overlaid (1 = single clustering; the
                            STRONGSOCIAL preset uses 18/17/15/7/3 for k <= 3/7/15/31/more) */
    int ip_tries;        /* BFS-grown starts per bisection of the initial partitioner (32) */
    int fm_rounds;       /* > 0: CPU 2-way FM over adjacent block pairs on every level after its
                            LP refinement, at most this many passes per level (0 = off) */
    int graph_refine;    /* must be 0 (the default): hipGraph-captured refinement rounds are
                            disabled after an unexplained host fault in that path under a
                            profiler; non-zero is rejected with an error */
    int group_items;     /* tile shape: nodes per lane group in flight in the 8..64-lane LP
                            kernels, 1, 2 (default) or 4 */
    int wave_small;      /* tile shape: clustering wave-bin nodes up to this degree use the
                            small-table wave kernel, 64, 128 (default) or 256 */
} kahip_params;

/* Edge ratings of the clustering. NONE uses the edge weights (the FASTSOCIAL preset);
 * EXPANSION_STAR2 is w(e)^2 / (c(u) c(v)) (lib/partition/coarsening/edge_rating/
 * edge_ratings.cpp:171-183) quantised as clamp(round(1024 w^2 / (c(u) c(v))), 1, 4096);
 * LEARNED is the MLP 8 -> 64 -> 64 -> 1 on the per-edge features of kahip_edge_features,
 * quantised as clamp(round(128 logit) + 2048, 1, 4096). Scores are shifted right by s >= 0,
 * the smallest shift with max_degree * (4096 >> s) < 2^31, and kept >= 1. */
#define KAHIP_RATING_NONE 0
#define KAHIP_RATING_EXPANSION_STAR2 1
#define KAHIP_RATING_LEARNED 2
#define KAHIP_MLP_FEATURES 8
#define KAHIP_MLP_HIDDEN 64
/* W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3, row-major FP32 */
#define KAHIP_MLP_PARAMS 4801

#define KAHIP_MAX_LEVELS 64

typedef struct {
    long long cut;
    double balance;
    unsigned upper_bound;   /* U */
    unsigned cluster_bound; /* B */
    int levels;             /* number of contractions */
    int level_n[KAHIP_MAX_LEVELS + 1];
    long long level_m2[KAHIP_MAX_LEVELS + 1];
    double ms_coarsen, ms_initial, ms_refine, ms_cut;
    long long edges_cluster, edges_contract, edges_refine, edges_cut;
    long long max_block_weight;
    double ms_rating; /* rating time, part of ms_coarsen */
    int reordered;    /* the input was renumbered by degree */
    double ms_fm;     /* CPU pairwise FM on every level (kahip_params.fm_rounds > 0) */
    long long fm_gain; /* its cut reduction */
} kahip_stats;

void kahip_default_params(kahip_params* p);
const char* kahip_last_error(void);

/* Status-returning kaffpa (no reference counterpart: the reference's kaffpa returns void,
 * interface/kaHIP_interface.cpp:146-166). stats may be NULL. */
int kaffpa_gpu_ex(int n, const int* vwgt, const int* xadj, const int* adjcwgt, const int* adjncy,
                  const kahip_params* params, int* edgecut, int* part, kahip_stats* stats);

/* Device-resident graphs. */
int kahip_graph_from_host(int n, const int* xadj, const int* adjncy, const int* vwgt,
                          const int* adjwgt, kahip_graph_t* out);
/* R-MAT generated on the GPU: 2^scale nodes, `samples` samples with quadrant probabilities
 * (a, b, c, 1-a-b-c); self-loops dropped, symmetrised, duplicates merged with weight =
 * multiplicity, rows sorted. The stream of samples is the counter-based one documented in
 * kahip_amd/graphs.py. */
int kahip_graph_rmat(int scale, long long samples, unsigned long long seed, double a, double b,
                     double c, kahip_graph_t* out);
/* Random geometric graph generated on the GPU: n points uniform in [0,1)^2, an edge between two
 * points at distance < radius (radius <= 0: 0.55 sqrt(ln n / n)), unit weights, nodes in
 * grid-cell order, rows sorted; the same graph as kahip_amd/graphs.py gen_rgg. */
int kahip_graph_rgg(int n, unsigned long long seed, double radius, kahip_graph_t* out);
int kahip_graph_info(kahip_graph_t g, int* n, long long* m2, long long* total_vwgt,
                     int* max_degree);
int kahip_graph_to_host(kahip_graph_t g, int* xadj, int* adjncy, int* vwgt, int* adjwgt);
void kahip_graph_free(kahip_graph_t g);

/* kaffpa's pipeline on a device graph. part: host array of n ints (may be NULL). */
int kahip_partition(kahip_graph_t g, const kahip_params* params, int* part, kahip_stats* stats);

/* kaffpa, kaffpa_balance(_NE), kahip_partition and kahip_batch_partition keep their device
 * workspaces (and stream) per device between calls, so that repeated calls do not allocate them
 * again; this frees those of the engines no call is using. No counterpart in the reference
 * (its kaffpa keeps no state between calls). */
int kahip_release_workspaces(void);
/* JSON object {name: {launches, ms, bytes, edges}} of the last profiled call on this thread. */
int kahip_profile_json(char* buf, int len);

/* Fine-grained steps (host arrays in and out; one GPU round trip each). */
/* Synchronous LP clustering + first-appearance relabel (kernel (a) of the path). */
int kahip_lp_cluster(int n, const int* xadj, const int* adjncy, const int* vwgt,
                     const int* adjwgt, int bound, int rounds, int subrounds, unsigned seed,
                     int* cluster, int* num_clusters, int* moves_per_round);
/* First-appearance relabel of arbitrary labels in [0, n). */
int kahip_remap(int n, const int* label, int* map, int* num_clusters);
/* Quotient graph of a clustering with ids 0..nc-1; cadjncy / cadjwgt hold xadj[n] entries. */
int kahip_contract(int n, const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                   const int* map, int nc, int* cxadj, int* cadjncy, int* cadjwgt, int* cvwgt);
/* Synchronous LP refinement of part (in/out) under block bound `bound`. */
int kahip_lp_refine(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, int bound, int rounds, int subrounds, unsigned seed,
                    int* part, int* moves_per_round);
int kahip_edge_cut(int n, const int* xadj, const int* adjncy, const int* adjwgt, const int* part,
                   long long* cut);
int kahip_block_weights(int n, const int* vwgt, int k, const int* part, long long* weights);
/* Edge cut and block weights of a partition in device memory, allocating nothing: no counterpart
 * in the reference. It is for callers that cannot allocate device memory (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc). Every pointer is a device pointer:
 * the CSR (n nodes, m2 = xadj[n] directed entries; adjwgt / vwgt NULL for unit weights), part[n]
 * with ids in [0, k), and the outputs *cut (one long long) and block_weights[k]. The work is
 * enqueued on `stream` (a hipStream_t of the current device; NULL: the null stream) and not
 * waited for. k <= 65536, the size of the library's per-GPU scratch array (2 + 65536 u64,
 * 512 KiB, held until the process ends); larger k is rejected. *cut receives -1 if a part
 * entry lies outside [0, k). Calls that overlap on one GPU share that scratch and race. */
int kahip_quality_noalloc(int n, long long m2, const int* xadj, const int* adjncy,
                          const int* adjwgt, const int* vwgt, const int* part, int k,
                          long long* cut, long long* block_weights, void* stream);
int kahip_project(int n_fine, const int* map, int n_coarse, const int* coarse_part,
                  int* fine_part);
/* CPU initial partitioner of the coarsest graph (no GPU). repetitions <= 0 and
 * bisection_tries <= 0 select the defaults (the reference's repetition count; 32 BFS-grown
 * starts per bisection, kahip_params.ip_tries). */
int kahip_initial_partition(int n, const int* xadj, const int* adjncy, const int* vwgt,
                            const int* adjwgt, int k, long long block_bound, int repetitions,
                            int bisection_tries, unsigned long long seed, int* part);

/* CPU post-pass of the perfectly balanced modes (the reference runs cycle_refinement there,
 * interface/kaHIP_interface.cpp:123-132): moves nodes out of blocks heavier than `bound`, best
 * cut change first with gains kept current, until every block is within the bound or no move
 * is possible, then refines by 2-way FM over adjacent block pairs under the same bound. part
 * (in/out) in [0, k); *cut (nullable) receives the resulting cut. No GPU. */
int kahip_rebalance(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, long long bound, unsigned long long seed,
                    int* part, long long* cut);

/* CPU refinement of a k-way partition by 2-way FM on the pairs of adjacent blocks (replaces
 * the quotient-graph refinement of lib/partition/uncoarsening/refinement/quotient_graph_refinement/
 * quotient_graph_refinement.cpp:48-130 with two_way_fm.cpp). part[n] in [0, k) is refined in
 * place; *gain (nullable) receives the cut reduction. No GPU. */
int kahip_fm_refine(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, long long block_bound, int rounds,
                    unsigned long long seed, int* part, long long* gain);

/* Learned edge rating (no reference counterpart; the reference rates edges in
 * lib/partition/coarsening/edge_rating/edge_ratings.cpp:171-183 and does not use the rating on
 * the LP path). features: xadj[n] x 8 floats per directed edge (u, v):
 *   0 log2(1 + min(deg u, deg v))  1 log2(1 + max deg)  2 log2(1 + w(e))
 *   3 log2(1 + min(c(u), c(v)))    4 log2(1 + max c)    5 popcount(sig(u) & sig(v)) / 8
 *   6 number of 3 one-round, one-sub-round LP clusterings (cluster bound `bound`, seeds
 *     hash3(seed, 0xA6, i)) that put u and v together
 *   7 log2(1 + w(e)^2 / max(c(u) c(v), 1))
 * with sig(x) = OR over neighbours y of 1 << (mix32(y) & 63). */
int kahip_edge_features(int n, const int* xadj, const int* adjncy, const int* vwgt,
                        const int* adjwgt, int bound, unsigned seed, float* features);
int kahip_graph_edge_features(kahip_graph_t g, int bound, unsigned seed, float* features);
/* MLP logits (BF16 operands on MFMA, FP32 accumulate) of count feature rows. */
int kahip_mlp_forward(long long count, const float* features, const float* params,
                      float* logits);
/* Integer scores of a rating (KAHIP_RATING_*) for every directed edge. */
int kahip_edge_scores(int n, const int* xadj, const int* adjncy, const int* vwgt,
                      const int* adjwgt, int rating, const float* params, int bound,
                      unsigned seed, int* scores);
/* Mean binary cross-entropy (labels in {0, 1}) of the MLP and its gradient, computed on the
 * GPU in fixed order. The forward rounds inputs, weights and layer-1 activations to BF16 like
 * the MFMA path; the backward treats those roundings as the identity. */
int kahip_mlp_gradient(int count, const float* features, const float* labels,
                       const float* params, float* grads, double* loss);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` passes of consecutive mini-batches of the
 * rows in the given order; params is updated in place, loss_per_epoch (may be NULL) receives
 * each pass's mean of batch losses. Deterministic. */
int kahip_mlp_train(long long count, const float* features, const float* labels, int batch,
                    int epochs, float lr, float* params, double* loss_per_epoch);

/* ----------------------------------------------------------------------------------------
 * Batched mode (no reference counterpart: the reference partitions one graph per call,
 * interface/kaHIP_interface.cpp:66-166). G independent graphs live in HBM as one disjoint union
 * (node ids of graph g offset by node_off[g]); each call partitions every graph into k blocks
 * with its own bounds (U_g, B_g from its own weight) and its own stop rule, sharing every GPU
 * launch. Per-graph results equal the union's synchronous schedule with per-label bounds, which
 * the oracle restates (orc_lp_cluster_sync_ex / orc_lp_refine_sync_ex). */
typedef struct kahip_batch* kahip_batch_t;

typedef struct {
    double ms_coarsen, ms_initial, ms_refine, ms_cut;
    int levels;
    long long edges_cluster, edges_contract, edges_refine, edges_cut;
} kahip_batch_stats;

/* xadj: the graphs' local CSR offsets concatenated (node_off[G] + G entries, graph g's at
 * xadj[node_off[g] + g ...]); adjncy: local neighbour ids concatenated (edge_off[G] entries);
 * vwgt / adjwgt concatenated or NULL. Every graph needs at least one node. */
int kahip_batch_from_host(int num_graphs, const long long* node_off, const long long* edge_off,
                          const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                          kahip_batch_t* out);
/* Graph i: RGG (even i) with n = floor(2^x + 0.5) nodes or R-MAT (odd i) of scale
 * floor(x + 0.5) with 16 samples per node, x = log2_min + u (log2_max - log2_min), u and the
 * graph's seed drawn from `seed` (kahip_amd/graphs.py batch_spec restates it). */
int kahip_batch_generate(int num_graphs, double log2_min, double log2_max, unsigned long long seed,
                         kahip_batch_t* out);
int kahip_batch_info(kahip_batch_t b, int* num_graphs, int* n, long long* m2, int* node_off);
/* The union CSR (global ids). */
int kahip_batch_to_host(kahip_batch_t b, int* xadj, int* adjncy, int* vwgt, int* adjwgt);
void kahip_batch_free(kahip_batch_t b);
/* part: n local block ids (may be NULL); cuts / max_block_weights: G values (may be NULL).
 * use_graph must be 0: hipGraph-captured round loops are disabled (non-zero is an error). */
int kahip_batch_partition(kahip_batch_t b, const kahip_params* params, int use_graph, int* part,
                          long long* cuts, long long* max_block_weights,
                          kahip_batch_stats* stats);
/* Autotuner of the LP degree bins for a device graph or batch (exactly one non-NULL): times two
 * LP clustering rounds per candidate upper bound of the wavefront bin (bins[4] in 96..512) and
 * returns the fastest bins; results of the LP do not depend on the bins. ms_per_candidate
 * (may be NULL) receives KAHIP_AUTOTUNE_CANDIDATES timings. */
#define KAHIP_AUTOTUNE_CANDIDATES 6
int kahip_autotune_bins(kahip_graph_t g, kahip_batch_t b, const kahip_params* params, int* bins,
                        double* ms_per_candidate);
/* Autotuner of the LP kernels' degree bins and tile shapes for a device graph or batch (exactly
 * one non-NULL), by coordinate descent, each candidate timed on two LP clustering rounds: the
 * wavefront bin's upper degree (bins[4] in 96..512), then group_items (1, 2, 4), then wave_small
 * (64, 128, 256). *tuned receives *params (or the defaults) with bins, group_items and wave_small
 * set to the fastest configuration; the LP results do not depend on them. ms (may be NULL)
 * receives KAHIP_AUTOTUNE_EVALS timings in that order. */
#define KAHIP_AUTOTUNE_EVALS 12
int kahip_autotune(kahip_graph_t g, kahip_batch_t b, const kahip_params* params,
                   kahip_params* tuned, double* ms);

#ifdef __cplusplus
}
#endif

#endif
