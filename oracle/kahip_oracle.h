/* The oracle: a CPU restatement of KaHIP's label-propagation (LP) coarsening and refinement path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load this library, and only as the checker. The shipped library (kahip_amd) never links it.
 *
 * Two families of functions:
 *   R1  restatement of the reference's sequential algorithm (asynchronous LP with the reference's
 *       std::mt19937 coin flips and std::sort degree order, first-appearance remap, quotient-graph
 *       contraction order, queue-driven LP refinement, edge cut, balance). Pinned bit-exact
 *       against the reference compiled from its own sources (oracle/_ref) and against closed
 *       forms / brute force in tests/.
 *   R2  the synchronous schedule the GPU runs (hash tie-breaks, sub-rounds, deterministic
 *       propose/admit under the size bound). There is no reference counterpart; it is pinned by
 *       invariants and special cases in tests/ and the GPU is checked bit-exact against it.
 *
 * All arrays are 0-based METIS CSR: xadj[n+1], adjncy[xadj[n]], int32. vwgt / adjwgt may be NULL
 * (unit weights), as in interface/kaHIP_interface.cpp:79-89.
 */
#ifndef KAHIP_ORACLE_H
#define KAHIP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- R1: reference restatement ---------------- */
int orc_lp_cluster_async(int n, const int* xadj, const int* adjncy, const int* vwgt,
                         const int* adjwgt, unsigned bound, int iterations, int seed,
                         int* cluster_out);
int orc_remap_first_appearance(int n, int* cluster);
int orc_contract(int n, const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                 const int* cluster, int nc, int* cxadj, int* cadjncy, int* cadjwgt, int* cvwgt);
void orc_lp_refine_async(int n, const int* xadj, const int* adjncy, const int* vwgt,
                         const int* adjwgt, int k, unsigned bound, int rounds, int seed,
                         int* part);
int64_t orc_edge_cut(int n, const int* xadj, const int* adjncy, const int* adjwgt,
                     const int* part);
void orc_block_weights(int n, const int* vwgt, int k, const int* part, int64_t* weights);
double orc_balance(int n, const int* vwgt, int k, const int* part);
unsigned orc_upper_bound(int64_t total_weight, int k, double imbalance);
unsigned orc_cluster_bound(unsigned upper_bound, int factor);
int orc_stop_rule_continue(int n_finer, int n_coarser, int k, int stop_factor);
void orc_project(int n_fine, const int* mapping, const int* coarse_part, int* fine_part);
void orc_order_nodes_degree(int n, const int* xadj, int* perm);

/* ---------------- R2: the synchronous schedule ---------------- */
uint32_t orc_hash3(uint32_t seed, uint32_t a, uint32_t b);
int orc_lp_cluster_sync(int n, const int* xadj, const int* adjncy, const int* vwgt,
                        const int* adjwgt, unsigned bound, int rounds, int subrounds,
                        uint32_t seed, int* cluster_out, int* moves_per_round);
void orc_lp_refine_sync(int n, const int* xadj, const int* adjncy, const int* vwgt,
                        const int* adjwgt, int k, unsigned bound, int rounds, int subrounds,
                        uint32_t seed, int* part, int* moves_per_round);
/* Per-label bounds and an optional inclusion mask (batched mode over a disjoint union). */
int orc_lp_cluster_sync_ex(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, const int* lbound, const uint8_t* include,
                           int rounds, int subrounds, uint32_t seed, int* cluster_out,
                           int* moves_per_round);
void orc_lp_refine_sync_ex(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, int k, const int* lbound, const uint8_t* include,
                           int rounds, int subrounds, uint32_t seed, int* part,
                           int* moves_per_round);

#ifdef __cplusplus
}
#endif
#endif
