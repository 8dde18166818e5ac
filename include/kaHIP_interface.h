/* Drop-in C interface of the MI355X LP partitioner.
 *
 * Declares kaffpa, kaffpa_balance and kaffpa_balance_NE with the same symbols and argument lists
 * as the reference's public header (interface/kaHIP_interface.h:30-47) and the same mode
 * constants (:17-22), so callers of those three functions link against libkahip.so from this
 * repository without source changes. The header's other exports are NOT provided:
 * process_mapping and the MAPMODE_* constants (:24-25, :52-58), node_separator (:61-64),
 * reduced_nd (:68-70), edge_partitioning (:72-75) and reduced_nd_fast (:77-81). Programs that
 * call them (e.g. misc/example_library_call/interface_test.cpp:58, which calls process_mapping)
 * do not link against this library; they are outside the accelerated path.
 *
 * Arrays are METIS CSR, 0-based, both edge directions stored, no self-loops; vwgt and adjcwgt
 * may be NULL (unit weights); part is an output array of *n ints allocated by the caller;
 * *edgecut receives the edge cut.
 *
 * Every mode coarsens and refines with the GPU label-propagation pipeline of the FASTSOCIAL
 * preset (LP coarsening, CPU initial partitioning, LP refinement). ECO, STRONG, ECOSOCIAL and
 * STRONGSOCIAL add a CPU 2-way FM pass over adjacent block pairs on the finest level;
 * STRONGSOCIAL also clusters with ensembles; an unknown mode falls back to ECO, as in the
 * reference (interface/kaHIP_interface.cpp:60-62). imbalance == 0 selects the perfectly balanced
 * mode as in the reference (app/balance_configuration.h:41-43): the pipeline runs with the bound
 * 1.01 ceil(W / k), then a CPU post-pass moves nodes until every block weighs at most
 * ceil(W / k) (a greedy, cut-aware stand-in for the reference's cycle_refinement,
 * interface/kaHIP_interface.cpp:123-132; its cuts can be higher). The functions return
 * nothing, like the reference's; on an error they print one line to stderr, set
 * *edgecut = -1 and leave part unspecified. Use kaffpa_gpu_ex() from kahip_amd.h for a status
 * code and a message.
 */
#ifndef KAHIP_AMD_KAHIP_INTERFACE_H
#define KAHIP_AMD_KAHIP_INTERFACE_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Preset ids (reference: interface/kaHIP_interface.h:17-22). */
static const int FAST = 0;
static const int ECO = 1;
static const int STRONG = 2;
static const int FASTSOCIAL = 3;
static const int ECOSOCIAL = 4;
static const int STRONGSOCIAL = 5;

/* Replaces kaffpa (reference: interface/kaHIP_interface.h:30-33,
 * interface/kaHIP_interface.cpp:146-166). imbalance is a fraction (0.03 = 3 %). */
void kaffpa(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
            double* imbalance, bool suppress_output, int seed, int mode, int* edgecut, int* part);

/* Replaces kaffpa_balance (reference: interface/kaHIP_interface.h:36-41,
 * interface/kaHIP_interface.cpp:168-186). With perfectly_balance the CPU post-pass enforces
 * blocks of at most (1 + imbalance) ceil(W / k) (the pipeline's own bound; the post-pass moves
 * nodes only if a block exceeds it). */
void kaffpa_balance(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
                    double* imbalance, bool perfectly_balance, bool suppress_output, int seed,
                    int mode, int* edgecut, int* part);

/* Replaces kaffpa_balance_NE (reference: interface/kaHIP_interface.h:44-47): blocks are balanced
 * on node weight plus weighted degree (app/balance_configuration.h:27-37), except at
 * imbalance == 0, where the node weights are kept (as the reference does). A node weight plus
 * weighted degree above INT_MAX is an error. */
void kaffpa_balance_NE(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
                       double* imbalance, bool suppress_output, int seed, int mode, int* edgecut,
                       int* part);

#ifdef __cplusplus
}
#endif

#endif
