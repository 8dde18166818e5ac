/* Drop-in C interface of the MI355X LP partitioner.
 *
 * Declares the same symbols, argument lists and mode constants as the reference's public header
 * (interface/kaHIP_interface.h:17-41 and :44-47), so a program built against the reference links
 * against libkahip.so from this repository unchanged. Arrays are METIS CSR, 0-based, both edge
 * directions stored, no self-loops; vwgt and adjcwgt may be NULL (unit weights); part is an
 * output array of *n ints allocated by the caller; *edgecut receives the edge cut.
 *
 * Every mode runs the GPU label-propagation pipeline of the FASTSOCIAL preset (LP coarsening,
 * CPU initial partitioning, LP refinement). The functions return nothing, like the reference's;
 * on an error they print one line to stderr, set *edgecut = -1 and leave part unspecified. Use
 * kaffpa_gpu_ex() from kahip_amd.h for a status code and a message.
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
 * interface/kaHIP_interface.cpp:168-186). With perfectly_balance the result is rebalanced on the
 * CPU until every block weighs at most (1 + imbalance) * ceil(W / k). */
void kaffpa_balance(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
                    double* imbalance, bool perfectly_balance, bool suppress_output, int seed,
                    int mode, int* edgecut, int* part);

/* Replaces kaffpa_balance_NE (reference: interface/kaHIP_interface.h:44-47): blocks are balanced
 * on node weight plus weighted degree (app/balance_configuration.h:27-37). */
void kaffpa_balance_NE(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
                       double* imbalance, bool suppress_output, int seed, int mode, int* edgecut,
                       int* part);

#ifdef __cplusplus
}
#endif

#endif
