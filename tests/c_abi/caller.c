/* A C caller of the drop-in interface, built the way a program built against the reference's
 * kaHIP_interface.h would be: it includes only include/kaHIP_interface.h and links -lkahip.
 * Partitions the 5-node graph of misc/example_library_call/interface_test.cpp:18-24 and an
 * n-node cycle with kaffpa, kaffpa_balance and kaffpa_balance_NE, printing one line per call:
 *   <function> <k> <imbalance> <edgecut> <part...> */
#include <stdio.h>
#include <stdlib.h>

#include "kaHIP_interface.h"

static void show(const char* fn, int k, double imb, int cut, const int* part, int n) {
    printf("%s %d %g %d", fn, k, imb, cut);
    for (int i = 0; i < n; ++i) printf(" %d", part[i]);
    printf("\n");
}

int main(void) {
    int n = 5;
    int xadj[] = {0, 2, 5, 7, 9, 12};
    int adjncy[] = {1, 4, 0, 2, 4, 1, 3, 2, 4, 0, 1, 3};
    int part[5];
    int nparts = 2, edgecut = 0;
    double imbalance = 0.03;
    kaffpa(&n, NULL, xadj, NULL, adjncy, &nparts, &imbalance, false, 0, FASTSOCIAL, &edgecut,
           part);
    show("kaffpa", nparts, imbalance, edgecut, part, n);

    /* a cycle of 1000 nodes, perfectly balanced into 3 blocks (imbalance 0) */
    int m = 1000;
    int* cx = malloc(sizeof(int) * (m + 1));
    int* ca = malloc(sizeof(int) * 2 * m);
    int* cp = malloc(sizeof(int) * m);
    for (int v = 0; v < m; ++v) {
        int a = (v + m - 1) % m, b = (v + 1) % m;
        cx[v] = 2 * v;
        ca[2 * v] = a < b ? a : b;
        ca[2 * v + 1] = a < b ? b : a;
    }
    cx[m] = 2 * m;
    int k3 = 3;
    double zero = 0.0;
    kaffpa(&m, NULL, cx, NULL, ca, &k3, &zero, true, 1, ECO, &edgecut, cp);
    show("kaffpa", k3, zero, edgecut, cp, m);
    kaffpa_balance(&m, NULL, cx, NULL, ca, &k3, &imbalance, true, true, 2, FAST, &edgecut, cp);
    show("kaffpa_balance", k3, imbalance, edgecut, cp, m);
    kaffpa_balance_NE(&m, NULL, cx, NULL, ca, &k3, &imbalance, true, 3, 99, &edgecut, cp);
    show("kaffpa_balance_NE", k3, imbalance, edgecut, cp, m);
    free(cx);
    free(ca);
    free(cp);
    return 0;
}
