// Degree ordering of the level-0 graph (see reorder.hip).
#pragma once

#include "contract.hpp"
#include "lp.hpp"

namespace kahip {

// mode: 0 never, 1 always, -1 automatic (maximum degree above 32 x the average degree).
bool reorder_wanted(const DGraph& g, int mode);
// Temporaries of reorder_by_degree, kept by the caller so that repeated calls reuse them.
struct ReorderScratch {
    DevBuf<unsigned> keys, skeys;
    DevBuf<int> ids, order, ndeg;
    DevBuf<char> tmp;
};
// out: g with node i = the i-th node by (degree desc, id asc), rows in their original order with
// mapped ids; inv[old] = new.
void reorder_by_degree(const DGraph& g, OwnedGraph& out, DevBuf<int>& inv, ReorderScratch& t,
                       hipStream_t s, Profiler* prof);
// part_old[v] = part_new[inv[v]]
void unpermute_parts(int n, const int* inv, const int* part_new, int* part_old, hipStream_t s);

}  // namespace kahip
