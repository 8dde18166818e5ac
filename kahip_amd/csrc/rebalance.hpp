// CPU post-pass of the perfectly balanced modes (see rebalance.cpp).
#pragma once

#include <cstdint>

namespace kahip {

// Moves nodes out of blocks heavier than `bound` (gain-ordered, gains kept current); returns the
// cut of the result. part (in/out) holds block ids in [0, k).
long long rebalance(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, long long bound, int* part);

// rebalance() followed by pairwise 2-way FM under the same bound; returns the cut.
long long balance_postpass(int n, const int* xadj, const int* adjncy, const int* vwgt,
                           const int* adjwgt, int k, long long bound, uint64_t seed, int* part);

}  // namespace kahip
