// CPU refinement of a k-way partition by 2-way FM on the pairs of adjacent blocks (the quotient
// graph's edges), after the reference's quotient-graph refinement
// (lib/partition/uncoarsening/refinement/quotient_graph_refinement/quotient_graph_refinement.cpp:48-130
// with two_way_fm.cpp:40-150 and the boundary of complete_boundary.h:104-214).
#pragma once

#include <cstdint>
#include <vector>

#include "initial_partition.hpp"

namespace kahip {

struct FmRefineParams {
    int k = 2;
    long long block_bound = 0;  // U: a move may not push its target block above U
    int rounds = 3;             // passes over all adjacent block pairs (stops early without gain)
    int stall_pct = 25;         // a pair's FM stops after max(15, pct % of its boundary) moves
                                // without improvement (two_way_fm.cpp:69-71, where the limit is a
                                // preset's percentage; 25 here, as FM runs on the finest level
                                // only: 5 gave 0.5-1 % less gain on the delaunay/rgg fixtures)
    uint64_t seed = 0;
    long long pass_slack = 0;   // within a pass a move may take its target up to U + pass_slack
                                // (so that a balanced pair can exchange nodes); only prefixes
                                // whose overload is no higher than at the pass start are kept
};

// Refines part in place; returns the cut reduction, which is >= 0 when every block respects U
// (an overloaded block is relieved first, possibly at a higher cut). Blocks that respect U keep
// respecting it; an overloaded block only loses weight.
long long fm_refine_pairs(const HostGraph& g, const FmRefineParams& p, std::vector<int>& part);

}  // namespace kahip
