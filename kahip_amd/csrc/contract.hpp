// Host interface of the remap / contraction / projection kernels.
#pragma once

#include "common.hpp"
#include "devbuf.hpp"

namespace kahip {

// A device graph that owns its arrays.
struct OwnedGraph {
    DevBuf<int> xadj, adjncy, adjwgt, vwgt;
    DGraph g;
    void alloc_nodes(int n) {
        g.n = n;
        g.xadj = xadj.reserve((long long)n + 1);
        g.vwgt = vwgt.reserve(n);
    }
    void alloc_edges(long long m2) {
        g.m2 = m2;
        g.adjncy = adjncy.reserve(m2);
        g.adjwgt = adjwgt.reserve(m2);
    }
    int* xadj_dev() { return xadj.p; }
};

class ContractWorkspace {
public:
    // First-appearance relabel: map[v] = rank of the smallest member of label[v]'s cluster among
    // all cluster minima. Returns the number of clusters (one host readback).
    int remap(const DGraph& g, const int* label, int* map, hipStream_t s, Profiler* prof);
    // Quotient graph of the clustering `map` (values 0..nc-1) in the reference's exact order.
    void contract(const DGraph& g, const int* map, int nc, OwnedGraph& out, hipStream_t s,
                  Profiler* prof);

private:
    void scan_i32(const int* in, int* out, int n, hipStream_t s);
    DevBuf<char> scan_tmp_;
    DevBuf<int> leader_, flag_, newid_;
    DevBuf<int> cnt_, fdeg_, moff_, tpos_, cdeg_, members_, tq_, tw_;
    DevBuf<int> iota_, mkeys_;               // radix sort of the node ids by cluster
    DevBuf<int> rlist_[5], rcount_;          // rows of the thread tier and the four LDS tiers
    DevBuf<int> ovl2_list_, ovl_count_;      // rows of the global tier
    DevBuf<long long> hcap_, hoff_;
    DevBuf<int> gchunk_prefix_, mdeg_, mpre_, ofd_;
    DevBuf<int> hk_, hv_, hm_;
    DevBuf<int> gtask_prefix_, goffm_, gcnt_, gpos_, gbig_, gnbig_;
    DevBuf<long long> goffm_base_, gct_base_;
    DevBuf<int4> gpart_;
    DevBuf<int> gcrow_, gtrow_;              // row of every global chunk / (row, bucket) task
    DevBuf<int4> gda_, gdb_;                 // per-chunk descriptors (gchunk_desc)
    DevBuf<int> stage_q_, stage_w_;          // merged entries at the f of their first edge
    DevBuf<int> ecnt_, epos_;                // staged entries per chunk and their scan
    DevBuf<int> err_;  // device error flag (kErrProbe) of the last contract()
};

// fine_part[v] = coarse_part[map[v]] (reference: graph_hierarchy.cpp:54-58).
void project_partition(int n_fine, const int* map, const int* coarse_part, int* fine_part,
                       hipStream_t s, Profiler* prof);

}  // namespace kahip
