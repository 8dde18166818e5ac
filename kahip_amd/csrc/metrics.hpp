// Host interface of the quality-metric kernels.
#pragma once

#include "common.hpp"
#include "devbuf.hpp"

namespace kahip {

class MetricsWorkspace {
public:
    // Sum over directed edges with endpoints in different blocks, halved (one readback).
    long long edge_cut(const DGraph& g, const int* part, hipStream_t s, Profiler* prof);

private:
    DevBuf<unsigned long long> acc_;
};

// Edge cut and block weights of a partition held in device memory, enqueued on s without any
// allocation or synchronisation (kahip_quality_noalloc). Uses the library's per-GPU scratch
// array (2 + 65536 u64): calls that overlap on one GPU race on it.
void quality_noalloc(int n, long long m2, const int* xadj, const int* adjncy, const int* adjwgt,
                     const int* vwgt, const int* part, int k, long long* cut, long long* weights,
                     hipStream_t s);

}  // namespace kahip
