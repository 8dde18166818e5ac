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

}  // namespace kahip
