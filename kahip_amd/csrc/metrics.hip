// Partition quality on gfx950: edge cut (reference: lib/tools/quality_metrics.cpp:23-37) and
// block weights (quality_metrics.cpp:265-288 uses them for balance).
//
// Edge cut is edge-parallel: each thread owns kChunk consecutive directed edges, finds the
// source of its first edge by binary search in xadj and walks forward; partial sums are reduced
// per workgroup and added with one 64-bit atomic. Integer sums: any order gives the same bits.

#include "common.hpp"
#include "metrics.hpp"
#include "lp.hpp"

namespace kahip {

namespace {

constexpr int kChunk = 8;
constexpr int kBlock = 256;

__device__ inline int source_of(const int* __restrict__ xadj, int n, long long e) {
    // largest u with xadj[u] <= e (and xadj[u+1] > e for a non-empty row)
    int lo = 0, hi = n;  // invariant: xadj[lo] <= e < xadj[hi]
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((long long)xadj[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void edge_cut_kernel(int n, long long m2,
                                                          const int* __restrict__ xadj,
                                                          const int* __restrict__ adjncy,
                                                          const int* __restrict__ adjwgt,
                                                          const int* __restrict__ part,
                                                          unsigned long long* __restrict__ out) {
    __shared__ long long red[kBlock / 64];
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long e0 = t * kChunk;
    long long acc = 0;
    if (e0 < m2) {
        int u = source_of(xadj, n, e0);
        long long next = xadj[u + 1];
        int pu = part[u];
        const long long e1 = e0 + kChunk < m2 ? e0 + kChunk : m2;
        for (long long e = e0; e < e1; ++e) {
            while (e >= next) {
                ++u;
                next = xadj[u + 1];
                pu = part[u];
            }
            if (part[adjncy[e]] != pu) acc += adjwgt[e];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long s = 0;
        for (int i = 0; i < kBlock / 64; ++i) s += red[i];
        if (s) atomicAdd(out, (unsigned long long)s);
    }
}

}  // namespace

long long MetricsWorkspace::edge_cut(const DGraph& g, const int* part, hipStream_t s,
                                     Profiler* prof) {
    acc_.reserve(1);
    fill_async(acc_.p, 1, 0ull, s);
    if (g.m2 > 0) {
        const long long threads = (g.m2 + kChunk - 1) / kChunk;
        // algorithmic bytes: 12 B per directed edge (target, weight, block of target) and
        // 12 B per node (offset, block of source, block weight read in the balance pass)
        ProfScope ps(prof, "edge_cut", 12.0 * (double)g.m2 + 12.0 * g.n, g.m2, s);
        hipLaunchKernelGGL(edge_cut_kernel, dim3((unsigned)((threads + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, g.n, g.m2, g.xadj, g.adjncy, g.adjwgt, part,
                           acc_.p);
    }
    unsigned long long h = 0;
    KH_CHECK(hipMemcpyAsync(&h, acc_.p, sizeof(h), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    return (long long)h / 2;
}

}  // namespace kahip
