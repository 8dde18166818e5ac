// Partition quality on gfx950: edge cut (reference: lib/tools/quality_metrics.cpp:23-37) and
// block weights (quality_metrics.cpp:265-288 uses them for balance).
//
// Edge cut is edge-parallel: each thread owns kChunk consecutive directed edges, finds the
// source of its first edge by binary search in xadj and walks forward; partial sums are reduced
// per workgroup and added with one 64-bit atomic. Integer sums: any order gives the same bits.

#include <algorithm>
#include <mutex>

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

// UNIT: adjwgt is null and every edge weighs 1.
template <bool UNIT = false>
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
            if (part[adjncy[e]] != pu) acc += UNIT ? 1 : adjwgt[e];
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

// ----------------------------------------------------------------------------- no allocation
// Scratch of kahip_quality_noalloc, one copy per GPU for the life of the process:
// [0] the doubled cut, [1] the number of out-of-range block ids, [2, 2 + k) the block weights.
constexpr int kNoAllocMaxK = 65536;
constexpr int kNoAllocLdsK = 4096;  // block histograms in LDS (u64, 32 KB) up to this k
static __device__ __align__(256) unsigned long long scratch[2 + kNoAllocMaxK];

// Block weights into sc[2 + b] and the count of ids outside [0, k) into sc[1]; node weights vwgt
// or 1 when vwgt is null.
__global__ __launch_bounds__(kBlock) void block_weights_noalloc(int n, const int* __restrict__ part,
                                                                const int* __restrict__ vwgt,
                                                                int k,
                                                                unsigned long long* __restrict__ sc) {
    __shared__ unsigned long long lw[kNoAllocLdsK];
    const bool lds = k <= kNoAllocLdsK;  // uniform
    if (lds)
        for (int i = threadIdx.x; i < k; i += kBlock) lw[i] = 0ull;
    __syncthreads();
    unsigned long long bad = 0;
    for (long long v = (long long)blockIdx.x * kBlock + threadIdx.x; v < n;
         v += (long long)gridDim.x * kBlock) {
        const int b = part[v];
        const unsigned long long w = vwgt ? (unsigned long long)(long long)vwgt[v] : 1ull;
        if ((unsigned)b >= (unsigned)k) {
            ++bad;
            continue;
        }
        if (lds) atomicAdd(&lw[b], w);
        else atomicAdd(&sc[2 + b], w);
    }
    if (bad) atomicAdd(&sc[1], bad);
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < k; i += kBlock)
            if (lw[i]) atomicAdd(&sc[2 + i], lw[i]);
}

// cut[0] = doubled cut / 2, or -1 if a block id was out of range; weights[b] = block weight.
__global__ void finish_quality_noalloc(int k, const unsigned long long* __restrict__ sc,
                                       long long* __restrict__ cut,
                                       long long* __restrict__ weights) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) cut[0] = sc[1] ? -1LL : (long long)(sc[0] / 2);
    if (i < k) weights[i] = (long long)sc[2 + i];
}

}  // namespace

// The scratch array's device address on each GPU, looked up once per device with that device
// current (the symbol's name is only a host-side handle, never a device address).
static std::mutex g_scratch_mu;
static unsigned long long* g_scratch_dev[64] = {};

static unsigned long long* scratch_on_current_device() {
    int dev = 0;
    KH_CHECK(hipGetDevice(&dev));
    KH_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        KH_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        KH_REQUIRE(ptr != nullptr, "scratch symbol has no device address");
        g_scratch_dev[dev] = static_cast<unsigned long long*>(ptr);
    }
    return g_scratch_dev[dev];
}

void quality_noalloc(int n, long long m2, const int* xadj, const int* adjncy, const int* adjwgt,
                     const int* vwgt, const int* part, int k, long long* cut, long long* weights,
                     hipStream_t s) {
    KH_REQUIRE(n >= 0 && m2 >= 0, "n and m2 must be non-negative");
    KH_REQUIRE(k >= 1 && k <= kNoAllocMaxK,
               "k must be in [1, 65536]: the no-allocation scratch holds 65536 block weights");
    KH_REQUIRE(cut && weights, "cut and weights must be device pointers");
    KH_REQUIRE(n == 0 || (xadj && part), "xadj and part must be device pointers");
    KH_REQUIRE(m2 == 0 || adjncy, "adjncy must be a device pointer");
    if (s) {
        int sdev = 0, dev = 0;
        KH_CHECK(hipStreamGetDevice(s, &sdev));
        KH_CHECK(hipGetDevice(&dev));
        KH_REQUIRE(sdev == dev, "the stream must belong to the current device");
    }
    unsigned long long* sc = scratch_on_current_device();
    fill_async(sc, 2LL + k, 0ull, s);  // a fill kernel: the call stays capturable
    if (m2 > 0) {
        const long long threads = (m2 + kChunk - 1) / kChunk;
        const dim3 grid((unsigned)((threads + kBlock - 1) / kBlock));
        if (adjwgt)
            hipLaunchKernelGGL(edge_cut_kernel<false>, grid, dim3(kBlock), 0, s, n, m2, xadj,
                               adjncy, adjwgt, part, sc);
        else
            hipLaunchKernelGGL(edge_cut_kernel<true>, grid, dim3(kBlock), 0, s, n, m2, xadj,
                               adjncy, adjwgt, part, sc);
    }
    if (n > 0) {
        const int grid = (int)std::min<long long>(((long long)n + kBlock - 1) / kBlock, 2048);
        hipLaunchKernelGGL(block_weights_noalloc, dim3(grid), dim3(kBlock), 0, s, n, part, vwgt,
                           k, sc);
    }
    hipLaunchKernelGGL(finish_quality_noalloc, dim3((k + 255) / 256), dim3(256), 0, s, k, sc, cut,
                       weights);
    KH_CHECK(hipGetLastError());
}

long long MetricsWorkspace::edge_cut(const DGraph& g, const int* part, hipStream_t s,
                                     Profiler* prof) {
    acc_.reserve(1);
    fill_async(acc_.p, 1, 0ull, s);
    if (g.m2 > 0) {
        const long long threads = (g.m2 + kChunk - 1) / kChunk;
        // algorithmic bytes: 12 B per directed edge (target, weight, block of target) and
        // 12 B per node (offset, block of source, block weight read in the balance pass)
        ProfScope ps(prof, "edge_cut", 12.0 * (double)g.m2 + 12.0 * g.n, g.m2, s);
        hipLaunchKernelGGL(edge_cut_kernel<false>, dim3((unsigned)((threads + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, g.n, g.m2, g.xadj, g.adjncy, g.adjwgt, part,
                           acc_.p);
    }
    unsigned long long h = 0;
    KH_CHECK(hipMemcpyAsync(&h, acc_.p, sizeof(h), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    return (long long)h / 2;
}

}  // namespace kahip
