// Degree ordering of the input graph before the multilevel pipeline (skewed graphs only).
//
// New id i is the rank of the node by degree, descending, ties by old id; new row i is the old row
// of order[i] in its original edge order with neighbour ids mapped through inv (no re-sort). On
// R-MAT graphs most edge endpoints fall on few high-degree nodes (the top 1 % of the nodes of a
// 2^20 instance carry 49 % of the endpoints, the top 10 % carry 87 %); numbered first, their labels
// form a small dense block of the label array that stays in L2, where the natural order scatters
// them over the whole array and every label gather touches its own cache line.

#include "reorder.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace kahip {

namespace {

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

__global__ void degree_keys(int n, const int* __restrict__ xadj, unsigned* __restrict__ keys,
                            int* __restrict__ ids) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    keys[v] = ~(unsigned)(xadj[v + 1] - xadj[v]);  // ascending key = descending degree
    ids[v] = v;
}

__global__ void invert_and_degrees(int n, const int* __restrict__ order,
                                   const int* __restrict__ xadj, const int* __restrict__ vwgt,
                                   int* __restrict__ inv, int* __restrict__ ndeg,
                                   int* __restrict__ nvwgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int u = order[i];
    inv[u] = i;
    ndeg[i] = xadj[u + 1] - xadj[u];
    nvwgt[i] = vwgt[u];
}

// Each thread fills 8 consecutive entries of the new CSR (row found by binary search).
__global__ void copy_rows(int n, long long m2, const int* __restrict__ nxadj,
                          const int* __restrict__ order, const int* __restrict__ xadj,
                          const int* __restrict__ adjncy, const int* __restrict__ adjwgt,
                          const int* __restrict__ inv, int* __restrict__ nadj,
                          int* __restrict__ nw) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long e0 = t * 8;
    if (e0 >= m2) return;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((long long)nxadj[mid] <= e0) lo = mid; else hi = mid;
    }
    int i = lo;
    long long next = nxadj[i + 1];
    long long src = (long long)xadj[order[i]] - nxadj[i];
    const long long e1 = min(e0 + 8, m2);
    for (long long e = e0; e < e1; ++e) {
        while (e >= next) {
            ++i;
            next = nxadj[i + 1];
            src = (long long)xadj[order[i]] - nxadj[i];
        }
        nadj[e] = inv[adjncy[src + e]];
        nw[e] = adjwgt[src + e];
    }
}

__global__ void gather_parts(int n, const int* __restrict__ inv, const int* __restrict__ in,
                             int* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = in[inv[v]];
}

}  // namespace

bool reorder_wanted(const DGraph& g, int mode) {
    if (mode == 0 || g.n == 0) return false;
    if (mode > 0) return true;
    // skewed degrees: maximum degree above 32 x the average
    return (long long)g.max_degree * (long long)g.n > 32LL * g.m2;
}

void reorder_by_degree(const DGraph& g, OwnedGraph& out, DevBuf<int>& inv, hipStream_t s,
                       Profiler* prof) {
    ProfScope ps(prof, "reorder", 0.0, g.m2, s);
    const int n = g.n;
    DevBuf<unsigned> keys, skeys;
    DevBuf<int> ids, order, ndeg;
    DevBuf<char> tmp;
    keys.reserve(n);
    skeys.reserve(n);
    ids.reserve(n);
    order.reserve(n);
    ndeg.reserve((long long)n + 1);
    inv.reserve(n);
    out.alloc_nodes(n);
    hipLaunchKernelGGL(degree_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n, g.xadj, keys.p,
                       ids.p);
    size_t bytes = 0;
    KH_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, keys.p, skeys.p, ids.p, order.p,
                                       (size_t)n, 0, 32, s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, keys.p, skeys.p, ids.p, order.p, (size_t)n,
                                       0, 32, s));
    hipLaunchKernelGGL(invert_and_degrees, dim3(grid_for(n, 256)), dim3(256), 0, s, n, order.p,
                       g.xadj, g.vwgt, inv.p, ndeg.p, out.g.vwgt);
    fill_async(ndeg.p + n, 1, 0, s);
    bytes = 0;
    KH_CHECK(rocprim::exclusive_scan(nullptr, bytes, ndeg.p, out.g.xadj, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::exclusive_scan(tmp.p, bytes, ndeg.p, out.g.xadj, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
    out.alloc_edges(g.m2);
    if (g.m2 > 0)
        hipLaunchKernelGGL(copy_rows, dim3(grid_for((g.m2 + 7) / 8, 256)), dim3(256), 0, s, n,
                           g.m2, out.g.xadj, order.p, g.xadj, g.adjncy, g.adjwgt, inv.p,
                           out.g.adjncy, out.g.adjwgt);
    out.g.total_vwgt = g.total_vwgt;
    out.g.max_degree = g.max_degree;
    KH_CHECK(hipStreamSynchronize(s));  // temporaries are freed on return
    KH_CHECK(hipGetLastError());
}

void unpermute_parts(int n, const int* inv, const int* part_new, int* part_old, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(gather_parts, dim3(grid_for(n, 256)), dim3(256), 0, s, n, inv,
                           part_new, part_old);
    KH_CHECK(hipGetLastError());
}

}  // namespace kahip
