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

// Workgroup w fills the new CSR entries [w * kCopyTile, (w + 1) * kCopyTile): the rows that
// overlap the tile (at most kCopyTile + 1 of them, rows being non-empty inside the range) are
// staged in LDS as (new offset, old offset), then thread t takes entries t, t + 256, ... and
// finds its row by a binary search in LDS, so consecutive threads read and write consecutive
// entries.
constexpr int kCopyTile = 2048;

__global__ __launch_bounds__(256) void copy_rows(int n, long long m2,
                                                 const int* __restrict__ nxadj,
                                                 const int* __restrict__ order,
                                                 const int* __restrict__ xadj,
                                                 const int* __restrict__ adjncy,
                                                 const int* __restrict__ adjwgt,
                                                 const int* __restrict__ inv,
                                                 int* __restrict__ nadj, int* __restrict__ nw) {
    __shared__ int nbeg[kCopyTile + 1];  // new row offsets of the tile's rows
    __shared__ int obeg[kCopyTile + 1];  // their old row offsets
    __shared__ int r0s, nrow;
    const long long t0 = (long long)blockIdx.x * kCopyTile;
    const long long t1 = min(t0 + kCopyTile, m2);
    if (threadIdx.x == 0) {
        int lo = 0, hi = n - 1;  // last row with nxadj[row] <= t0
        while (lo < hi) {
            const int mid = lo + ((hi - lo + 1) >> 1);
            if ((long long)nxadj[mid] <= t0) lo = mid; else hi = mid - 1;
        }
        int hi2 = n - 1, lo2 = lo;  // last row with nxadj[row] < t1
        while (lo2 < hi2) {
            const int mid = lo2 + ((hi2 - lo2 + 1) >> 1);
            if ((long long)nxadj[mid] < t1) lo2 = mid; else hi2 = mid - 1;
        }
        r0s = lo;
        // rows are sorted by degree, descending, so empty rows come last and never start
        // inside a tile: every row but the first starts at its own offset in (t0, t1), hence
        // at most kCopyTile rows (the clamp only guards the LDS arrays)
        nrow = min(lo2 - lo + 1, kCopyTile + 1);
    }
    __syncthreads();
    const int r0 = r0s, nr = nrow;
    for (int i = threadIdx.x; i < nr; i += 256) {
        nbeg[i] = nxadj[r0 + i];
        obeg[i] = xadj[order[r0 + i]];
    }
    __syncthreads();
    for (long long e = t0 + threadIdx.x; e < t1; e += 256) {
        int lo = 0, hi = nr - 1;  // last staged row starting at or before e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((long long)nbeg[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const long long src = (long long)obeg[lo] + (e - nbeg[lo]);
        nadj[e] = inv[adjncy[src]];
        nw[e] = adjwgt[src];
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

void reorder_by_degree(const DGraph& g, OwnedGraph& out, DevBuf<int>& inv, ReorderScratch& t,
                       hipStream_t s, Profiler* prof) {
    ProfScope ps(prof, "reorder", 0.0, g.m2, s);
    const int n = g.n;
    DevBuf<unsigned>& keys = t.keys;
    DevBuf<unsigned>& skeys = t.skeys;
    DevBuf<int>& ids = t.ids;
    DevBuf<int>& order = t.order;
    DevBuf<int>& ndeg = t.ndeg;
    DevBuf<char>& tmp = t.tmp;
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
        hipLaunchKernelGGL(copy_rows, dim3(grid_for(g.m2, kCopyTile)), dim3(256), 0, s, n, g.m2,
                           out.g.xadj, order.p, g.xadj, g.adjncy, g.adjwgt, inv.p, out.g.adjncy,
                           out.g.adjwgt);
    out.g.total_vwgt = g.total_vwgt;
    out.g.max_degree = g.max_degree;
    KH_CHECK(hipGetLastError());
}

void unpermute_parts(int n, const int* inv, const int* part_new, int* part_old, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(gather_parts, dim3(grid_for(n, 256)), dim3(256), 0, s, n, inv,
                           part_new, part_old);
    KH_CHECK(hipGetLastError());
}

}  // namespace kahip
