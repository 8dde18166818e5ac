// Synthetic graphs generated on the GPU (the graph never leaves HBM).
//
// R-MAT (the 2^24-node / 2^28-sample workload): Sample j draws 64-bit words SplitMix64(seed*K1 + 2*K2 + idx) with
// idx = j*ceil(scale/2) + level/2 (low half for even levels, high half for odd levels) and picks
// a quadrant per level by integer thresholds floor(p * 2^32). Self-loops are dropped, both
// directions stored, duplicates merged with weight = multiplicity, rows sorted ascending; the
// same definition is restated in numpy in kahip_amd/graphs.py.
//
// Random geometric graph: point i has x = (d(2i) >> 11) 2^-53, y = (d(2i+1) >> 11) 2^-53 with
// d(idx) = SplitMix64(seed*K1 + K2 + idx); an edge joins two points at squared distance < r^2
// (r = 0.55 sqrt(ln n / n) unless given); nodes are numbered in grid-cell order (g = floor(1/r)
// cells per side, row-major, ties by generation index) and rows are sorted. All double
// arithmetic is rounded per operation (no fused multiply-add) so that the graph equals the
// numpy definition in kahip_amd/graphs.py bit for bit.

#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_run_length_encode.hpp>

#include "gen.hpp"
#include "lp.hpp"

namespace kahip {

namespace {

constexpr unsigned long long K1 = 0xD1B54A32D192ED03ULL;
constexpr unsigned long long K2 = 0x8CB92BA72F3D8DD7ULL;

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

__device__ inline unsigned long long sm64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void rmat_samples(long long samples, int scale, unsigned long long base,
                             unsigned t1, unsigned t2, unsigned t3,
                             unsigned long long* __restrict__ keys) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= samples) return;
    const unsigned long long per = (unsigned long long)((scale + 1) / 2);
    unsigned long long u = 0, v = 0, d = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
        if ((lvl & 1) == 0) d = sm64(base + (unsigned long long)j * per + (unsigned long long)(lvl >> 1));
        const unsigned r = (lvl & 1) == 0 ? (unsigned)(d & 0xffffffffu) : (unsigned)(d >> 32);
        const unsigned q = (unsigned)(r >= t1) + (unsigned)(r >= t2) + (unsigned)(r >= t3);
        u = u * 2 + (q >> 1);
        v = v * 2 + (q & 1);
    }
    if (u == v) {
        keys[2 * j] = ~0ull;
        keys[2 * j + 1] = ~0ull;
    } else {
        keys[2 * j] = (u << 32) | v;
        keys[2 * j + 1] = (v << 32) | u;
    }
}

__global__ void split_runs(long long m2, const unsigned long long* __restrict__ uniq,
                           const unsigned* __restrict__ counts, int* __restrict__ adjncy,
                           int* __restrict__ adjwgt, int* __restrict__ deg) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m2) return;
    const unsigned long long k = uniq[i];
    adjncy[i] = (int)(k & 0xffffffffu);
    adjwgt[i] = (int)counts[i];
    atomicAdd(&deg[(int)(k >> 32)], 1);
}

__global__ void max_deg(int n, const int* __restrict__ deg, int* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicMax(out, deg[v]);
}

__global__ void ones(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1;
}

// ------------------------------------------------------------------ random geometric graph
__global__ void rgg_points(int n, unsigned long long base, int g, double* __restrict__ px,
                           double* __restrict__ py, unsigned* __restrict__ cell,
                           int* __restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double scale = 1.0 / 9007199254740992.0;
    const double x = __dmul_rn((double)(sm64(base + 2ull * (unsigned long long)i) >> 11), scale);
    const double y = __dmul_rn((double)(sm64(base + 2ull * (unsigned long long)i + 1) >> 11), scale);
    const long long cx = min((long long)__dmul_rn(x, (double)g), (long long)g - 1);
    const long long cy = min((long long)__dmul_rn(y, (double)g), (long long)g - 1);
    px[i] = x;
    py[i] = y;
    cell[i] = (unsigned)(cy * g + cx);
    idx[i] = i;
}

__global__ void rgg_gather(int n, const int* __restrict__ order, const double* __restrict__ px,
                           const double* __restrict__ py, double* __restrict__ sx,
                           double* __restrict__ sy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sx[i] = px[order[i]];
    sy[i] = py[order[i]];
}

// start[c] = first position of cell c in the sorted cell array (cells 0..g*g)
__global__ void rgg_cell_start(int n, long long cells, const unsigned* __restrict__ scell,
                               int* __restrict__ start) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > cells) return;
    int lo = 0, hi = n;  // first i with scell[i] >= c
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)scell[mid] < c) lo = mid + 1; else hi = mid;
    }
    start[c] = lo;
}

// FILL = false: deg[i] = number of neighbours; FILL = true: writes them at adj[xadj[i]...] in
// ascending order (cells are visited in increasing id order and ids follow cell order).
template <bool FILL>
__global__ void rgg_edges(int n, int g, double r2, const double* __restrict__ sx,
                          const double* __restrict__ sy, const unsigned* __restrict__ scell,
                          const int* __restrict__ start, int* __restrict__ deg,
                          const int* __restrict__ xadj, int* __restrict__ adj) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)scell[i];
    const int cx = c % g, cy = c / g;
    const double x = sx[i], y = sy[i];
    int cnt = 0;
    int pos = FILL ? xadj[i] : 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int ny = cy + dy;
        if (ny < 0 || ny >= g) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int nx = cx + dx;
            if (nx < 0 || nx >= g) continue;
            const int nc = ny * g + nx;
            for (int j = start[nc]; j < start[nc + 1]; ++j) {
                if (j == i) continue;
                const double ddx = __dsub_rn(x, sx[j]);
                const double ddy = __dsub_rn(y, sy[j]);
                const double d2 = __dadd_rn(__dmul_rn(ddx, ddx), __dmul_rn(ddy, ddy));
                if (d2 < r2) {
                    if (FILL) adj[pos++] = j;
                    else ++cnt;
                }
            }
        }
    }
    if (!FILL) deg[i] = cnt;
}

}  // namespace

namespace {
unsigned long long sm64_host(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
}  // namespace

void batch_graph_spec(unsigned long long seed, int i, double log2_min, double log2_max, int* kind,
                      int* size, unsigned long long* graph_seed) {
    const unsigned long long d = sm64_host(seed * K1 + 3ULL * K2 + (unsigned long long)i);
    const double u = (double)(d >> 11) * (1.0 / 9007199254740992.0);
    const double x = log2_min + u * (log2_max - log2_min);
    *kind = i & 1;
    *size = *kind == 0 ? (int)std::floor(std::pow(2.0, x) + 0.5) : (int)std::floor(x + 0.5);
    *graph_seed = sm64_host(seed * K1 + 4ULL * K2 + (unsigned long long)i);
}

void generate_rgg(int n, unsigned long long seed, double radius, OwnedGraph& out) {
    KH_REQUIRE(n >= 2, "RGG needs n >= 2");
    const double r = radius > 0 ? radius : 0.55 * std::sqrt(std::log((double)n) / (double)n);
    const int g = std::max(1, (int)std::floor(1.0 / r));
    KH_REQUIRE((long long)g * g < (1LL << 31), "RGG grid too fine");
    const long long cells = (long long)g * g;
    const unsigned long long base = seed * K1 + 1ULL * K2;
    hipStream_t s;
    KH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf<double> px, py, sx, sy;
    DevBuf<unsigned> cell, scell;
    DevBuf<int> idx, order, start, deg;
    DevBuf<char> tmp;
    px.reserve(n);
    py.reserve(n);
    sx.reserve(n);
    sy.reserve(n);
    cell.reserve(n);
    scell.reserve(n);
    idx.reserve(n);
    order.reserve(n);
    start.reserve(cells + 1);
    deg.reserve((long long)n + 1);
    hipLaunchKernelGGL(rgg_points, dim3(grid_for(n, 256)), dim3(256), 0, s, n, base, g, px.p,
                       py.p, cell.p, idx.p);
    // stable sort by cell keeps the generation order inside a cell
    int bits = 1;
    while (bits < 32 && (1LL << bits) < cells) ++bits;
    size_t bytes = 0;
    KH_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, cell.p, scell.p, idx.p, order.p,
                                       (size_t)n, 0, bits, s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, cell.p, scell.p, idx.p, order.p, (size_t)n,
                                       0, bits, s));
    hipLaunchKernelGGL(rgg_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, n, order.p, px.p,
                       py.p, sx.p, sy.p);
    hipLaunchKernelGGL(rgg_cell_start, dim3(grid_for(cells + 1, 256)), dim3(256), 0, s, n, cells,
                       scell.p, start.p);
    const double r2 = r * r;
    hipLaunchKernelGGL(rgg_edges<false>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, g, r2, sx.p,
                       sy.p, scell.p, start.p, deg.p, nullptr, nullptr);
    fill_async(deg.p + n, 1, 0, s);
    out.alloc_nodes(n);
    bytes = 0;
    KH_CHECK(rocprim::exclusive_scan(nullptr, bytes, deg.p, out.g.xadj, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::exclusive_scan(tmp.p, bytes, deg.p, out.g.xadj, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
    int m2 = 0;
    KH_CHECK(hipMemcpyAsync(&m2, out.g.xadj + n, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    KH_REQUIRE(m2 >= 0, "RGG has too many edges for int32 offsets");
    out.alloc_edges(m2);
    hipLaunchKernelGGL(rgg_edges<true>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, g, r2, sx.p,
                       sy.p, scell.p, start.p, nullptr, out.g.xadj, out.g.adjncy);
    if (m2 > 0)
        hipLaunchKernelGGL(ones, dim3(grid_for(m2, 256)), dim3(256), 0, s, out.g.adjwgt, m2);
    hipLaunchKernelGGL(ones, dim3(grid_for(n, 256)), dim3(256), 0, s, out.g.vwgt, n);
    DevBuf<int> mx;
    mx.reserve(1);
    fill_async(mx.p, 1, 0, s);
    hipLaunchKernelGGL(max_deg, dim3(grid_for(n, 256)), dim3(256), 0, s, n, deg.p, mx.p);
    int maxd = 0;
    KH_CHECK(hipMemcpyAsync(&maxd, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    KH_CHECK(hipGetLastError());
    KH_CHECK(hipStreamDestroy(s));
    out.g.total_vwgt = n;
    out.g.max_degree = maxd;
}

void generate_rmat(int scale, long long samples, unsigned long long seed, double a, double b,
                   double c, OwnedGraph& out) {
    KH_REQUIRE(scale >= 1 && scale <= 30, "scale must be in [1, 30]");
    KH_REQUIRE(samples >= 0 && 2 * samples < (1LL << 31), "2*samples must be < 2^31");
    KH_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "bad R-MAT probabilities");
    const unsigned t1 = (unsigned)std::floor(a * 4294967296.0);
    const unsigned t2 = (unsigned)std::floor((a + b) * 4294967296.0);
    const unsigned t3 = (unsigned)std::floor((a + b + c) * 4294967296.0);
    const unsigned long long base = seed * K1 + 2ULL * K2;
    const int n = 1 << scale;
    hipStream_t s;
    KH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const long long nk = 2 * samples;
    DevBuf<unsigned long long> k1, k2, uniq;
    DevBuf<unsigned> counts;
    DevBuf<unsigned> nruns;
    DevBuf<char> tmp;
    k1.reserve(nk);
    k2.reserve(nk);
    if (samples > 0)
        hipLaunchKernelGGL(rmat_samples, dim3(grid_for(samples, 256)), dim3(256), 0, s, samples,
                           scale, base, t1, t2, t3, k1.p);
    size_t bytes = 0;
    KH_CHECK(rocprim::radix_sort_keys(nullptr, bytes, k1.p, k2.p, nk, 0, 64, s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::radix_sort_keys(tmp.p, bytes, k1.p, k2.p, nk, 0, 64, s));
    k1.free();
    uniq.reserve(nk);
    counts.reserve(nk);
    nruns.reserve(1);
    bytes = 0;
    KH_CHECK(rocprim::run_length_encode(nullptr, bytes, k2.p, (unsigned)nk, uniq.p, counts.p,
                                        nruns.p, s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::run_length_encode(tmp.p, bytes, k2.p, (unsigned)nk, uniq.p, counts.p,
                                        nruns.p, s));
    unsigned runs = 0;
    KH_CHECK(hipMemcpyAsync(&runs, nruns.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    long long m2 = runs;
    if (m2 > 0) {
        unsigned long long last = 0;
        KH_CHECK(hipMemcpy(&last, uniq.p + (m2 - 1), sizeof(last), hipMemcpyDeviceToHost));
        if (last == ~0ull) --m2;  // the self-loop sentinel run
    }
    k2.free();
    out.alloc_nodes(n);
    out.alloc_edges(m2);
    DevBuf<int> deg, mx;
    deg.reserve(n + 1);
    mx.reserve(1);
    fill_async(deg.p, n + 1, 0, s);
    fill_async(mx.p, 1, 0, s);
    if (m2 > 0)
        hipLaunchKernelGGL(split_runs, dim3(grid_for(m2, 256)), dim3(256), 0, s, m2, uniq.p,
                           counts.p, out.g.adjncy, out.g.adjwgt, deg.p);
    hipLaunchKernelGGL(max_deg, dim3(grid_for(n, 256)), dim3(256), 0, s, n, deg.p, mx.p);
    bytes = 0;
    KH_CHECK(rocprim::exclusive_scan(nullptr, bytes, deg.p, out.g.xadj, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
    tmp.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::exclusive_scan(tmp.p, bytes, deg.p, out.g.xadj, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
    hipLaunchKernelGGL(ones, dim3(grid_for(n, 256)), dim3(256), 0, s, out.g.vwgt, n);
    int maxd = 0;
    KH_CHECK(hipMemcpyAsync(&maxd, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    KH_CHECK(hipGetLastError());
    KH_CHECK(hipStreamDestroy(s));
    out.g.total_vwgt = n;
    out.g.max_degree = maxd;
}

}  // namespace kahip
