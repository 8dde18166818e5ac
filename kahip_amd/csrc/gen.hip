// Synthetic R-MAT graphs generated on the GPU (input for the 2^24-node / 2^28-sample workload;
// the graph never leaves HBM). Sample j draws 64-bit words SplitMix64(seed*K1 + 2*K2 + idx) with
// idx = j*ceil(scale/2) + level/2 (low half for even levels, high half for odd levels) and picks
// a quadrant per level by integer thresholds floor(p * 2^32). Self-loops are dropped, both
// directions stored, duplicates merged with weight = multiplicity, rows sorted ascending; the
// same definition is restated in numpy in kahip_amd/graphs.py.

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

}  // namespace

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
