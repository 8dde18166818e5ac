// Batched mode (see batch.hpp). Level loop, initial partitioning and uncoarsening follow the
// single-graph driver (driver.hip) with per-graph bounds, stop rules and coarsest levels.

#include "batch.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>

#include <rocprim/device/device_scan.hpp>

#include "gen.hpp"

namespace kahip {

namespace {

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

// node_graph[v] = g with off[g] <= v < off[g + 1]
__global__ void node_graph_kernel(int n, int G, const int* __restrict__ off,
                                  int* __restrict__ node_graph) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int lo = 0, hi = G;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= v) lo = mid; else hi = mid;
    }
    node_graph[v] = lo;
}

// Per node: the cluster bound of its graph and whether its graph takes part in this level.
__global__ void node_tables(int n, const int* __restrict__ node_graph,
                            const int* __restrict__ graph_bound,
                            const uint8_t* __restrict__ graph_on, int* __restrict__ nbound,
                            uint8_t* __restrict__ include) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int g = node_graph[v];
    nbound[v] = graph_bound[g];
    include[v] = graph_on[g];
}

__global__ void graph_weights(int n, const int* __restrict__ node_graph,
                              const int* __restrict__ vwgt, unsigned long long* __restrict__ W) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicAdd(&W[node_graph[v]], (unsigned long long)vwgt[v]);
}

// coarse offsets: the first node of graph g is the smallest member of its cluster, so its coarse
// id is the number of coarse nodes of graphs before g
__global__ void coarse_offsets(int G, const int* __restrict__ off, const int* __restrict__ map,
                               int nc, int* __restrict__ coff) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < G) coff[g] = map[off[g]];
    if (g == 0) coff[G] = nc;
}

// part[v] = g * k + local part (for graphs whose coarsest level is this one), or g * k for
// graphs not yet present (kept in range; they are excluded from refinement)
__global__ void place_parts(int n, int k, const int* __restrict__ node_graph,
                            const int* __restrict__ src, const uint8_t* __restrict__ take,
                            int* __restrict__ part) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int g = node_graph[v];
    if (take[g]) part[v] = g * k + src[v];
}

__global__ void fill_graph_base(int n, int k, const int* __restrict__ node_graph,
                                int* __restrict__ part) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) part[v] = node_graph[v] * k;
}

__global__ void to_local(int n, int k, const int* __restrict__ node_graph,
                         const int* __restrict__ part, int* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = part[v] - node_graph[v] * k;
}

// Per-graph cut: each thread walks 8 consecutive directed edges (sources found by binary
// search), summing into the source's graph; sums are flushed when the graph changes.
__global__ void graph_cuts(int n, long long m2, const int* __restrict__ xadj,
                           const int* __restrict__ adjncy, const int* __restrict__ adjwgt,
                           const int* __restrict__ part, const int* __restrict__ node_graph,
                           unsigned long long* __restrict__ cut) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long e0 = t * 8;
    if (e0 >= m2) return;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((long long)xadj[mid] <= e0) lo = mid; else hi = mid;
    }
    int u = lo;
    long long next = xadj[u + 1];
    int g = node_graph[u];
    unsigned long long acc = 0;
    const long long e1 = min(e0 + 8, m2);
    for (long long e = e0; e < e1; ++e) {
        while (e >= next) {
            ++u;
            next = xadj[u + 1];
        }
        const int gu = node_graph[u];
        if (gu != g) {
            if (acc) atomicAdd(&cut[g], acc);
            acc = 0;
            g = gu;
        }
        if (part[u] != part[adjncy[e]]) acc += (unsigned long long)adjwgt[e];
    }
    if (acc) atomicAdd(&cut[g], acc);
}

__global__ void shift_copy(long long count, const int* __restrict__ src, int add,
                           int* __restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i] + add;
}

struct Ev {
    hipEvent_t a = nullptr, b = nullptr;
    Ev() {
        KH_CHECK(hipEventCreate(&a));
        KH_CHECK(hipEventCreate(&b));
    }
    ~Ev() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    float ms() const {
        float t = 0;
        KH_CHECK(hipEventElapsedTime(&t, a, b));
        return t;
    }
};

// Runs the LP rounds with per-label bounds and checks the device error flag.
void run_lp(LpWorkspace& lp, LpMode mode, const DGraph& g, int* label, int* size, int rounds,
            uint32_t seed, hipStream_t s, const int* lbound, const char* phase) {
    lp.run(mode, g, label, size, 0, rounds, seed, s, nullptr, phase, lbound);
    lp.check(s);
}

}  // namespace

BinConfig autotune_bins(const DGraph& g, const PartitionParams& p, bool shapes, double* ms) {
    static const int wave_hi[kAutotuneCandidates] = {96, 128, 192, 256, 384, 512};
    static const int items[3] = {1, 2, 4};
    static const int small[3] = {64, 128, 256};
    hipStream_t s;
    KH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    LpWorkspace lp;
    DevBuf<int> label, size;
    label.reserve(std::max(g.n, 1));
    size.reserve(std::max(g.n, 1));
    const unsigned U = upper_bound_for(g.total_vwgt, p.k, p.imbalance);
    const int B = (int)cluster_bound_for(U, p.cluster_factor);
    const uint32_t seed = cluster_seed(p.seed, 0);
    Ev ev;
    int evals = 0;
    // two LP clustering rounds under bc, timed
    auto time_config = [&](const BinConfig& bc) {
        LpWorkspace::init_cluster_labels(g, label.p, size.p, s);
        lp.prepare(LpMode::Cluster, g, g.n, p.subrounds, seed, bc, s);
        KH_CHECK(hipEventRecord(ev.a, s));
        lp.run(LpMode::Cluster, g, label.p, size.p, B, 2, seed, s, nullptr, "autotune");
        KH_CHECK(hipEventRecord(ev.b, s));
        KH_CHECK(hipStreamSynchronize(s));
        lp.check(s);
        const float t = ev.ms();
        if (ms) ms[evals] = t;
        ++evals;
        return t;
    };
    // coordinate descent: the wave bin's upper degree, then (with shapes) the group kernels'
    // items per lane group, then the small wave kernel's degree limit
    BinConfig best = p.bins;
    float best_ms = 1e30f;
    for (int c = 0; c < kAutotuneCandidates; ++c) {
        BinConfig bc = best;
        bc.hi[kNumBins - 2] = wave_hi[c];
        const float t = time_config(bc);
        if (t < best_ms) {
            best_ms = t;
            best = bc;
        }
    }
    if (shapes) {
        const BinConfig base = best;
        for (int c = 0; c < 3; ++c) {
            BinConfig bc = base;
            bc.group_items = items[c];
            const float t = time_config(bc);
            if (t < best_ms) {
                best_ms = t;
                best = bc;
            }
        }
        const BinConfig base2 = best;
        for (int c = 0; c < 3; ++c) {
            BinConfig bc = base2;
            bc.wave_small = small[c];
            const float t = time_config(bc);
            if (t < best_ms) {
                best_ms = t;
                best = bc;
            }
        }
    }
    KH_CHECK(hipStreamDestroy(s));
    return best;
}

void upload_batch(int G, const long long* node_off, const long long* edge_off, const int* xadj,
                  const int* adjncy, const int* vwgt, const int* adjwgt, OwnedGraph& out,
                  hipStream_t s) {
    KH_REQUIRE(G >= 1, "a batch needs at least one graph");
    const long long N = node_off[G], M = edge_off[G];
    KH_REQUIRE(N < (1LL << 31) - 1 && M < (1LL << 31), "batch exceeds int32 union offsets");
    std::vector<int> gx((size_t)N + 1), ga((size_t)std::max<long long>(M, 1));
    for (int g = 0; g < G; ++g) {
        const long long n0 = node_off[g], n1 = node_off[g + 1], e0 = edge_off[g];
        KH_REQUIRE(n1 > n0, "every graph of a batch needs at least one node");
        const int* lx = xadj + n0 + g;  // graph g's local offsets: n_g + 1 entries
        KH_REQUIRE(lx[0] == 0 && lx[n1 - n0] == edge_off[g + 1] - e0, "batch xadj mismatch");
        for (long long v = 0; v < n1 - n0; ++v) gx[n0 + v] = (int)(e0 + lx[v]);
        for (long long e = e0; e < edge_off[g + 1]; ++e) {
            const int t = adjncy[e];
            KH_REQUIRE(t >= 0 && t < n1 - n0, "batch neighbour id out of range");
            ga[e] = (int)(t + n0);
        }
    }
    gx[N] = (int)M;
    upload_graph((int)N, gx.data(), ga.data(), vwgt, adjwgt, out, s);
}

void generate_batch(int G, double log2_min, double log2_max, unsigned long long seed,
                    OwnedGraph& out, std::vector<int>& node_off) {
    KH_REQUIRE(G >= 1 && log2_min >= 1 && log2_max >= log2_min && log2_max <= 24,
               "bad batch parameters");
    std::vector<OwnedGraph> parts(G);
    node_off.assign(G + 1, 0);
    std::vector<long long> eoff(G + 1, 0);
    for (int i = 0; i < G; ++i) {
        int kind, size;
        unsigned long long gseed;
        batch_graph_spec(seed, i, log2_min, log2_max, &kind, &size, &gseed);
        if (kind == 0) generate_rgg(size, gseed, 0.0, parts[i]);
        else generate_rmat(size, 16LL << size, gseed, 0.57, 0.19, 0.19, parts[i]);
        KH_REQUIRE((long long)node_off[i] + parts[i].g.n < (1LL << 31) - 1, "batch too large");
        node_off[i + 1] = node_off[i] + parts[i].g.n;
        eoff[i + 1] = eoff[i] + parts[i].g.m2;
    }
    const long long N = node_off[G], M = eoff[G];
    KH_REQUIRE(M < (1LL << 31), "batch exceeds 2^31 directed edges");
    hipStream_t s;
    KH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    out.alloc_nodes((int)N);
    out.alloc_edges(M);
    int maxd = 0;
    for (int i = 0; i < G; ++i) {
        const DGraph& p = parts[i].g;
        hipLaunchKernelGGL(shift_copy, dim3(grid_for(p.n + 1, 256)), dim3(256), 0, s,
                           (long long)p.n + 1, p.xadj, (int)eoff[i], out.g.xadj + node_off[i]);
        if (p.m2 > 0) {
            hipLaunchKernelGGL(shift_copy, dim3(grid_for(p.m2, 256)), dim3(256), 0, s, p.m2,
                               p.adjncy, node_off[i], out.g.adjncy + eoff[i]);
            KH_CHECK(hipMemcpyAsync(out.g.adjwgt + eoff[i], p.adjwgt, sizeof(int) * p.m2,
                                    hipMemcpyDeviceToDevice, s));
        }
        KH_CHECK(hipMemcpyAsync(out.g.vwgt + node_off[i], p.vwgt, sizeof(int) * p.n,
                                hipMemcpyDeviceToDevice, s));
        maxd = std::max(maxd, p.max_degree);
    }
    KH_CHECK(hipStreamSynchronize(s));
    KH_CHECK(hipGetLastError());
    KH_CHECK(hipStreamDestroy(s));
    out.g.total_vwgt = N;
    out.g.max_degree = maxd;
}

void partition_batch(Engine& eng, const DGraph& g0, int G, const int* node_off,
                     const PartitionParams& p, int* part_out, BatchStats* st, bool use_graph) {
    KH_REQUIRE(!use_graph, "hipGraph-captured LP rounds are disabled: the captured path hit a host segmentation "
                   "fault under rocprofv3 whose cause was not found");
    KH_REQUIRE(G >= 1 && node_off[0] == 0 && node_off[G] == g0.n, "bad batch offsets");
    KH_REQUIRE(p.k >= 1 && (long long)G * p.k < (1LL << 31), "bad k for the batch");
    KH_REQUIRE(p.rating == RatingMode::None, "ratings are not supported in batched mode");
    for (int i = 0; i < G; ++i) KH_REQUIRE(node_off[i + 1] > node_off[i], "empty graph in batch");
    BatchStats local;
    BatchStats& S = st ? *st : local;
    S = BatchStats();
    S.cut.assign(G, 0);
    S.max_block_weight.assign(G, 0);
    S.coarsest_level.assign(G, 0);
    S.coarsest_n.assign(G, 0);
    S.upper_bound.assign(G, 0);
    hipStream_t s = eng.stream();
    const int k = p.k;

    // ------------------------------------------------------------------ per-graph bounds
    DevBuf<int> off_d, ng0;
    off_d.reserve(G + 1);
    KH_CHECK(hipMemcpyAsync(off_d.p, node_off, sizeof(int) * (G + 1), hipMemcpyHostToDevice, s));
    ng0.reserve(g0.n);
    hipLaunchKernelGGL(node_graph_kernel, dim3(grid_for(g0.n, 256)), dim3(256), 0, s, g0.n, G,
                       off_d.p, ng0.p);
    DevBuf<unsigned long long> Wg;
    Wg.reserve(G);
    fill_async(Wg.p, G, 0ull, s);
    hipLaunchKernelGGL(graph_weights, dim3(grid_for(g0.n, 256)), dim3(256), 0, s, g0.n, ng0.p,
                       g0.vwgt, Wg.p);
    std::vector<unsigned long long> hW(G);
    KH_CHECK(hipMemcpyAsync(hW.data(), Wg.p, sizeof(unsigned long long) * G,
                            hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    std::vector<int> Bg(G), Ug(G);
    for (int i = 0; i < G; ++i) {
        KH_REQUIRE(hW[i] < (1ULL << 31), "graph weight must be < 2^31");
        const unsigned U = upper_bound_for((long long)hW[i], k, p.imbalance);
        KH_REQUIRE(U < (1u << 31), "block bound exceeds int32");
        Ug[i] = (int)U;
        Bg[i] = (int)cluster_bound_for(U, p.cluster_factor);
        S.upper_bound[i] = U;
    }
    DevBuf<int> gbound;
    gbound.reserve(G);
    KH_CHECK(hipMemcpyAsync(gbound.p, Bg.data(), sizeof(int) * G, hipMemcpyHostToDevice, s));

    // ------------------------------------------------------------------ coarsening
    Ev ev_c;
    KH_CHECK(hipEventRecord(ev_c.a, s));
    std::vector<std::unique_ptr<OwnedGraph>> coarse;
    std::vector<DevBuf<int>> maps, ngs;       // ngs[l]: node -> graph at level l (l >= 1)
    std::vector<std::vector<int>> offs{std::vector<int>(node_off, node_off + G + 1)};
    auto level_graph = [&](int l) -> const DGraph& { return l == 0 ? g0 : coarse[l - 1]->g; };
    auto level_ng = [&](int l) -> const int* { return l == 0 ? ng0.p : ngs[l - 1].p; };
    std::vector<uint8_t> on(G, 1);
    DevBuf<uint8_t> on_d, include;
    on_d.reserve(G);
    DevBuf<int> label, size, nbound, coff;
    const unsigned num_stop = (unsigned)p.stop_factor * (unsigned)k;
    int level = 0;
    int n_on = G;
    while (n_on > 0) {
        KH_REQUIRE(level < kMaxLevels, "too many coarsening levels");
        const DGraph& fine = level_graph(level);
        const int* ng = level_ng(level);
        KH_CHECK(hipMemcpyAsync(on_d.p, on.data(), G, hipMemcpyHostToDevice, s));
        nbound.reserve(fine.n);
        include.reserve(fine.n);
        hipLaunchKernelGGL(node_tables, dim3(grid_for(fine.n, 256)), dim3(256), 0, s, fine.n, ng,
                           gbound.p, on_d.p, nbound.p, include.p);
        label.reserve(fine.n);
        size.reserve(fine.n);
        LpWorkspace::init_cluster_labels(fine, label.p, size.p, s);
        const uint32_t cseed = cluster_seed(p.seed, level);
        eng.lp.prepare(LpMode::Cluster, fine, fine.n, p.subrounds, cseed, p.bins, s, include.p);
        run_lp(eng.lp, LpMode::Cluster, fine, label.p, size.p, p.cluster_rounds, cseed,
               s, nbound.p, "lp_cluster");
        long long er = 0;
        for (int i = 0; i < p.subrounds * kNumBins; ++i) er += eng.lp.host_edges[i];
        S.edges_cluster += er * p.cluster_rounds;
        maps.emplace_back();
        maps.back().reserve(fine.n);
        const int nc = eng.cw.remap(fine, label.p, maps.back().p, s, nullptr);
        coarse.emplace_back(new OwnedGraph());
        eng.cw.contract(fine, maps.back().p, nc, *coarse.back(), s, nullptr);
        S.edges_contract += fine.m2;
        DGraph& cg = coarse.back()->g;
        cg.total_vwgt = fine.total_vwgt;
        coff.reserve(G + 1);
        hipLaunchKernelGGL(coarse_offsets, dim3(grid_for(G, 256)), dim3(256), 0, s, G, off_d.p,
                           maps.back().p, nc, coff.p);
        std::vector<int> hc(G + 1);
        KH_CHECK(hipMemcpyAsync(hc.data(), coff.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost,
                                s));
        KH_CHECK(hipStreamSynchronize(s));
        // the next level's offsets and node -> graph table
        off_d.reserve(G + 1);
        KH_CHECK(hipMemcpyAsync(off_d.p, hc.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice,
                                s));
        ngs.emplace_back();
        ngs.back().reserve(nc);
        hipLaunchKernelGGL(node_graph_kernel, dim3(grid_for(nc, 256)), dim3(256), 0, s, nc, G,
                           off_d.p, ngs.back().p);
        const std::vector<int>& fo = offs.back();
        ++level;
        n_on = 0;
        for (int i = 0; i < G; ++i) {
            if (!on[i]) continue;
            const unsigned nf = (unsigned)(fo[i + 1] - fo[i]);
            const unsigned ncg = (unsigned)(hc[i + 1] - hc[i]);
            const double rate = 1.0 * nf / (double)ncg;  // stop_rules.h:108-111
            if (rate >= 1.1 && ncg >= num_stop) {
                ++n_on;
            } else {
                on[i] = 0;
                S.coarsest_level[i] = level;
                S.coarsest_n[i] = (int)ncg;
            }
        }
        offs.push_back(hc);
        KH_CHECK(hipStreamSynchronize(s));  // hc must outlive its copy
    }
    S.levels = level;
    KH_CHECK(hipEventRecord(ev_c.b, s));
    KH_CHECK(hipStreamSynchronize(s));
    S.ms_coarsen = ev_c.ms();

    // ------------------------------------------------------------------ initial partitioning
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::vector<int>> ip_part(G);
    for (int l = 1; l <= level; ++l) {
        std::vector<int> todo;
        for (int i = 0; i < G; ++i)
            if (S.coarsest_level[i] == l) todo.push_back(i);
        if (todo.empty()) continue;
        HostGraph hg;
        download_graph(level_graph(l), hg, s);
        const std::vector<int>& o = offs[l];
        std::atomic<int> next{0};
        const int reps = p.ip_repetitions > 0 ? p.ip_repetitions : ip_repetitions_for(k);
        auto worker = [&]() {
            for (int t = next++; t < (int)todo.size(); t = next++) {
                const int i = todo[t];
                HostGraph sub;
                const int a = o[i], b = o[i + 1];
                sub.n = b - a;
                const int e0 = hg.xadj[a], e1 = hg.xadj[b];
                sub.xadj.resize(sub.n + 1);
                for (int v = 0; v <= sub.n; ++v) sub.xadj[v] = hg.xadj[a + v] - e0;
                sub.adjncy.resize(e1 - e0);
                for (int e = e0; e < e1; ++e) sub.adjncy[e - e0] = hg.adjncy[e] - a;
                sub.adjwgt.assign(hg.adjwgt.begin() + e0, hg.adjwgt.begin() + e1);
                sub.vwgt.assign(hg.vwgt.begin() + a, hg.vwgt.begin() + b);
                IpParams ip;
                ip.k = k;
                ip.block_bound = Ug[i];
                ip.repetitions = reps;
                ip.seed = ((uint64_t)(uint32_t)p.seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL) ^
                          ((uint64_t)hash3((uint32_t)p.seed, 0xB7u, (uint32_t)i) << 1);
                ip.threads = 1;
                ip.bisection_tries = p.ip_tries;
                initial_partition(sub, ip, ip_part[i]);
            }
        };
        unsigned hw = std::thread::hardware_concurrency();
        const int T = std::max(1, std::min<int>(p.ip_threads > 0 ? p.ip_threads
                                                                  : (int)std::min(16u, hw ? hw : 1u),
                                                (int)todo.size()));
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
    }
    S.ms_initial =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    // ------------------------------------------------------------------ uncoarsening
    Ev ev_r;
    KH_CHECK(hipEventRecord(ev_r.a, s));
    const long long L = (long long)G * k;
    std::vector<int> lb((size_t)L);
    for (int i = 0; i < G; ++i)
        for (int b = 0; b < k; ++b) lb[(size_t)i * k + b] = Ug[i];
    DevBuf<int> lbound_d, W, part_c, part_f, src;
    DevBuf<uint8_t> take;
    lbound_d.reserve(L);
    W.reserve(L);
    take.reserve(G);
    KH_CHECK(hipMemcpyAsync(lbound_d.p, lb.data(), sizeof(int) * L, hipMemcpyHostToDevice, s));
    for (int l = level; l >= 0; --l) {
        const DGraph& gl = level_graph(l);
        const int* ng = level_ng(l);
        part_f.reserve(std::max(gl.n, 1));
        if (l == level) {
            hipLaunchKernelGGL(fill_graph_base, dim3(grid_for(gl.n, 256)), dim3(256), 0, s, gl.n,
                               k, ng, part_f.p);
        } else {
            project_partition(gl.n, maps[l].p, part_c.p, part_f.p, s, nullptr);
        }
        std::swap(part_c, part_f);
        // graphs whose coarsest level is l take their initial partition here
        std::vector<uint8_t> tk(G, 0);
        std::vector<int> hsrc;
        bool any = false;
        for (int i = 0; i < G; ++i)
            if (S.coarsest_level[i] == l && l > 0) {
                tk[i] = 1;
                any = true;
            }
        if (any) {
            const std::vector<int>& o = offs[l];
            hsrc.assign(gl.n, 0);
            for (int i = 0; i < G; ++i)
                if (tk[i]) std::copy(ip_part[i].begin(), ip_part[i].end(), hsrc.begin() + o[i]);
            src.reserve(gl.n);
            KH_CHECK(hipMemcpyAsync(src.p, hsrc.data(), sizeof(int) * gl.n,
                                    hipMemcpyHostToDevice, s));
            KH_CHECK(hipMemcpyAsync(take.p, tk.data(), G, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(place_parts, dim3(grid_for(gl.n, 256)), dim3(256), 0, s, gl.n, k,
                               ng, src.p, take.p, part_c.p);
        }
        // refinement covers the graphs present at this level
        for (int i = 0; i < G; ++i) on[i] = S.coarsest_level[i] >= l ? 1 : 0;
        KH_CHECK(hipMemcpyAsync(on_d.p, on.data(), G, hipMemcpyHostToDevice, s));
        nbound.reserve(gl.n);
        include.reserve(gl.n);
        hipLaunchKernelGGL(node_tables, dim3(grid_for(gl.n, 256)), dim3(256), 0, s, gl.n, ng,
                           gbound.p, on_d.p, nbound.p, include.p);
        compute_block_weights(gl, part_c.p, (int)L, W.p, s);
        const uint32_t rseed = refine_seed(p.seed, l);
        eng.lp.prepare(LpMode::Refine, gl, (int)L, p.subrounds, rseed, p.bins, s, include.p);
        run_lp(eng.lp, LpMode::Refine, gl, part_c.p, W.p, p.refine_rounds, rseed, s,
               lbound_d.p, "lp_refine");
        S.edges_refine += eng.lp.traversed(s);
        KH_CHECK(hipStreamSynchronize(s));  // host vectors above must outlive their copies
    }
    KH_CHECK(hipEventRecord(ev_r.b, s));
    KH_CHECK(hipStreamSynchronize(s));
    S.ms_refine = ev_r.ms();

    // ------------------------------------------------------------------ metrics
    Ev ev_m;
    KH_CHECK(hipEventRecord(ev_m.a, s));
    DevBuf<unsigned long long> cut;
    cut.reserve(G);
    fill_async(cut.p, G, 0ull, s);
    if (g0.m2 > 0)
        hipLaunchKernelGGL(graph_cuts, dim3(grid_for((g0.m2 + 7) / 8, 256)), dim3(256), 0, s,
                           g0.n, g0.m2, g0.xadj, g0.adjncy, g0.adjwgt, part_c.p, ng0.p, cut.p);
    compute_block_weights(g0, part_c.p, (int)L, W.p, s);
    KH_CHECK(hipEventRecord(ev_m.b, s));
    S.edges_cut = g0.m2;
    std::vector<unsigned long long> hcut(G);
    std::vector<int> hw((size_t)L);
    KH_CHECK(hipMemcpyAsync(hcut.data(), cut.p, sizeof(unsigned long long) * G,
                            hipMemcpyDeviceToHost, s));
    KH_CHECK(hipMemcpyAsync(hw.data(), W.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
    part_f.reserve(g0.n);
    hipLaunchKernelGGL(to_local, dim3(grid_for(g0.n, 256)), dim3(256), 0, s, g0.n, k, ng0.p,
                       part_c.p, part_f.p);
    KH_CHECK(hipMemcpyAsync(part_out, part_f.p, sizeof(int) * g0.n, hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    KH_CHECK(hipGetLastError());
    S.ms_cut = ev_m.ms();
    for (int i = 0; i < G; ++i) {
        S.cut[i] = (long long)(hcut[i] / 2);
        long long mx = 0;
        for (int b = 0; b < k; ++b) mx = std::max(mx, (long long)hw[(size_t)i * k + b]);
        S.max_block_weight[i] = mx;
    }
}

}  // namespace kahip
