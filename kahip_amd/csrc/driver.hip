// Multilevel driver (reference: graph_partitioner.cpp:240-273 single_run for the FASTSOCIAL
// preset; coarsening.cpp:30-97; uncoarsening.cpp:43-121; interface/kaHIP_interface.cpp:95-166).
//
// Level loop, as in coarsening.cpp:59-88: cluster (synchronous LP on the GPU), relabel,
// contract, push; continue while the multiple-k stop rule holds (stop_rules.h:108-111). The
// coarsest graph is copied to the host for the CPU initial partitioning, then uncoarsening
// refines the coarsest level and, level by level, projects and refines (LP refinement on the
// GPU with the block bound U; balance_factor is 0 for FASTSOCIAL so U is used at every level,
// uncoarsening.cpp:65,84-85). The edge cut of the finest level is computed on the GPU.

#include "driver.hpp"
#include "fm_refine.hpp"

#include "reorder.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>

namespace kahip {

namespace {

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

__global__ void fill_ones(int* p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1;
}

// keys[v] = (a[v] << 32) | b[v], ids[v] = v
__global__ void pair_keys(int n, const int* __restrict__ a, const int* __restrict__ b,
                          unsigned long long* __restrict__ keys, int* __restrict__ ids) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    keys[v] = ((unsigned long long)(uint32_t)a[v] << 32) | (uint32_t)b[v];
    ids[v] = v;
}

// After a stable sort by pair key: every node gets the smallest node id of its pair group (the
// first element of its run).
__global__ void run_leaders(int n, const unsigned long long* __restrict__ skeys,
                            const int* __restrict__ sids, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i > 0 && skeys[i] == skeys[i - 1]) return;
    const int lead = sids[i];
    for (int j = i; j < n && skeys[j] == skeys[i]; ++j) out[sids[j]] = lead;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        KH_CHECK(hipEventCreate(&a));
        KH_CHECK(hipEventCreate(&b));
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    float ms() const {
        float t = 0;
        KH_CHECK(hipEventElapsedTime(&t, a, b));
        return t;
    }
};

}  // namespace

int ensemble_size_for(int k) {
    if (k <= 3) return 18;
    if (k <= 7) return 17;
    if (k <= 15) return 15;
    if (k <= 31) return 7;
    return 3;
}

unsigned upper_bound_for(long long total_weight, int k, double imbalance) {
    const double imb_percent = 100 * imbalance;
    const double epsilon = imb_percent / 100.0;
    const unsigned load = (unsigned)total_weight;
    if (imb_percent == 0) return (unsigned)((1 + epsilon + 0.01) * std::ceil(load / (double)k));
    return (unsigned)((1 + epsilon) * std::ceil(load / (double)k));
}

unsigned cluster_bound_for(unsigned upper, int factor) {
    return (unsigned)std::ceil(upper / (double)factor);
}

int ip_repetitions_for(int k) {
    int reps = 2;
    if (k > 1) reps = std::max((int)std::ceil(16 / std::log2((double)k)), 2);
    return std::min(4, reps);
}

void upload_graph(int n, const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                  OwnedGraph& out, hipStream_t s) {
    KH_REQUIRE(n >= 0, "n must be non-negative");
    out.alloc_nodes(n);
    const long long m2 = n > 0 ? (long long)xadj[n] : 0;
    out.alloc_edges(m2);
    KH_CHECK(hipMemcpyAsync(out.g.xadj, xadj, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (m2 > 0)
        KH_CHECK(hipMemcpyAsync(out.g.adjncy, adjncy, sizeof(int) * m2, hipMemcpyHostToDevice, s));
    long long total = 0;
    if (vwgt) {
        KH_CHECK(hipMemcpyAsync(out.g.vwgt, vwgt, sizeof(int) * n, hipMemcpyHostToDevice, s));
        for (int v = 0; v < n; ++v) total += vwgt[v];
    } else {
        if (n > 0) hipLaunchKernelGGL(fill_ones, dim3(grid_for(n, 256)), dim3(256), 0, s, out.g.vwgt, (long long)n);
        total = n;
    }
    if (adjwgt) {
        if (m2 > 0)
            KH_CHECK(hipMemcpyAsync(out.g.adjwgt, adjwgt, sizeof(int) * m2, hipMemcpyHostToDevice, s));
    } else if (m2 > 0) {
        hipLaunchKernelGGL(fill_ones, dim3(grid_for(m2, 256)), dim3(256), 0, s, out.g.adjwgt, m2);
    }
    int maxdeg = 0;
    for (int v = 0; v < n; ++v) maxdeg = std::max(maxdeg, xadj[v + 1] - xadj[v]);
    out.g.total_vwgt = total;
    out.g.max_degree = maxdeg;
    KH_CHECK(hipStreamSynchronize(s));
}

void download_graph(const DGraph& g, HostGraph& out, hipStream_t s) {
    out.n = g.n;
    out.xadj.resize(g.n + 1);
    out.adjncy.resize(g.m2);
    out.adjwgt.resize(g.m2);
    out.vwgt.resize(g.n);
    KH_CHECK(hipMemcpyAsync(out.xadj.data(), g.xadj, sizeof(int) * (g.n + 1), hipMemcpyDeviceToHost, s));
    if (g.m2 > 0) {
        KH_CHECK(hipMemcpyAsync(out.adjncy.data(), g.adjncy, sizeof(int) * g.m2, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipMemcpyAsync(out.adjwgt.data(), g.adjwgt, sizeof(int) * g.m2, hipMemcpyDeviceToHost, s));
    }
    if (g.n > 0)
        KH_CHECK(hipMemcpyAsync(out.vwgt.data(), g.vwgt, sizeof(int) * g.n, hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
}

Engine::Engine() { KH_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }

Engine::~Engine() {
    if (s_) (void)hipStreamDestroy(s_);
}

void Engine::partition(const DGraph& input, const PartitionParams& p, int* part_out,
                       PartitionStats* st) {
    KH_REQUIRE(p.k >= 1, "nparts must be >= 1");
    KH_REQUIRE(p.k <= 65536, "nparts must be <= 65536");
    KH_REQUIRE(p.imbalance >= 0, "imbalance must be >= 0");
    KH_REQUIRE(input.total_vwgt < (1LL << 31), "total node weight must be < 2^31");
    KH_REQUIRE(p.rating != RatingMode::Learned || p.mlp, "learned rating needs MLP weights");
    KH_REQUIRE(p.ensemble >= 1 && p.ensemble <= 64, "ensemble must be in [1, 64]");
    PartitionStats local;
    PartitionStats& S = st ? *st : local;
    S = PartitionStats();
    hipStream_t s = s_;
    Profiler* prof = p.profile ? &this->prof : nullptr;
    this->prof.enabled = p.profile != 0;
    this->prof.light = p.profile == 2;
    this->prof.reset();

    const unsigned U = upper_bound_for(input.total_vwgt, p.k, p.imbalance);
    const unsigned B = cluster_bound_for(U, p.cluster_factor);
    KH_REQUIRE(U < (1u << 31), "block bound exceeds int32");
    S.upper_bound = U;
    S.cluster_bound = B;

    // ------------------------------------------------------------------ coarsening
    // Skewed graphs are renumbered by degree first (part of the coarsening time); the pipeline
    // runs on the renumbered graph and the partition is mapped back at the end.
    OwnedGraph& reordered = buf_.reordered;
    DevBuf<int>& inv = buf_.inv;
    const bool reorder = reorder_wanted(input, p.reorder);
    EventPair ev_coarsen;
    KH_CHECK(hipEventRecord(ev_coarsen.a, s));
    if (reorder) reorder_by_degree(input, reordered, inv, buf_.rtmp, s, prof);
    const DGraph& g0 = reorder ? reordered.g : input;
    S.reordered = reorder;
    auto& coarse = buf_.coarse;  // coarse[l] = level l+1 (entries past the last level: unused)
    auto& maps = buf_.maps;      // maps[l]: level l -> level l+1
    auto level_graph = [&](int l) -> const DGraph& { return l == 0 ? g0 : coarse[l - 1]->g; };
    DevBuf<int>& label = buf_.label;
    DevBuf<int>& size = buf_.size;
    DevBuf<int>& score = buf_.score;
    DevBuf<int>& elabel = buf_.elabel;
    // label := the smallest node id among the nodes sharing (label, other)
    DevBuf<unsigned long long>& okeys = buf_.okeys;
    DevBuf<unsigned long long>& okeys2 = buf_.okeys2;
    DevBuf<int>& oids = buf_.oids;
    DevBuf<int>& oids2 = buf_.oids2;
    DevBuf<char>& otmp = buf_.otmp;
    auto overlay = [&](int n, int* lab, const int* other, hipStream_t st) {
        okeys.reserve(n);
        okeys2.reserve(n);
        oids.reserve(n);
        oids2.reserve(n);
        hipLaunchKernelGGL(pair_keys, dim3(grid_for(n, 256)), dim3(256), 0, st, n, lab, other,
                           okeys.p, oids.p);
        size_t bytes = 0;
        KH_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, okeys.p, okeys2.p, oids.p, oids2.p,
                                           (size_t)n, 0, 64, st));
        otmp.reserve((long long)bytes + 1);
        KH_CHECK(rocprim::radix_sort_pairs(otmp.p, bytes, okeys.p, okeys2.p, oids.p, oids2.p,
                                           (size_t)n, 0, 64, st));
        hipLaunchKernelGGL(run_leaders, dim3(grid_for(n, 256)), dim3(256), 0, st, n, okeys2.p,
                           oids2.p, lab);
        KH_CHECK(hipGetLastError());
    };
    const unsigned num_stop = (unsigned)p.stop_factor * (unsigned)p.k;
    int level = 0;
    S.level_n[0] = g0.n;
    S.level_m2[0] = g0.m2;
    bool again = g0.n > 0;
    while (again) {
        KH_REQUIRE(level < kMaxLevels, "too many coarsening levels");
        const DGraph& fine = level_graph(level);
        label.reserve(fine.n);
        size.reserve(fine.n);
        LpWorkspace::init_cluster_labels(fine, label.p, size.p, s);
        const uint32_t cseed = cluster_seed(p.seed, level);
        // The clustering sees the rating in place of the edge weights; contraction keeps them.
        DGraph rated = fine;
        if (p.rating != RatingMode::None && level < p.rating_levels && fine.m2 > 0) {
            EventPair ev;
            KH_CHECK(hipEventRecord(ev.a, s));
            score.reserve(fine.m2);
            rw.scores(p.rating, fine, (int)B, rating_seed(p.seed, level), p.mlp, score.p, s, lp,
                      prof);
            KH_CHECK(hipEventRecord(ev.b, s));
            KH_CHECK(hipEventSynchronize(ev.b));
            S.ms_rating += ev.ms();
            rated.adjwgt = score.p;
        }
        ProfScope ps_cluster(prof, "lp_cluster_total", 0.0, 0, s);  // the level's clustering
        lp.prepare(LpMode::Cluster, rated, fine.n, p.subrounds, cseed, p.bins, s);
        lp.run(LpMode::Cluster, rated, label.p, size.p, (int)B, p.cluster_rounds, cseed, s, prof,
               "lp_cluster");
        lp.check(s);
        long long edges_round = 0;
        for (int i = 0; i < p.subrounds * kNumBins; ++i) edges_round += lp.host_edges[i];
        S.edges_cluster += edges_round * p.cluster_rounds;
        for (int run = 1; run < p.ensemble; ++run) {
            // another clustering with its own seed and cluster factor, overlaid on the result
            const uint32_t eseed = ensemble_seed(p.seed, level, run);
            const int eb = (int)cluster_bound_for(U, ensemble_factor(p.seed, level, run));
            elabel.reserve(fine.n);
            LpWorkspace::init_cluster_labels(fine, elabel.p, size.p, s);
            lp.prepare(LpMode::Cluster, rated, fine.n, p.subrounds, eseed, p.bins, s);
            lp.run(LpMode::Cluster, rated, elabel.p, size.p, eb, p.cluster_rounds, eseed, s, prof,
                   "lp_cluster");
            lp.check(s);
            S.edges_cluster += edges_round * p.cluster_rounds;
            overlay(fine.n, label.p, elabel.p, s);
        }
        ps_cluster.close();
        if ((int)maps.size() <= level) maps.emplace_back();
        if ((int)coarse.size() <= level) coarse.emplace_back(new OwnedGraph());
        maps[level].reserve(fine.n);
        const int nc = cw.remap(fine, label.p, maps[level].p, s, prof);
        cw.contract(fine, maps[level].p, nc, *coarse[level], s, prof);
        S.edges_contract += fine.m2;
        DGraph& cg = coarse[level]->g;
        cg.total_vwgt = fine.total_vwgt;
        ++level;
        S.level_n[level] = cg.n;
        S.level_m2[level] = cg.m2;
        // stop_rules.h:108-111
        const double rate = 1.0 * (unsigned)fine.n / (double)(unsigned)nc;
        again = rate >= 1.1 && (unsigned)nc >= num_stop;
    }
    S.levels = level;
    KH_CHECK(hipEventRecord(ev_coarsen.b, s));
    KH_CHECK(hipStreamSynchronize(s));
    S.ms_coarsen = ev_coarsen.ms();

    // ------------------------------------------------------------------ initial partitioning
    const DGraph& cg = level_graph(level);
    const auto t0 = std::chrono::steady_clock::now();
    HostGraph hg;
    download_graph(cg, hg, s);
    std::vector<int> hpart;
    IpParams ip;
    ip.k = p.k;
    ip.block_bound = U;
    ip.repetitions = p.ip_repetitions > 0 ? p.ip_repetitions : ip_repetitions_for(p.k);
    ip.seed = (uint64_t)(uint32_t)p.seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL;
    ip.threads = p.ip_threads;
    ip.bisection_tries = p.ip_tries;
    initial_partition(hg, ip, hpart);
    DevBuf<int>& part_c = buf_.part_c;
    DevBuf<int>& part_f = buf_.part_f;
    part_c.reserve(std::max(cg.n, 1));
    if (cg.n > 0)
        KH_CHECK(hipMemcpyAsync(part_c.p, hpart.data(), sizeof(int) * cg.n, hipMemcpyHostToDevice, s));
    KH_CHECK(hipStreamSynchronize(s));
    S.ms_initial = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    // ------------------------------------------------------------------ uncoarsening
    // Every level: projection, LP refinement on the GPU (timed by an event pair per level) and,
    // for the FM-based presets, the CPU pairwise FM of the level's partition, as the reference
    // refines every level after projection (uncoarsening.cpp:58-95).
    DevBuf<int>& W = buf_.W;
    W.reserve(p.k);
    HostGraph hl;
    std::vector<int> hp;
    S.ms_refine = 0;
    for (int l = level; l >= 0; --l) {
        const DGraph& gl = level_graph(l);
        EventPair ev;
        KH_CHECK(hipEventRecord(ev.a, s));
        int* part_l;
        if (l == level) {
            part_l = part_c.p;
        } else {
            part_f.reserve(std::max(gl.n, 1));
            project_partition(gl.n, maps[l].p, part_c.p, part_f.p, s, prof);
            std::swap(part_c, part_f);
            part_l = part_c.p;
        }
        compute_block_weights(gl, part_l, p.k, W.p, s);
        const uint32_t rseed = refine_seed(p.seed, l);
        lp.prepare(LpMode::Refine, gl, p.k, p.subrounds, rseed, p.bins, s);
        // the profiler times the clustering and contraction kernels, not the refinement rounds
        lp.run(LpMode::Refine, gl, part_l, W.p, (int)U, p.refine_rounds, rseed, s, nullptr,
               "lp_refine");
        KH_CHECK(hipEventRecord(ev.b, s));
        lp.check(s);
        S.edges_refine += lp.traversed(s);
        S.ms_refine += ev.ms();
        if (p.fm_rounds > 0 && gl.n > 0) {
            const auto t1 = std::chrono::steady_clock::now();
            download_graph(gl, hl, s);
            hp.resize(gl.n);
            KH_CHECK(hipMemcpyAsync(hp.data(), part_l, sizeof(int) * gl.n, hipMemcpyDeviceToHost,
                                    s));
            KH_CHECK(hipStreamSynchronize(s));
            FmRefineParams fp;
            fp.k = p.k;
            fp.block_bound = U;
            fp.rounds = p.fm_rounds;
            fp.seed = fm_seed(p.seed) + (uint64_t)l;
            S.fm_gain += fm_refine_pairs(hl, fp, hp);
            KH_CHECK(hipMemcpyAsync(part_l, hp.data(), sizeof(int) * gl.n, hipMemcpyHostToDevice,
                                    s));
            KH_CHECK(hipStreamSynchronize(s));
            S.ms_fm += std::chrono::duration<double, std::milli>(
                           std::chrono::steady_clock::now() - t1).count();
        }
    }

    // ------------------------------------------------------------------ metrics
    EventPair ev_cut;
    KH_CHECK(hipEventRecord(ev_cut.a, s));
    S.cut = mw.edge_cut(g0, part_c.p, s, prof);
    KH_CHECK(hipEventRecord(ev_cut.b, s));
    S.edges_cut = g0.m2;
    compute_block_weights(g0, part_c.p, p.k, W.p, s);
    std::vector<int> hw(p.k);
    KH_CHECK(hipMemcpyAsync(hw.data(), W.p, sizeof(int) * p.k, hipMemcpyDeviceToHost, s));
    if (g0.n > 0) {
        const int* out_part = part_c.p;
        if (reorder) {
            part_f.reserve(g0.n);
            unpermute_parts(g0.n, inv.p, part_c.p, part_f.p, s);
            out_part = part_f.p;
        }
        KH_CHECK(hipMemcpyAsync(part_out, out_part, sizeof(int) * g0.n, hipMemcpyDeviceToHost, s));
    }
    KH_CHECK(hipStreamSynchronize(s));
    S.ms_cut = ev_cut.ms();
    long long mx = 0;
    for (int b = 0; b < p.k; ++b) mx = std::max(mx, (long long)hw[b]);
    S.max_block_weight = mx;
    // quality_metrics.cpp:265-288
    S.balance = g0.n > 0 ? (double)mx / std::ceil((double)g0.total_vwgt / (double)p.k) : 0.0;
}

}  // namespace kahip
