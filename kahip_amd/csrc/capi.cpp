// extern "C" entry points of libkahip.so (declared in include/kaHIP_interface.h and
// include/kahip_amd.h). Host arrays are copied to the GPU, the work runs on a stream, results
// are copied back. State that survives a call: the thread-local last error and last profile
// summary, and the partition entry points' cached engine per device (its stream and device
// workspaces, reused by the next call on that device; kahip_release_workspaces frees them).

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#include "batch.hpp"
#include "driver.hpp"
#include "fm_refine.hpp"
#include "gen.hpp"
#include "rebalance.hpp"
#include "kahip_amd.h"

using namespace kahip;

struct kahip_graph {
    OwnedGraph og;
};

struct kahip_batch {
    OwnedGraph og;
    std::vector<int> node_off;
};

namespace {

thread_local std::string g_last_error;
thread_local std::string g_last_profile = "{}";

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return 1;
}

// The partition entry points (kaffpa, kahip_partition, kahip_batch_partition) run on a cached
// engine of the current device, so that a repeated call reuses the device workspaces sized by
// the previous ones instead of allocating (and freeing) them inside the call; every workspace is
// re-initialised by the kernels that use it, exactly as between two levels of one call. A call
// that finds the device's engine held by another thread runs on an engine of its own, and a call
// that fails drops the cached engine. The cache is never destroyed at exit (its memory goes with
// the process); kahip_release_workspaces frees the idle engines.
struct CachedEngine {
    std::mutex mu;
    std::unique_ptr<Engine> eng;
};
std::mutex g_cache_mu;
std::map<int, CachedEngine*>& engine_cache() {
    static auto* cache = new std::map<int, CachedEngine*>();  // left to the process exit
    return *cache;
}

class EngineLease {
public:
    EngineLease() : unwinding_(std::uncaught_exceptions()) {
        int dev = 0;
        KH_CHECK(hipGetDevice(&dev));
        CachedEngine* c;
        {
            std::lock_guard<std::mutex> g(g_cache_mu);
            CachedEngine*& slot = engine_cache()[dev];
            if (!slot) slot = new CachedEngine();
            c = slot;
        }
        std::unique_lock<std::mutex> lk(c->mu, std::try_to_lock);
        if (lk.owns_lock()) {
            if (!c->eng) c->eng.reset(new Engine());
            cached_ = c;
            e_ = c->eng.get();
            lk_ = std::move(lk);
        } else {
            own_.reset(new Engine());
            e_ = own_.get();
        }
    }
    ~EngineLease() {
        if (cached_ && std::uncaught_exceptions() > unwinding_) cached_->eng.reset();
    }
    EngineLease(const EngineLease&) = delete;
    EngineLease& operator=(const EngineLease&) = delete;
    Engine& operator*() { return *e_; }
    Engine* operator->() { return e_; }

private:
    int unwinding_;
    CachedEngine* cached_ = nullptr;
    Engine* e_ = nullptr;
    std::unique_ptr<Engine> own_;
    std::unique_lock<std::mutex> lk_;
};

PartitionParams to_params(const kahip_params* q) {
    kahip_params d;
    kahip_default_params(&d);
    if (!q) q = &d;
    PartitionParams p;
    p.k = q->k;
    p.imbalance = q->imbalance;
    p.seed = q->seed;
    p.cluster_rounds = q->cluster_rounds;
    p.refine_rounds = q->refine_rounds;
    p.subrounds = q->subrounds;
    p.cluster_factor = q->cluster_factor;
    p.stop_factor = q->stop_factor;
    p.ip_repetitions = q->ip_repetitions;
    p.ip_threads = q->ip_threads;
    for (int b = 0; b < kNumBins; ++b) p.bins.hi[b] = q->bins[b];
    KH_REQUIRE(q->group_items == 1 || q->group_items == 2 || q->group_items == 4,
               "group_items must be 1, 2 or 4");
    KH_REQUIRE(q->wave_small == 64 || q->wave_small == 128 || q->wave_small == 256,
               "wave_small must be 64, 128 or 256");
    p.bins.group_items = q->group_items;
    p.bins.wave_small = q->wave_small;
    KH_REQUIRE(q->profile >= 0 && q->profile <= 2, "profile must be 0, 1 or 2");
    p.profile = q->profile;
    KH_REQUIRE(q->rating >= 0 && q->rating <= 2, "bad rating");
    p.rating = (RatingMode)q->rating;
    p.rating_levels = q->rating_levels;
    p.mlp = reinterpret_cast<const MlpWeights*>(q->mlp_weights);
    p.reorder = q->reorder;
    p.ensemble = q->ensemble;
    p.ip_tries = q->ip_tries;
    KH_REQUIRE(p.ip_tries >= 1, "ip_tries must be >= 1");
    p.fm_rounds = q->fm_rounds;
    KH_REQUIRE(p.fm_rounds >= 0, "fm_rounds must be >= 0");
    KH_REQUIRE(q->graph_refine == 0,
               "hipGraph-captured LP rounds are disabled: the captured path hit a host "
               "segmentation fault under rocprofv3 whose cause was not found");
    p.graph_refine = 0;
    KH_REQUIRE(p.rating != RatingMode::Learned || p.mlp, "learned rating needs mlp_weights");
    KH_REQUIRE(p.bins.hi[4] <= 512, "bins[4] must be <= 512 (LDS hash capacity)");
    for (int b = 0; b + 1 < kNumBins; ++b)
        KH_REQUIRE(p.bins.hi[b] < p.bins.hi[b + 1], "bins must be strictly increasing");
    KH_REQUIRE(p.bins.hi[0] == 8 && p.bins.hi[1] == 16 && p.bins.hi[2] == 32 &&
                   p.bins.hi[3] == 64,
               "bins[0..3] are the lane-group widths 8, 16, 32, 64");
    KH_REQUIRE(p.cluster_rounds >= 0 && p.cluster_rounds <= kMaxRounds, "bad cluster_rounds");
    KH_REQUIRE(p.refine_rounds >= 0 && p.refine_rounds <= kMaxRounds, "bad refine_rounds");
    KH_REQUIRE(p.cluster_factor >= 1, "cluster_factor must be >= 1");
    return p;
}

void check_csr(int n, const int* xadj, const int* adjncy) {
    KH_REQUIRE(n >= 0, "n must be non-negative");
    KH_REQUIRE(xadj != nullptr, "xadj is NULL");
    KH_REQUIRE(xadj[0] == 0, "xadj[0] must be 0");
    for (int v = 0; v < n; ++v) KH_REQUIRE(xadj[v + 1] >= xadj[v], "xadj must be non-decreasing");
    const long long m2 = xadj[n];
    KH_REQUIRE(m2 == 0 || adjncy != nullptr, "adjncy is NULL");
    for (long long e = 0; e < m2; ++e)
        KH_REQUIRE(adjncy[e] >= 0 && adjncy[e] < n, "adjncy entry out of range");
}

std::string profile_json(const Profiler& prof) {
    std::ostringstream os;
    os << "{";
    bool first = true;
    for (const auto& kv : prof.summarize()) {
        if (!first) os << ",";
        first = false;
        os << "\"" << kv.first << "\":{\"launches\":" << kv.second.launches
           << ",\"ms\":" << kv.second.ms << ",\"bytes\":" << kv.second.bytes
           << ",\"edges\":" << kv.second.edges << "}";
    }
    os << "}";
    return os.str();
}

void fill_stats(const PartitionStats& s, kahip_stats* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->cut = s.cut;
    o->balance = s.balance;
    o->upper_bound = s.upper_bound;
    o->cluster_bound = s.cluster_bound;
    o->levels = s.levels;
    for (int i = 0; i <= kMaxLevels; ++i) {
        o->level_n[i] = s.level_n[i];
        o->level_m2[i] = s.level_m2[i];
    }
    o->ms_coarsen = s.ms_coarsen;
    o->ms_initial = s.ms_initial;
    o->ms_refine = s.ms_refine;
    o->ms_cut = s.ms_cut;
    o->edges_cluster = s.edges_cluster;
    o->edges_contract = s.edges_contract;
    o->edges_refine = s.edges_refine;
    o->edges_cut = s.edges_cut;
    o->max_block_weight = s.max_block_weight;
    o->ms_rating = s.ms_rating;
    o->reordered = s.reordered ? 1 : 0;
    o->ms_fm = s.ms_fm;
    o->fm_gain = s.fm_gain;
}

int run_kaffpa(int n, const int* vwgt, const int* xadj, const int* adjcwgt, const int* adjncy,
               const kahip_params* params, int* edgecut, int* part, kahip_stats* stats) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        PartitionParams p = to_params(params);
        EngineLease eng;
        OwnedGraph g;
        upload_graph(n, xadj, adjncy, vwgt, adjcwgt, g, eng->stream());
        PartitionStats st;
        eng->partition(g.g, p, part, &st);
        if (p.profile) g_last_profile = profile_json(eng->prof);
        if (edgecut) *edgecut = (int)st.cut;
        fill_stats(st, stats);
    });
}

void report(const char* fn, int* edgecut) {
    std::fprintf(stderr, "%s: %s\n", fn, g_last_error.c_str());
    if (edgecut) *edgecut = -1;
}

// The drop-in entry points' common path: the preset of `mode` (an unknown mode falls back to
// ECO, interface/kaHIP_interface.cpp:60-62), the pipeline, and the perfectly balanced post-pass
// (kaHIP_interface.cpp:123-132). With imbalance 0 the pipeline runs with the bound
// 1.01 ceil(W / k) (app/balance_configuration.h:41-43) and the post-pass enforces ceil(W / k);
// with perfectly_balance it enforces (1 + imbalance) ceil(W / k).
void dropin_call(const char* fn, int n, const int* vwgt, const int* xadj, const int* adjcwgt,
                 const int* adjncy, int nparts, double imbalance, bool perfectly_balance,
                 int seed, int mode, int* edgecut, int* part) {
    if (mode < FAST || mode > STRONGSOCIAL) mode = ECO;
    kahip_params p;
    kahip_default_params(&p);
    p.k = nparts;
    p.imbalance = imbalance;
    p.seed = seed;
    // STRONGSOCIAL clusters with ensembles (app/configuration.h:584-591); the presets whose
    // reference refinement is FM-based (ECO, STRONG, ECOSOCIAL, STRONGSOCIAL) add the CPU
    // pairwise FM on the finest level after the LP refinement
    if (mode == STRONGSOCIAL) p.ensemble = ensemble_size_for(p.k);
    if (mode == ECO || mode == STRONG || mode == ECOSOCIAL || mode == STRONGSOCIAL)
        p.fm_rounds = kPresetFmRounds;
    if (run_kaffpa(n, vwgt, xadj, adjcwgt, adjncy, &p, edgecut, part, nullptr)) {
        report(fn, edgecut);
        return;
    }
    if (imbalance == 0 || perfectly_balance) {
        long long W = 0;
        for (int v = 0; v < n; ++v) W += vwgt ? vwgt[v] : 1;
        const long long bound =
            (long long)((1 + imbalance) * std::ceil((double)W / (double)nparts));
        const long long cut = balance_postpass(n, xadj, adjncy, vwgt, adjcwgt, nparts, bound,
                                               fm_seed(seed) ^ 0xBA1A'4CE5ULL, part);
        if (edgecut) *edgecut = (int)cut;
    }
}

}  // namespace

extern "C" {

void kahip_default_params(kahip_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    PartitionParams d;
    p->k = d.k;
    p->imbalance = d.imbalance;
    p->seed = d.seed;
    p->cluster_rounds = d.cluster_rounds;
    p->refine_rounds = d.refine_rounds;
    p->subrounds = d.subrounds;
    p->cluster_factor = d.cluster_factor;
    p->stop_factor = d.stop_factor;
    p->ip_repetitions = d.ip_repetitions;
    p->ip_threads = d.ip_threads;
    for (int b = 0; b < kNumBins; ++b) p->bins[b] = d.bins.hi[b];
    p->group_items = d.bins.group_items;
    p->wave_small = d.bins.wave_small;
    p->profile = 0;
    p->rating = (int)d.rating;
    p->rating_levels = d.rating_levels;
    p->mlp_weights = nullptr;
    p->reorder = d.reorder;
    p->ensemble = d.ensemble;
    p->ip_tries = d.ip_tries;
    p->fm_rounds = d.fm_rounds;
    p->graph_refine = d.graph_refine;
}

const char* kahip_last_error(void) { return g_last_error.c_str(); }

int kaffpa_gpu_ex(int n, const int* vwgt, const int* xadj, const int* adjcwgt, const int* adjncy,
                  const kahip_params* params, int* edgecut, int* part, kahip_stats* stats) {
    return run_kaffpa(n, vwgt, xadj, adjcwgt, adjncy, params, edgecut, part, stats);
}

void kaffpa(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
            double* imbalance, bool suppress_output, int seed, int mode, int* edgecut,
            int* part) {
    (void)suppress_output;  // ignored by the reference too (kaHIP_interface.cpp:108-113)
    dropin_call("kaffpa", *n, vwgt, xadj, adjcwgt, adjncy, *nparts, *imbalance, false, seed,
                mode, edgecut, part);
}

void kaffpa_balance(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
                    double* imbalance, bool perfectly_balance, bool suppress_output, int seed,
                    int mode, int* edgecut, int* part) {
    (void)suppress_output;
    dropin_call("kaffpa_balance", *n, vwgt, xadj, adjcwgt, adjncy, *nparts, *imbalance,
                perfectly_balance, seed, mode, edgecut, part);
}

void kaffpa_balance_NE(int* n, int* vwgt, int* xadj, int* adjcwgt, int* adjncy, int* nparts,
                       double* imbalance, bool suppress_output, int seed, int mode, int* edgecut,
                       int* part) {
    (void)suppress_output;
    // app/balance_configuration.h:27-37: node weight += weighted degree, except in the perfectly
    // balanced case (imbalance 0), which keeps the node weights
    if (*imbalance == 0) {
        dropin_call("kaffpa_balance_NE", *n, vwgt, xadj, adjcwgt, adjncy, *nparts, *imbalance,
                    false, seed, mode, edgecut, part);
        return;
    }
    std::vector<int> w(std::max(*n, 0));
    for (int v = 0; v < *n; ++v) {
        long long s = vwgt ? vwgt[v] : 1;
        for (int e = xadj[v]; e < xadj[v + 1]; ++e) s += adjcwgt ? adjcwgt[e] : 1;
        if (s > INT_MAX) {
            g_last_error = "kaffpa_balance_NE: node weight plus weighted degree exceeds INT_MAX";
            report("kaffpa_balance_NE", edgecut);
            return;
        }
        w[v] = (int)s;
    }
    dropin_call("kaffpa_balance_NE", *n, w.data(), xadj, adjcwgt, adjncy, *nparts, *imbalance,
                false, seed, mode, edgecut, part);
}

int kahip_rebalance(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, long long bound, unsigned long long seed,
                    int* part, long long* cut) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        KH_REQUIRE(k >= 1, "k must be >= 1");
        for (int v = 0; v < n; ++v) KH_REQUIRE(part[v] >= 0 && part[v] < k, "part out of range");
        const long long c = balance_postpass(n, xadj, adjncy, vwgt, adjwgt, k, bound, seed, part);
        if (cut) *cut = c;
    });
}

int kahip_graph_from_host(int n, const int* xadj, const int* adjncy, const int* vwgt,
                          const int* adjwgt, kahip_graph_t* out) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        auto* g = new kahip_graph();
        hipStream_t s;
        KH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            upload_graph(n, xadj, adjncy, vwgt, adjwgt, g->og, s);
        } catch (...) {
            (void)hipStreamDestroy(s);
            delete g;
            throw;
        }
        KH_CHECK(hipStreamDestroy(s));
        *out = g;
    });
}

int kahip_graph_rmat(int scale, long long samples, unsigned long long seed, double a, double b,
                     double c, kahip_graph_t* out) {
    return guarded([&] {
        auto* g = new kahip_graph();
        try {
            generate_rmat(scale, samples, seed, a, b, c, g->og);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int kahip_graph_rgg(int n, unsigned long long seed, double radius, kahip_graph_t* out) {
    return guarded([&] {
        auto* g = new kahip_graph();
        try {
            generate_rgg(n, seed, radius, g->og);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int kahip_graph_info(kahip_graph_t g, int* n, long long* m2, long long* total_vwgt,
                     int* max_degree) {
    return guarded([&] {
        KH_REQUIRE(g, "null graph");
        if (n) *n = g->og.g.n;
        if (m2) *m2 = g->og.g.m2;
        if (total_vwgt) *total_vwgt = g->og.g.total_vwgt;
        if (max_degree) *max_degree = g->og.g.max_degree;
    });
}

int kahip_graph_to_host(kahip_graph_t g, int* xadj, int* adjncy, int* vwgt, int* adjwgt) {
    return guarded([&] {
        KH_REQUIRE(g, "null graph");
        const DGraph& d = g->og.g;
        if (xadj) KH_CHECK(hipMemcpy(xadj, d.xadj, sizeof(int) * (d.n + 1), hipMemcpyDeviceToHost));
        if (adjncy && d.m2) KH_CHECK(hipMemcpy(adjncy, d.adjncy, sizeof(int) * d.m2, hipMemcpyDeviceToHost));
        if (adjwgt && d.m2) KH_CHECK(hipMemcpy(adjwgt, d.adjwgt, sizeof(int) * d.m2, hipMemcpyDeviceToHost));
        if (vwgt && d.n) KH_CHECK(hipMemcpy(vwgt, d.vwgt, sizeof(int) * d.n, hipMemcpyDeviceToHost));
    });
}

void kahip_graph_free(kahip_graph_t g) { delete g; }

int kahip_release_workspaces(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(g_cache_mu);
        for (auto& kv : engine_cache()) {
            std::unique_lock<std::mutex> lk(kv.second->mu, std::try_to_lock);
            if (lk.owns_lock()) kv.second->eng.reset();  // an engine in use stays
        }
    });
}

int kahip_partition(kahip_graph_t g, const kahip_params* params, int* part, kahip_stats* stats) {
    return guarded([&] {
        KH_REQUIRE(g, "null graph");
        PartitionParams p = to_params(params);
        EngineLease eng;
        std::vector<int> tmp;
        if (!part) {
            tmp.resize(std::max(g->og.g.n, 1));
            part = tmp.data();
        }
        PartitionStats st;
        eng->partition(g->og.g, p, part, &st);
        if (p.profile) g_last_profile = profile_json(eng->prof);
        fill_stats(st, stats);
    });
}

int kahip_profile_json(char* buf, int len) {
    if (!buf || len <= 0) return (int)g_last_profile.size() + 1;
    std::snprintf(buf, (size_t)len, "%s", g_last_profile.c_str());
    return (int)g_last_profile.size() + 1 <= len ? 0 : (int)g_last_profile.size() + 1;
}

int kahip_lp_cluster(int n, const int* xadj, const int* adjncy, const int* vwgt,
                     const int* adjwgt, int bound, int rounds, int subrounds, unsigned seed,
                     int* cluster, int* num_clusters, int* moves_per_round) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        Engine eng;
        hipStream_t s = eng.stream();
        OwnedGraph g;
        upload_graph(n, xadj, adjncy, vwgt, adjwgt, g, s);
        DevBuf<int> label, size, map;
        label.reserve(n);
        size.reserve(n);
        map.reserve(n);
        LpWorkspace::init_cluster_labels(g.g, label.p, size.p, s);
        BinConfig bins;
        eng.lp.prepare(LpMode::Cluster, g.g, n, subrounds, seed, bins, s);
        eng.lp.run(LpMode::Cluster, g.g, label.p, size.p, bound, rounds, seed, s, nullptr,
                   "lp_cluster");
        eng.lp.check(s);
        if (moves_per_round) eng.lp.read_moves(rounds, moves_per_round, s);
        const int nc = eng.cw.remap(g.g, label.p, map.p, s, nullptr);
        if (n) KH_CHECK(hipMemcpyAsync(cluster, map.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        if (num_clusters) *num_clusters = nc;
    });
}

int kahip_remap(int n, const int* label, int* map, int* num_clusters) {
    return guarded([&] {
        KH_REQUIRE(n >= 0, "n must be non-negative");
        for (int v = 0; v < n; ++v) KH_REQUIRE(label[v] >= 0 && label[v] < n, "label out of range");
        Engine eng;
        hipStream_t s = eng.stream();
        DevBuf<int> dl, dm;
        dl.reserve(n);
        dm.reserve(n);
        if (n) KH_CHECK(hipMemcpyAsync(dl.p, label, sizeof(int) * n, hipMemcpyHostToDevice, s));
        DGraph g;
        g.n = n;
        const int nc = eng.cw.remap(g, dl.p, dm.p, s, nullptr);
        if (n) KH_CHECK(hipMemcpyAsync(map, dm.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        if (num_clusters) *num_clusters = nc;
    });
}

int kahip_contract(int n, const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                   const int* map, int nc, int* cxadj, int* cadjncy, int* cadjwgt, int* cvwgt) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        KH_REQUIRE(nc >= 0 && (n == 0 || nc >= 1), "bad number of clusters");
        for (int v = 0; v < n; ++v) KH_REQUIRE(map[v] >= 0 && map[v] < nc, "map out of range");
        Engine eng;
        hipStream_t s = eng.stream();
        OwnedGraph g, c;
        upload_graph(n, xadj, adjncy, vwgt, adjwgt, g, s);
        DevBuf<int> dm;
        dm.reserve(n);
        if (n) KH_CHECK(hipMemcpyAsync(dm.p, map, sizeof(int) * n, hipMemcpyHostToDevice, s));
        eng.cw.contract(g.g, dm.p, nc, c, s, nullptr);
        KH_CHECK(hipMemcpyAsync(cxadj, c.g.xadj, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost, s));
        if (c.g.m2) {
            KH_CHECK(hipMemcpyAsync(cadjncy, c.g.adjncy, sizeof(int) * c.g.m2, hipMemcpyDeviceToHost, s));
            KH_CHECK(hipMemcpyAsync(cadjwgt, c.g.adjwgt, sizeof(int) * c.g.m2, hipMemcpyDeviceToHost, s));
        }
        if (nc) KH_CHECK(hipMemcpyAsync(cvwgt, c.g.vwgt, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_lp_refine(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, int bound, int rounds, int subrounds, unsigned seed,
                    int* part, int* moves_per_round) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        KH_REQUIRE(k >= 1 && k <= 65536, "k must be in [1, 65536]");
        for (int v = 0; v < n; ++v) KH_REQUIRE(part[v] >= 0 && part[v] < k, "part out of range");
        Engine eng;
        hipStream_t s = eng.stream();
        OwnedGraph g;
        upload_graph(n, xadj, adjncy, vwgt, adjwgt, g, s);
        DevBuf<int> dp, W;
        dp.reserve(n);
        W.reserve(k);
        if (n) KH_CHECK(hipMemcpyAsync(dp.p, part, sizeof(int) * n, hipMemcpyHostToDevice, s));
        compute_block_weights(g.g, dp.p, k, W.p, s);
        BinConfig bins;
        eng.lp.prepare(LpMode::Refine, g.g, k, subrounds, seed, bins, s);
        eng.lp.run(LpMode::Refine, g.g, dp.p, W.p, bound, rounds, seed, s, nullptr, "lp_refine");
        eng.lp.check(s);
        if (moves_per_round) eng.lp.read_moves(rounds, moves_per_round, s);
        if (n) KH_CHECK(hipMemcpyAsync(part, dp.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_edge_cut(int n, const int* xadj, const int* adjncy, const int* adjwgt, const int* part,
                   long long* cut) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        Engine eng;
        hipStream_t s = eng.stream();
        OwnedGraph g;
        upload_graph(n, xadj, adjncy, nullptr, adjwgt, g, s);
        DevBuf<int> dp;
        dp.reserve(n);
        if (n) KH_CHECK(hipMemcpyAsync(dp.p, part, sizeof(int) * n, hipMemcpyHostToDevice, s));
        *cut = eng.mw.edge_cut(g.g, dp.p, s, nullptr);
    });
}

int kahip_quality_noalloc(int n, long long m2, const int* xadj, const int* adjncy,
                          const int* adjwgt, const int* vwgt, const int* part, int k,
                          long long* cut, long long* block_weights, void* stream) {
    return guarded([&] {
        quality_noalloc(n, m2, xadj, adjncy, adjwgt, vwgt, part, k, cut, block_weights,
                        static_cast<hipStream_t>(stream));
    });
}

int kahip_block_weights(int n, const int* vwgt, int k, const int* part, long long* weights) {
    return guarded([&] {
        KH_REQUIRE(n >= 0 && k >= 1 && k <= 65536, "bad n or k");
        for (int v = 0; v < n; ++v) KH_REQUIRE(part[v] >= 0 && part[v] < k, "part out of range");
        Engine eng;
        hipStream_t s = eng.stream();
        DevBuf<int> dp, dv, W;
        dp.reserve(n);
        dv.reserve(n);
        W.reserve(k);
        DGraph g;
        g.n = n;
        g.vwgt = dv.p;
        if (n) {
            KH_CHECK(hipMemcpyAsync(dp.p, part, sizeof(int) * n, hipMemcpyHostToDevice, s));
            if (vwgt) {
                KH_CHECK(hipMemcpyAsync(dv.p, vwgt, sizeof(int) * n, hipMemcpyHostToDevice, s));
            } else {
                std::vector<int> ones(n, 1);
                KH_CHECK(hipMemcpyAsync(dv.p, ones.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
                KH_CHECK(hipStreamSynchronize(s));
            }
        }
        compute_block_weights(g, dp.p, k, W.p, s);
        std::vector<int> h(k);
        KH_CHECK(hipMemcpyAsync(h.data(), W.p, sizeof(int) * k, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        for (int b = 0; b < k; ++b) weights[b] = h[b];
    });
}

int kahip_project(int n_fine, const int* map, int n_coarse, const int* coarse_part,
                  int* fine_part) {
    return guarded([&] {
        KH_REQUIRE(n_fine >= 0 && n_coarse >= 0, "bad sizes");
        for (int v = 0; v < n_fine; ++v) KH_REQUIRE(map[v] >= 0 && map[v] < n_coarse, "map out of range");
        Engine eng;
        hipStream_t s = eng.stream();
        DevBuf<int> dm, dc, df;
        dm.reserve(n_fine);
        dc.reserve(n_coarse);
        df.reserve(n_fine);
        if (n_fine) KH_CHECK(hipMemcpyAsync(dm.p, map, sizeof(int) * n_fine, hipMemcpyHostToDevice, s));
        if (n_coarse) KH_CHECK(hipMemcpyAsync(dc.p, coarse_part, sizeof(int) * n_coarse, hipMemcpyHostToDevice, s));
        project_partition(n_fine, dm.p, dc.p, df.p, s, nullptr);
        if (n_fine) KH_CHECK(hipMemcpyAsync(fine_part, df.p, sizeof(int) * n_fine, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_initial_partition(int n, const int* xadj, const int* adjncy, const int* vwgt,
                            const int* adjwgt, int k, long long block_bound, int repetitions,
                            int bisection_tries, unsigned long long seed, int* part) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        KH_REQUIRE(k >= 1, "k must be >= 1");
        HostGraph g;
        g.n = n;
        g.xadj.assign(xadj, xadj + n + 1);
        g.adjncy.assign(adjncy, adjncy + xadj[n]);
        g.adjwgt.resize(xadj[n]);
        g.vwgt.resize(n);
        for (int e = 0; e < xadj[n]; ++e) g.adjwgt[e] = adjwgt ? adjwgt[e] : 1;
        for (int v = 0; v < n; ++v) g.vwgt[v] = vwgt ? vwgt[v] : 1;
        IpParams p;
        p.k = k;
        p.block_bound = block_bound;
        p.repetitions = repetitions > 0 ? repetitions : ip_repetitions_for(k);
        p.seed = seed;
        if (bisection_tries > 0) p.bisection_tries = bisection_tries;
        std::vector<int> out;
        initial_partition(g, p, out);
        std::copy(out.begin(), out.end(), part);
    });
}

int kahip_fm_refine(int n, const int* xadj, const int* adjncy, const int* vwgt,
                    const int* adjwgt, int k, long long block_bound, int rounds,
                    unsigned long long seed, int* part, long long* gain) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        KH_REQUIRE(k >= 1, "k must be >= 1");
        KH_REQUIRE(rounds >= 0, "rounds must be >= 0");
        HostGraph g;
        g.n = n;
        g.xadj.assign(xadj, xadj + n + 1);
        g.adjncy.assign(adjncy, adjncy + xadj[n]);
        g.adjwgt.resize(xadj[n]);
        g.vwgt.resize(n);
        for (int e = 0; e < xadj[n]; ++e) g.adjwgt[e] = adjwgt ? adjwgt[e] : 1;
        for (int v = 0; v < n; ++v) g.vwgt[v] = vwgt ? vwgt[v] : 1;
        std::vector<int> pv(part, part + n);
        for (int v = 0; v < n; ++v) KH_REQUIRE(pv[v] >= 0 && pv[v] < k, "block id out of range");
        FmRefineParams fp;
        fp.k = k;
        fp.block_bound = block_bound;
        fp.rounds = rounds;
        fp.seed = seed;
        const long long gn = fm_refine_pairs(g, fp, pv);
        std::copy(pv.begin(), pv.end(), part);
        if (gain) *gain = gn;
    });
}

int kahip_edge_features(int n, const int* xadj, const int* adjncy, const int* vwgt,
                        const int* adjwgt, int bound, unsigned seed, float* features) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        Engine eng;
        hipStream_t s = eng.stream();
        OwnedGraph g;
        upload_graph(n, xadj, adjncy, vwgt, adjwgt, g, s);
        DevBuf<float> f;
        f.reserve(std::max<long long>(g.g.m2 * kFeatures, 1));
        eng.rw.features(g.g, bound, seed, f.p, s, eng.lp, nullptr);
        if (g.g.m2)
            KH_CHECK(hipMemcpyAsync(features, f.p, sizeof(float) * kFeatures * g.g.m2,
                                    hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_graph_edge_features(kahip_graph_t g, int bound, unsigned seed, float* features) {
    return guarded([&] {
        KH_REQUIRE(g, "null graph");
        Engine eng;
        hipStream_t s = eng.stream();
        const DGraph& d = g->og.g;
        DevBuf<float> f;
        f.reserve(std::max<long long>(d.m2 * kFeatures, 1));
        eng.rw.features(d, bound, seed, f.p, s, eng.lp, nullptr);
        if (d.m2)
            KH_CHECK(hipMemcpyAsync(features, f.p, sizeof(float) * kFeatures * d.m2,
                                    hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_mlp_forward(long long count, const float* features, const float* params,
                      float* logits) {
    return guarded([&] {
        KH_REQUIRE(count >= 0, "count must be non-negative");
        KH_REQUIRE(params, "params is NULL");
        Engine eng;
        hipStream_t s = eng.stream();
        DevBuf<float> f, o;
        f.reserve(std::max<long long>(count * kFeatures, 1));
        o.reserve(std::max<long long>(count, 1));
        if (count)
            KH_CHECK(hipMemcpyAsync(f.p, features, sizeof(float) * kFeatures * count,
                                    hipMemcpyHostToDevice, s));
        eng.rw.mlp_forward(f.p, count, *reinterpret_cast<const MlpWeights*>(params), o.p, s,
                           nullptr);
        if (count)
            KH_CHECK(hipMemcpyAsync(logits, o.p, sizeof(float) * count, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_edge_scores(int n, const int* xadj, const int* adjncy, const int* vwgt,
                      const int* adjwgt, int rating, const float* params, int bound,
                      unsigned seed, int* scores) {
    return guarded([&] {
        check_csr(n, xadj, adjncy);
        KH_REQUIRE(rating == 1 || rating == 2, "rating must be EXPANSION_STAR2 or LEARNED");
        KH_REQUIRE(rating != 2 || params, "learned rating needs params");
        Engine eng;
        hipStream_t s = eng.stream();
        OwnedGraph g;
        upload_graph(n, xadj, adjncy, vwgt, adjwgt, g, s);
        DevBuf<int> o;
        o.reserve(std::max<long long>(g.g.m2, 1));
        eng.rw.scores((RatingMode)rating, g.g, bound, seed,
                      reinterpret_cast<const MlpWeights*>(params), o.p, s, eng.lp, nullptr);
        if (g.g.m2)
            KH_CHECK(hipMemcpyAsync(scores, o.p, sizeof(int) * g.g.m2, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_mlp_gradient(int count, const float* features, const float* labels,
                       const float* params, float* grads, double* loss) {
    return guarded([&] {
        KH_REQUIRE(count > 0, "count must be positive");
        Engine eng;
        hipStream_t s = eng.stream();
        DevBuf<float> f, y, w, g, l;
        f.reserve((long long)count * kFeatures);
        y.reserve(count);
        w.reserve(kMlpParams);
        g.reserve(kMlpParams);
        l.reserve(1);
        KH_CHECK(hipMemcpyAsync(f.p, features, sizeof(float) * kFeatures * count,
                                hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(y.p, labels, sizeof(float) * count, hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(w.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
        eng.rw.train_step(f.p, y.p, count, w.p, nullptr, nullptr, g.p, 1, 0.f, l.p, s);
        float hl = 0;
        KH_CHECK(hipMemcpyAsync(grads, g.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipMemcpyAsync(&hl, l.p, sizeof(float), hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        if (loss) *loss = hl;
    });
}

int kahip_mlp_train(long long count, const float* features, const float* labels, int batch,
                    int epochs, float lr, float* params, double* loss_per_epoch) {
    return guarded([&] {
        KH_REQUIRE(count > 0 && batch > 0 && epochs >= 0, "bad count / batch / epochs");
        KH_REQUIRE(lr > 0, "lr must be positive");
        Engine eng;
        hipStream_t s = eng.stream();
        DevBuf<float> f, y, w, m, v, g, l;
        f.reserve(count * kFeatures);
        y.reserve(count);
        w.reserve(kMlpParams);
        m.reserve(kMlpParams);
        v.reserve(kMlpParams);
        g.reserve(kMlpParams);
        const long long steps_per_epoch = (count + batch - 1) / batch;
        l.reserve(std::max<long long>(steps_per_epoch, 1));
        KH_CHECK(hipMemcpyAsync(f.p, features, sizeof(float) * kFeatures * count,
                                hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(y.p, labels, sizeof(float) * count, hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(w.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
        std::vector<float> zeros(kMlpParams, 0.f), hl(steps_per_epoch);
        KH_CHECK(hipMemcpyAsync(m.p, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(v.p, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
        int step = 0;
        for (int ep = 0; ep < epochs; ++ep) {
            for (long long i = 0; i < steps_per_epoch; ++i) {
                const long long lo = i * batch;
                const int b = (int)std::min<long long>(batch, count - lo);
                eng.rw.train_step(f.p + lo * kFeatures, y.p + lo, b, w.p, m.p, v.p, g.p, ++step,
                                  lr, l.p + i, s);
            }
            KH_CHECK(hipMemcpyAsync(hl.data(), l.p, sizeof(float) * steps_per_epoch,
                                    hipMemcpyDeviceToHost, s));
            KH_CHECK(hipStreamSynchronize(s));
            double tot = 0;
            for (float x : hl) tot += x;
            if (loss_per_epoch) loss_per_epoch[ep] = tot / (double)steps_per_epoch;
        }
        KH_CHECK(hipMemcpyAsync(params, w.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
    });
}

int kahip_batch_from_host(int num_graphs, const long long* node_off, const long long* edge_off,
                          const int* xadj, const int* adjncy, const int* vwgt, const int* adjwgt,
                          kahip_batch_t* out) {
    return guarded([&] {
        KH_REQUIRE(num_graphs >= 1 && node_off && edge_off && xadj, "bad batch arguments");
        KH_REQUIRE(node_off[0] == 0 && edge_off[0] == 0, "batch offsets must start at 0");
        for (int g = 0; g < num_graphs; ++g)
            KH_REQUIRE(node_off[g + 1] > node_off[g] && edge_off[g + 1] >= edge_off[g],
                       "batch offsets must increase (every graph needs a node)");
        auto* b = new kahip_batch();
        try {
            hipStream_t s;
            KH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            upload_batch(num_graphs, node_off, edge_off, xadj, adjncy, vwgt, adjwgt, b->og, s);
            (void)hipStreamDestroy(s);
            b->node_off.assign(num_graphs + 1, 0);
            for (int g = 0; g <= num_graphs; ++g) b->node_off[g] = (int)node_off[g];
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int kahip_batch_generate(int num_graphs, double log2_min, double log2_max, unsigned long long seed,
                         kahip_batch_t* out) {
    return guarded([&] {
        auto* b = new kahip_batch();
        try {
            generate_batch(num_graphs, log2_min, log2_max, seed, b->og, b->node_off);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int kahip_batch_info(kahip_batch_t b, int* num_graphs, int* n, long long* m2, int* node_off) {
    return guarded([&] {
        KH_REQUIRE(b, "null batch");
        const int G = (int)b->node_off.size() - 1;
        if (num_graphs) *num_graphs = G;
        if (n) *n = b->og.g.n;
        if (m2) *m2 = b->og.g.m2;
        if (node_off) std::copy(b->node_off.begin(), b->node_off.end(), node_off);
    });
}

int kahip_batch_to_host(kahip_batch_t b, int* xadj, int* adjncy, int* vwgt, int* adjwgt) {
    return guarded([&] {
        KH_REQUIRE(b, "null batch");
        const DGraph& d = b->og.g;
        if (xadj) KH_CHECK(hipMemcpy(xadj, d.xadj, sizeof(int) * (d.n + 1), hipMemcpyDeviceToHost));
        if (adjncy && d.m2) KH_CHECK(hipMemcpy(adjncy, d.adjncy, sizeof(int) * d.m2, hipMemcpyDeviceToHost));
        if (adjwgt && d.m2) KH_CHECK(hipMemcpy(adjwgt, d.adjwgt, sizeof(int) * d.m2, hipMemcpyDeviceToHost));
        if (vwgt && d.n) KH_CHECK(hipMemcpy(vwgt, d.vwgt, sizeof(int) * d.n, hipMemcpyDeviceToHost));
    });
}

void kahip_batch_free(kahip_batch_t b) { delete b; }

int kahip_batch_partition(kahip_batch_t b, const kahip_params* params, int use_graph, int* part,
                          long long* cuts, long long* max_block_weights,
                          kahip_batch_stats* stats) {
    return guarded([&] {
        KH_REQUIRE(b, "null batch");
        PartitionParams p = to_params(params);
        EngineLease eng;
        std::vector<int> tmp;
        if (!part) {
            tmp.resize(std::max(b->og.g.n, 1));
            part = tmp.data();
        }
        BatchStats st;
        const int G = (int)b->node_off.size() - 1;
        partition_batch(*eng, b->og.g, G, b->node_off.data(), p, part, &st, use_graph != 0);
        for (int g = 0; g < G; ++g) {
            if (cuts) cuts[g] = st.cut[g];
            if (max_block_weights) max_block_weights[g] = st.max_block_weight[g];
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->ms_coarsen = st.ms_coarsen;
            stats->ms_initial = st.ms_initial;
            stats->ms_refine = st.ms_refine;
            stats->ms_cut = st.ms_cut;
            stats->levels = st.levels;
            stats->edges_cluster = st.edges_cluster;
            stats->edges_contract = st.edges_contract;
            stats->edges_refine = st.edges_refine;
            stats->edges_cut = st.edges_cut;
        }
    });
}

int kahip_autotune_bins(kahip_graph_t g, kahip_batch_t b, const kahip_params* params, int* bins,
                        double* ms_per_candidate) {
    return guarded([&] {
        KH_REQUIRE((g != nullptr) != (b != nullptr), "pass exactly one of graph and batch");
        PartitionParams p = to_params(params);
        const DGraph& d = g ? g->og.g : b->og.g;
        BinConfig best = autotune_bins(d, p, false, ms_per_candidate);
        for (int i = 0; i < kNumBins; ++i) bins[i] = best.hi[i];
    });
}

int kahip_autotune(kahip_graph_t g, kahip_batch_t b, const kahip_params* params,
                   kahip_params* tuned, double* ms) {
    return guarded([&] {
        KH_REQUIRE((g != nullptr) != (b != nullptr), "pass exactly one of graph and batch");
        KH_REQUIRE(tuned != nullptr, "tuned must not be NULL");
        PartitionParams p = to_params(params);
        const DGraph& d = g ? g->og.g : b->og.g;
        const BinConfig best = autotune_bins(d, p, true, ms);
        kahip_params out;
        if (params) out = *params;
        else kahip_default_params(&out);
        for (int i = 0; i < kNumBins; ++i) out.bins[i] = best.hi[i];
        out.group_items = best.group_items;
        out.wave_small = best.wave_small;
        *tuned = out;
    });
}

}  // extern "C"
