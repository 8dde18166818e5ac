// Growable device buffers and the launch profiler (hipEvent pairs with algorithmic bytes).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace kahip {

template <class T>
struct DevBuf {
    T* p = nullptr;
    long long cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap) {
        o.p = nullptr;
        o.cap = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            free();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        }
        return *this;
    }
    ~DevBuf() { free(); }
    // Ensures room for n elements (contents are not preserved on growth).
    T* reserve(long long n) {
        if (n < 1) n = 1;
        if (n > cap) {
            free();
            KH_CHECK(hipMalloc((void**)&p, sizeof(T) * n));
            cap = n;
        }
        return p;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    operator T*() const { return p; }
};

// Kernel-launch profiler. Each record brackets one launch (or one fused phase) on a stream with a
// pair of events and carries the launch's algorithmic HBM bytes and traversed edges, so that
// achieved GB/s = bytes / duration is measured on the stream the kernel ran on.
struct Profiler {
    bool enabled = false;
    // phases and proposal kernels only: the whole clustering of a level, the contraction,
    // renumbering and cut, and every clustering proposal launch (the candidates for the dominant
    // kernel); fewer event pairs, so the timing perturbs the run less
    bool light = false;
    static bool phase_scope(const std::string& n) {
        static const std::string prop = "lp_cluster_propose_";
        if (n.compare(0, prop.size(), prop) == 0) return n != "lp_cluster_propose_hub_merge";
        return n == "lp_cluster_total" || n == "contract" || n == "remap" || n == "reorder" ||
               n == "edge_cut";
    }
    struct Rec {
        std::string name;
        hipEvent_t a, b;
        double bytes;
        long long edges;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;

    ~Profiler() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    hipEvent_t take() {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            KH_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[pool_used++];
    }
    int begin(const std::string& name, double bytes, long long edges, hipStream_t s) {
        if (!enabled || (light && !phase_scope(name))) return -1;
        Rec r{name, take(), take(), bytes, edges};
        KH_CHECK(hipEventRecord(r.a, s));
        recs.push_back(r);
        return (int)recs.size() - 1;
    }
    void end(int idx, hipStream_t s) {
        if (idx < 0) return;
        KH_CHECK(hipEventRecord(recs[idx].b, s));
    }
    void reset() {
        recs.clear();
        pool_used = 0;
    }
    struct Sum {
        long long launches = 0;
        double ms = 0, bytes = 0;
        long long edges = 0;
    };
    // Call after the stream has been synchronised.
    std::map<std::string, Sum> summarize() const {
        std::map<std::string, Sum> m;
        for (const Rec& r : recs) {
            float ms = 0;
            KH_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
            Sum& s = m[r.name];
            s.launches++;
            s.ms += ms;
            s.bytes += r.bytes;
            s.edges += r.edges;
        }
        return m;
    }
};

struct ProfScope {
    Profiler* p;
    int idx;
    hipStream_t s;
    ProfScope(Profiler* prof, const std::string& name, double bytes, long long edges,
              hipStream_t st)
        : p(prof), idx(prof ? prof->begin(name, bytes, edges, st) : -1), s(st) {}
    ~ProfScope() { close(); }
    // Ends the timed range early (idempotent).
    void close() {
        if (p) p->end(idx, s);
        p = nullptr;
    }
};

}  // namespace kahip
