// CPU initial partitioning of the coarsest graph: recursive multilevel bisection, after the
// reference's scheme (initial_partitioning.cpp:39-134 repeats
// initial_partition_bipartition.cpp:22-60, which runs graph_partitioner.cpp:110-238: bisect the
// block range [lb, ub] at floor((lb+ub)/2), extract both sides, recurse). Each bisection is
// multilevel: size-constrained LP coarsening (cluster bound ceil(max side / 12), as
// initial_partition_bipartition.cpp:39-43), BFS-grown starts (bipartition_tries of them, as
// bipartition.cpp:24-75) refined by 2-way FM, and FM at every uncoarsening level.
//
// The implementation is this library's own: sequential code, one independent RNG stream per
// (repetition, block range), so results do not depend on the number of threads used. Nodes of
// degree 0 do not affect the cut; they are left out of the bisections and packed into the
// lightest blocks at the end (heaviest first).

#include "initial_partition.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <functional>
#include <future>
#include <numeric>
#include <queue>
#include <thread>

namespace kahip {

long long HostGraph::total_weight() const {
    long long s = 0;
    for (int v = 0; v < n; ++v) s += vwgt[v];
    return s;
}

long long host_edge_cut(const HostGraph& g, const std::vector<int>& part) {
    long long c = 0;
    for (int v = 0; v < g.n; ++v)
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
            if (part[v] != part[g.adjncy[e]]) c += g.adjwgt[e];
    return c / 2;
}

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return n ? (uint32_t)(next() % n) : 0; }
    bool coin() { return next() & 1; }
};

uint64_t stream_seed(uint64_t seed, uint64_t a, uint64_t b) {
    Rng r(seed ^ (a * 0xD1B54A32D192ED03ULL) ^ (b * 0x8CB92BA72F3D8DD7ULL));
    r.next();
    return r.next();
}

// ------------------------------------------------------------------ coarsening for bisection
// Sequential size-constrained LP in a random order (the reference's IP uses the same clustering
// with factor 12 inside the bisection hierarchy). Returns the number of clusters; label is
// renumbered to 0..nc-1 in first-appearance order.
int lp_cluster(const HostGraph& g, long long bound, int iters, Rng& rng, std::vector<int>& label) {
    const int n = g.n;
    label.resize(n);
    // int32 sizes and connection sums: the library keeps total node and edge weight below 2^31
    std::vector<int> size(n);
    std::vector<int> order(n);
    for (int v = 0; v < n; ++v) {
        label[v] = v;
        size[v] = g.vwgt[v];
        order[v] = v;
    }
    for (int i = n - 1; i > 0; --i) std::swap(order[i], order[rng.below(i + 1)]);
    std::vector<int> acc(n, 0);
    std::vector<int> touched;
    for (int it = 0; it < iters; ++it) {
        int moved = 0;
        for (int v : order) {
            const int own = label[v];
            for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int c = label[g.adjncy[e]];
                if (acc[c] == 0) touched.push_back(c);
                acc[c] += g.adjwgt[e];
            }
            int best = own;
            int bv = acc[own];
            for (int c : touched) {
                // the connection test first: it rejects most candidates without reading size[c]
                if (c == own || acc[c] < bv) continue;
                if ((long long)size[c] + g.vwgt[v] >= bound) continue;
                if (acc[c] > bv || rng.coin()) {
                    best = c;
                    bv = acc[c];
                }
            }
            for (int c : touched) acc[c] = 0;
            touched.clear();
            if (best != own) {
                size[own] -= g.vwgt[v];
                size[best] += g.vwgt[v];
                label[v] = best;
                ++moved;
            }
        }
        if (moved == 0) break;
    }
    std::vector<int> remap(n, -1);
    int nc = 0;
    for (int v = 0; v < n; ++v) {
        int& r = remap[label[v]];
        if (r < 0) r = nc++;
        label[v] = r;
    }
    return nc;
}

HostGraph contract(const HostGraph& g, const std::vector<int>& label, int nc) {
    HostGraph c;
    c.n = nc;
    c.vwgt.assign(nc, 0);
    c.xadj.assign(nc + 1, 0);
    std::vector<int> start(nc + 1, 0), members(g.n);
    for (int v = 0; v < g.n; ++v) {
        c.vwgt[label[v]] += g.vwgt[v];
        start[label[v] + 1]++;
    }
    for (int p = 0; p < nc; ++p) start[p + 1] += start[p];
    std::vector<int> pos(start.begin(), start.end() - 1);
    for (int v = 0; v < g.n; ++v) members[pos[label[v]]++] = v;
    std::vector<int> acc(nc, 0);
    std::vector<int> touched;
    c.adjncy.reserve(g.adjncy.size() / 2);
    c.adjwgt.reserve(g.adjncy.size() / 2);
    for (int p = 0; p < nc; ++p) {
        for (int i = start[p]; i < start[p + 1]; ++i) {
            const int v = members[i];
            for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int q = label[g.adjncy[e]];
                if (q == p) continue;
                if (acc[q] == 0) touched.push_back(q);
                acc[q] += g.adjwgt[e];
            }
        }
        for (int q : touched) {
            c.adjncy.push_back(q);
            c.adjwgt.push_back(acc[q]);
            acc[q] = 0;
        }
        touched.clear();
        c.xadj[p + 1] = (int)c.adjncy.size();
    }
    return c;
}

// ------------------------------------------------------------------ 2-way FM
struct Bisection {
    std::vector<char> side;
    long long w[2] = {0, 0};
    long long cut = 0;
};

long long bis_cut(const HostGraph& g, const std::vector<char>& side) {
    long long c = 0;
    for (int v = 0; v < g.n; ++v)
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
            if (side[v] != side[g.adjncy[e]]) c += g.adjwgt[e];
    return c / 2;
}

inline long long overload(const long long w[2], const long long maxw[2]) {
    return std::max(0LL, w[0] - maxw[0]) + std::max(0LL, w[1] - maxw[1]);
}

// Addressable binary max-heap of the FM queue of one side: one entry per node, ordered by
// (gain, random tiebreak, node). A gain change draws a new tiebreak and moves the entry in place.
struct FmHeap {
    std::vector<int> h;  // node ids in heap order
    int* pos = nullptr;  // pos[v]: index of v in h, -1 if absent (shared by both sides)
    const long long* gain = nullptr;
    const uint64_t* key = nullptr;  // tie << 32 | node
    bool less(int a, int b) const {
        return gain[a] < gain[b] || (gain[a] == gain[b] && key[a] < key[b]);
    }
    void up(int i) {
        const int v = h[i];
        while (i > 0) {
            const int p = (i - 1) >> 1;
            if (!less(h[p], v)) break;
            h[i] = h[p];
            pos[h[i]] = i;
            i = p;
        }
        h[i] = v;
        pos[v] = i;
    }
    void down(int i) {
        const int v = h[i];
        const int sz = (int)h.size();
        while (true) {
            int c = 2 * i + 1;
            if (c >= sz) break;
            if (c + 1 < sz && less(h[c], h[c + 1])) ++c;
            if (!less(v, h[c])) break;
            h[i] = h[c];
            pos[h[i]] = i;
            i = c;
        }
        h[i] = v;
        pos[v] = i;
    }
    void insert(int v) {
        h.push_back(v);
        up((int)h.size() - 1);
    }
    void update(int v) {  // after key/gain of v changed
        const int i = pos[v];
        up(i);
        down(pos[v]);
    }
    void pop() {
        const int v = h[0];
        pos[v] = -1;
        const int last = h.back();
        h.pop_back();
        if (!h.empty()) {
            h[0] = last;
            pos[last] = 0;
            down(0);
        }
    }
};

// Buffers of fm_pass, kept per thread across calls (the tries call it thousands of times on
// graphs of a few hundred nodes, where allocation would dominate).
struct FmScratch {
    std::vector<long long> gain;
    std::vector<uint64_t> key;
    std::vector<int> pos, moved;
    std::vector<char> locked;
    std::vector<int> heap[2];
};

// One FM pass with rollback to the best prefix. State order: lower overload first, then lower
// cut. Returns true when the state improved. Queues: boundary nodes of each side by gain; a node
// enters its queue when it first has a moved neighbour or at the start if it is on the boundary.
bool fm_pass(const HostGraph& g, Bisection& b, const long long maxw[2], int stall_limit,
             int stall_pct, Rng& rng) {
    static thread_local FmScratch sc;
    const int n = g.n;
    // moved into locals for the pass (and back at the end), so that the compiler can keep their
    // data pointers in registers across the char stores into b.side
    std::vector<long long> gain = std::move(sc.gain);
    std::vector<uint64_t> key = std::move(sc.key);
    std::vector<int> pos = std::move(sc.pos);
    std::vector<int> moved = std::move(sc.moved);
    std::vector<char> locked = std::move(sc.locked);
    FmHeap pq[2];
    pq[0].h = std::move(sc.heap[0]);
    pq[1].h = std::move(sc.heap[1]);
    gain.assign(n, 0);
    key.resize(n);
    pos.assign(n, -1);
    locked.assign(n, 0);
    moved.clear();
    for (int s = 0; s < 2; ++s) {
        pq[s].h.clear();
        pq[s].pos = pos.data();
        pq[s].gain = gain.data();
        pq[s].key = key.data();
    }
    int nb = 0;
    for (int v = 0; v < n; ++v) {
        long long gv = 0;
        bool bnd = false;
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            if (b.side[g.adjncy[e]] != b.side[v]) {
                gv += g.adjwgt[e];
                bnd = true;
            } else {
                gv -= g.adjwgt[e];
            }
        }
        gain[v] = gv;
        if (bnd) {
            key[v] = ((uint64_t)(uint32_t)rng.next() << 32) | (uint32_t)v;
            pq[(int)b.side[v]].insert(v);
            ++nb;
        }
    }
    if (stall_pct > 0) stall_limit = std::max(15, (int)((long long)stall_pct * nb / 100));
    long long cut = b.cut, w[2] = {b.w[0], b.w[1]};
    long long best_cut = cut, best_over = overload(w, maxw);
    size_t best_len = 0;
    int since_best = 0;
    while (true) {
        const int v0 = pq[0].h.empty() ? -1 : pq[0].h[0];
        const int v1 = pq[1].h.empty() ? -1 : pq[1].h[0];
        // a move from s to 1-s is allowed if the target stays within its bound or the source
        // is overloaded
        auto ok = [&](int v, int s) {
            if (v < 0) return false;
            const int t = 1 - s;
            return w[t] + g.vwgt[v] <= maxw[t] || w[s] > maxw[s];
        };
        const bool a0 = ok(v0, 0), a1 = ok(v1, 1);
        int v = -1, s = 0;
        if (a0 && a1) {
            if (gain[v0] > gain[v1] || (gain[v0] == gain[v1] && w[0] >= w[1])) {
                v = v0;
                s = 0;
            } else {
                v = v1;
                s = 1;
            }
        } else if (a0) {
            v = v0;
            s = 0;
        } else if (a1) {
            v = v1;
            s = 1;
        }
        if (v < 0) break;
        pq[s].pop();
        const long long gv = gain[v];
        b.side[v] = (char)(1 - s);
        locked[v] = 1;
        w[s] -= g.vwgt[v];
        w[1 - s] += g.vwgt[v];
        cut -= gv;
        moved.push_back(v);
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int u = g.adjncy[e];
            if (locked[u]) continue;
            // v left u's side (if u on s) or joined u's side (if u on 1-s)
            const long long d = 2LL * g.adjwgt[e];
            if (b.side[u] == s)
                gain[u] += d;
            else
                gain[u] -= d;
            key[u] = ((uint64_t)(uint32_t)rng.next() << 32) | (uint32_t)u;
            if (pos[u] < 0)
                pq[(int)b.side[u]].insert(u);
            else
                pq[(int)b.side[u]].update(u);
        }
        const long long ov = overload(w, maxw);
        if (ov < best_over || (ov == best_over && cut < best_cut)) {
            best_over = ov;
            best_cut = cut;
            best_len = moved.size();
            since_best = 0;
        } else if (++since_best > stall_limit) {
            break;
        }
    }
    // roll back to the best prefix
    for (size_t i = moved.size(); i > best_len; --i) {
        const int v = moved[i - 1];
        const int s = b.side[v];
        b.side[v] = (char)(1 - s);
        w[s] -= g.vwgt[v];
        w[1 - s] += g.vwgt[v];
    }
    const long long old_over = overload(b.w, maxw);
    const bool improved = best_over < old_over || (best_over == old_over && best_cut < b.cut);
    b.w[0] = w[0];
    b.w[1] = w[1];
    b.cut = best_cut;
    sc.gain = std::move(gain);
    sc.key = std::move(key);
    sc.pos = std::move(pos);
    sc.moved = std::move(moved);
    sc.locked = std::move(locked);
    sc.heap[0] = std::move(pq[0].h);
    sc.heap[1] = std::move(pq[1].h);
    return improved;
}

void fm_refine(const HostGraph& g, Bisection& b, const long long maxw[2], int stall, int pct,
               Rng& rng) {
    for (int pass = 0; pass < 8; ++pass)
        if (!fm_pass(g, b, maxw, stall, pct, rng)) break;
}

// BFS growing of side 0 from random starts until it reaches target0.
Bisection grow(const HostGraph& g, long long target0, Rng& rng) {
    Bisection b;
    b.side.assign(g.n, 1);
    long long w0 = 0;
    std::vector<char> seen(g.n, 0);
    std::vector<int> queue;
    queue.reserve(g.n);
    size_t head = 0;
    int remaining = g.n;
    while (w0 < target0 && remaining > 0) {
        if (head == queue.size()) {
            int s = (int)rng.below((uint32_t)g.n);
            while (seen[s]) s = (s + 1) % g.n;
            seen[s] = 1;
            queue.push_back(s);
        }
        const int v = queue[head++];
        --remaining;
        b.side[v] = 0;
        w0 += g.vwgt[v];
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int u = g.adjncy[e];
            if (!seen[u]) {
                seen[u] = 1;
                queue.push_back(u);
            }
        }
    }
    long long total = 0;
    for (int v = 0; v < g.n; ++v) total += g.vwgt[v];
    b.w[0] = w0;
    b.w[1] = total - w0;
    b.cut = bis_cut(g, b.side);
    return b;
}

bool better_bis(const Bisection& a, const Bisection& b, const long long maxw[2]) {
    const long long oa = overload(a.w, maxw), ob = overload(b.w, maxw);
    return oa < ob || (oa == ob && a.cut < b.cut);
}

// Multilevel bisection of g. side[v] = 0 for the first part.
std::vector<char> ml_bisect(const HostGraph& g, long long target0, const long long maxw[2],
                            const IpParams& p, Rng& rng) {
    std::vector<HostGraph> levels;
    std::vector<std::vector<int>> maps;
    const HostGraph* cur = &g;
    const long long cbound = std::max(1LL, (std::min(maxw[0], maxw[1]) + p.coarsen_factor - 1) /
                                               p.coarsen_factor);
    while (cur->n > p.stop_nodes) {
        std::vector<int> label;
        const int nc = lp_cluster(*cur, cbound, 3, rng, label);
        if ((double)cur->n / (double)nc < 1.1) break;
        HostGraph c = contract(*cur, label, nc);
        maps.push_back(std::move(label));
        levels.push_back(std::move(c));
        cur = &levels.back();
    }
    Bisection best;
    bool have = false;
    for (int t = 0; t < p.bisection_tries; ++t) {
        Bisection b = grow(*cur, target0, rng);
        fm_refine(*cur, b, maxw, p.fm_stall, p.fm_pct_coarsest, rng);
        if (!have || better_bis(b, best, maxw)) {
            best = std::move(b);
            have = true;
        }
    }
    for (int l = (int)levels.size() - 1; l >= 0; --l) {
        const HostGraph& fine = l == 0 ? g : levels[l - 1];
        const std::vector<int>& map = maps[l];
        Bisection b;
        b.side.resize(fine.n);
        for (int v = 0; v < fine.n; ++v) b.side[v] = best.side[map[v]];
        b.w[0] = best.w[0];
        b.w[1] = best.w[1];
        b.cut = best.cut;
        fm_refine(fine, b, maxw, p.fm_stall, p.fm_pct_levels, rng);
        best = std::move(b);
    }
    return best.side;
}

HostGraph extract(const HostGraph& g, const std::vector<char>& side, char which,
                  std::vector<int>& ids) {
    std::vector<int> local(g.n, -1);
    ids.clear();
    for (int v = 0; v < g.n; ++v)
        if (side[v] == which) {
            local[v] = (int)ids.size();
            ids.push_back(v);
        }
    HostGraph s;
    s.n = (int)ids.size();
    s.xadj.assign(s.n + 1, 0);
    s.vwgt.resize(s.n);
    for (int i = 0; i < s.n; ++i) {
        const int v = ids[i];
        s.vwgt[i] = g.vwgt[v];
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int u = local[g.adjncy[e]];
            if (u < 0) continue;
            s.adjncy.push_back(u);
            s.adjwgt.push_back(g.adjwgt[e]);
        }
        s.xadj[i + 1] = (int)s.adjncy.size();
    }
    return s;
}

struct RecCtx {
    const IpParams* p;
    uint64_t rep_seed;
    std::atomic<int> spawn_budget{0};
};

// Partitions g into blocks lb..ub; part receives block ids for g's nodes.
void recurse(const HostGraph& g, int lb, int ub, RecCtx& ctx, std::vector<int>& part) {
    part.assign(g.n, lb);
    const int kk = ub - lb + 1;
    if (kk <= 1 || g.n == 0) return;
    const int mid = (lb + ub) / 2;
    const int k1 = mid - lb + 1, k2 = ub - mid;
    const long long W = g.total_weight();
    const long long target0 = (long long)std::llround((double)W * k1 / kk);
    // Each side may hold up to (number of its final blocks) x U: the imbalance budget is taken
    // against the global ideal block weight at every level, so every sub-problem stays feasible
    // and final blocks respect U whenever each bisection respects its bounds.
    long long maxw[2] = {(long long)k1 * ctx.p->block_bound, (long long)k2 * ctx.p->block_bound};
    maxw[0] = std::max(maxw[0], target0);
    maxw[1] = std::max(maxw[1], W - target0);
    Rng rng(stream_seed(ctx.rep_seed, (uint64_t)lb, (uint64_t)ub));
    std::vector<char> side;
    if (g.n == 1) {
        side.assign(1, 0);
    } else {
        side = ml_bisect(g, target0, maxw, *ctx.p, rng);
    }
    std::vector<int> ids0, ids1;
    HostGraph g0 = extract(g, side, 0, ids0);
    HostGraph g1 = extract(g, side, 1, ids1);
    std::vector<int> p0, p1;
    bool spawned = false;
    std::future<void> fut;
    if (k1 > 1 && k2 > 1 && g.n > 20000 && ctx.spawn_budget.fetch_sub(1) > 0) {
        fut = std::async(std::launch::async, [&] { recurse(g0, lb, mid, ctx, p0); });
        spawned = true;
    } else {
        recurse(g0, lb, mid, ctx, p0);
    }
    recurse(g1, mid + 1, ub, ctx, p1);
    if (spawned) fut.get();
    for (size_t i = 0; i < ids0.size(); ++i) part[ids0[i]] = p0[i];
    for (size_t i = 0; i < ids1.size(); ++i) part[ids1[i]] = p1[i];
}

// Moves nodes out of blocks heavier than `bound` until none is (or no move is possible): nodes
// of overloaded blocks are ranked by the cut increase of their best move into a block with room
// (neighbouring blocks, else the lightest block) and moved greedily; rankings are recomputed
// when a target fills up. Isolated nodes cost nothing and go first.
void rebalance(const HostGraph& g, int k, long long bound, std::vector<int>& part) {
    std::vector<long long> W(k, 0);
    for (int v = 0; v < g.n; ++v) W[part[v]] += g.vwgt[v];
    std::vector<long long> conn(k, 0);
    std::vector<int> touched;
    for (int pass = 0; pass < 64; ++pass) {
        bool any_over = false;
        for (int b = 0; b < k; ++b) any_over |= W[b] > bound;
        if (!any_over) return;
        const int lightest = (int)(std::min_element(W.begin(), W.end()) - W.begin());
        struct Move {
            long long loss;
            int v, to;
        };
        std::vector<Move> moves;
        for (int v = 0; v < g.n; ++v) {
            const int own = part[v];
            if (W[own] <= bound) continue;
            for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int b = part[g.adjncy[e]];
                if (conn[b] == 0) touched.push_back(b);
                conn[b] += g.adjwgt[e];
            }
            int t = -1;
            long long tc = -1;
            for (int b : touched)
                if (b != own && W[b] + g.vwgt[v] <= bound && conn[b] > tc) {
                    t = b;
                    tc = conn[b];
                }
            if (t < 0 && lightest != own && W[lightest] + g.vwgt[v] <= bound) {
                t = lightest;
                tc = conn[lightest];
            }
            const long long own_conn = conn[own];
            for (int b : touched) conn[b] = 0;
            touched.clear();
            if (t >= 0) moves.push_back({own_conn - tc, v, t});
        }
        if (moves.empty()) return;
        std::sort(moves.begin(), moves.end(), [](const Move& a, const Move& b) {
            return a.loss < b.loss || (a.loss == b.loss && a.v < b.v);
        });
        bool moved = false;
        for (const Move& m : moves) {
            const int own = part[m.v];
            if (W[own] <= bound) continue;
            if (W[m.to] + g.vwgt[m.v] > bound) continue;  // target filled up: next pass re-ranks
            W[own] -= g.vwgt[m.v];
            W[m.to] += g.vwgt[m.v];
            part[m.v] = m.to;
            moved = true;
        }
        if (!moved) return;
    }
}

}  // namespace

void initial_partition(const HostGraph& g, const IpParams& p, std::vector<int>& part) {
    const int k = std::max(1, p.k);
    part.assign(g.n, 0);
    if (k == 1 || g.n == 0) return;
    // split off isolated nodes
    std::vector<int> core_ids, iso_ids;
    for (int v = 0; v < g.n; ++v)
        (g.xadj[v + 1] > g.xadj[v] ? core_ids : iso_ids).push_back(v);
    std::vector<char> all_core(g.n, 0);
    for (int v : core_ids) all_core[v] = 1;
    std::vector<int> ids;
    HostGraph core = extract(g, all_core, 1, ids);
    unsigned hw = std::thread::hardware_concurrency();
    int threads = p.threads > 0 ? p.threads : (int)std::min(16u, hw ? hw : 1u);
    const int reps = std::max(1, p.repetitions);

    std::vector<std::vector<int>> results(reps);
    std::vector<long long> cuts(reps, 0);
    auto run_rep = [&](int r) {
        RecCtx ctx;
        ctx.p = &p;
        ctx.rep_seed = stream_seed(p.seed, 0x1e5u, (uint64_t)r);
        ctx.spawn_budget = std::max(0, threads / reps - 1);
        std::vector<int> cp;
        recurse(core, 0, k - 1, ctx, cp);
        results[r] = std::move(cp);
        std::vector<int> full(core.n);
        cuts[r] = 0;
        for (int v = 0; v < core.n; ++v)
            for (int e = core.xadj[v]; e < core.xadj[v + 1]; ++e)
                if (results[r][v] != results[r][core.adjncy[e]]) cuts[r] += core.adjwgt[e];
    };
    if (threads <= 1) {  // the caller parallelises (e.g. over the graphs of a batch)
        for (int r = 0; r < reps; ++r) run_rep(r);
    } else {
        // an exception (e.g. bad_alloc) is rethrown here, not left to end the process
        std::vector<std::exception_ptr> err(reps);
        std::vector<std::thread> pool;
        for (int r = 0; r < reps; ++r)
            pool.emplace_back([&, r] {
                try {
                    run_rep(r);
                } catch (...) {
                    err[r] = std::current_exception();
                }
            });
        for (auto& t : pool) t.join();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
    }
    int best = 0;
    for (int r = 1; r < reps; ++r)
        if (cuts[r] < cuts[best]) best = r;
    std::vector<long long> bw(k, 0);
    for (size_t i = 0; i < ids.size(); ++i) {
        part[ids[i]] = results[best][i];
        bw[results[best][i]] += g.vwgt[ids[i]];
    }
    // isolated nodes: heaviest first into the lightest block (lowest id on ties)
    std::stable_sort(iso_ids.begin(), iso_ids.end(),
                     [&](int a, int b) { return g.vwgt[a] > g.vwgt[b]; });
    using E = std::pair<long long, int>;
    std::priority_queue<E, std::vector<E>, std::greater<E>> heap;
    for (int b = 0; b < k; ++b) heap.emplace(bw[b], b);
    for (int v : iso_ids) {
        auto [w, b] = heap.top();
        heap.pop();
        part[v] = b;
        heap.emplace(w + g.vwgt[v], b);
    }
    if (p.block_bound > 0) rebalance(g, k, p.block_bound, part);
}

}  // namespace kahip
