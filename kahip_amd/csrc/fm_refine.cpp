// Pairwise 2-way FM over the quotient graph (see fm_refine.hpp).
//
// A round collects, for every pair of adjacent blocks (A, B), the nodes of A with a neighbour in
// B and of B with a neighbour in A (the pair's boundary, complete_boundary.h:104-214), orders the
// pairs by boundary size, and runs one FM pass per pair: moves between A and B only, gains
// counted on the edges into A and B (edges into third blocks stay cut either way), each node
// moved at most once, the best prefix kept (lower overload first, then lower cut), stop after a
// stall limit of moves without improvement (two_way_fm.cpp:69-71). Nodes that become boundary
// nodes during the pass join it. Ties between equal gains are broken by a per-(round, pair, step)
// hash, so the result depends only on the seed. The code is this library's own; it follows the
// reference's scheme (quotient-graph scheduling of 2-way FM), not its code.

#include "fm_refine.hpp"

#include <algorithm>
#include <utility>

namespace kahip {

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Addressable binary max-heap over (gain, key); pos is shared by both sides' heaps.
struct Heap {
    std::vector<int> h;
    int* pos = nullptr;
    const long long* gain = nullptr;
    const uint64_t* key = nullptr;
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
    void update(int v) {
        up(pos[v]);
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

struct PairFm {
    const HostGraph& g;
    std::vector<int>& part;
    std::vector<long long>& W;
    long long U;
    long long slack = 0;
    std::vector<long long> gain;
    std::vector<uint64_t> key;
    std::vector<int> pos;
    std::vector<char> locked;
    std::vector<int> touched, moved;
    Heap heap[2];
    uint64_t salt = 0, step = 0;

    PairFm(const HostGraph& g_, std::vector<int>& part_, std::vector<long long>& W_, long long U_)
        : g(g_), part(part_), W(W_), U(U_), gain(g_.n, 0), key(g_.n, 0), pos(g_.n, -1),
          locked(g_.n, 0) {
        for (int s = 0; s < 2; ++s) {
            heap[s].pos = pos.data();
            heap[s].gain = gain.data();
            heap[s].key = key.data();
        }
    }

    long long overload(int a, int b) const {
        return std::max(0LL, W[a] - U) + std::max(0LL, W[b] - U);
    }

    // gain of moving u to the other block of the pair, and whether u touches the other block
    long long pair_gain(int u, int own, int other, bool* boundary) const {
        long long gv = 0;
        bool bnd = false;
        for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
            const int b = part[g.adjncy[e]];
            if (b == other) {
                gv += g.adjwgt[e];
                bnd = true;
            } else if (b == own) {
                gv -= g.adjwgt[e];
            }
        }
        *boundary = bnd;
        return gv;
    }

    void enter(int u, int side) {
        key[u] = (mix64(salt ^ (step++ * 0xD1B54A32D192ED03ULL)) & ~0xFFFFFFFFULL) | (uint32_t)u;
        if (pos[u] < 0) {
            heap[side].insert(u);
            touched.push_back(u);
        } else {
            heap[side].update(u);
        }
    }

    // One FM pass on the pair (blk[0], blk[1]) from the candidate nodes; returns the cut gain.
    long long run(const int blk[2], const std::vector<int>& cand, int stall_pct) {
        heap[0].h.clear();
        heap[1].h.clear();
        touched.clear();
        moved.clear();
        int nb = 0;
        for (int u : cand) {
            const int pu = part[u];
            if ((pu != blk[0] && pu != blk[1]) || pos[u] >= 0) continue;
            const int s = pu == blk[0] ? 0 : 1;
            bool bnd = false;
            gain[u] = pair_gain(u, blk[s], blk[1 - s], &bnd);
            if (!bnd) continue;
            enter(u, s);
            ++nb;
        }
        const int stall = std::max(15, (int)((long long)stall_pct * nb / 100));
        long long cut_delta = 0, best_delta = 0;
        long long best_over = overload(blk[0], blk[1]);
        size_t best_len = 0;
        int since = 0;
        while (true) {
            const int v0 = heap[0].h.empty() ? -1 : heap[0].h[0];
            const int v1 = heap[1].h.empty() ? -1 : heap[1].h[0];
            auto ok = [&](int v, int s) {
                if (v < 0) return false;
                return W[blk[1 - s]] + g.vwgt[v] <= U + slack || W[blk[s]] > U;
            };
            const bool a0 = ok(v0, 0), a1 = ok(v1, 1);
            int v = -1, s = 0;
            if (a0 && a1) {
                if (gain[v0] > gain[v1] || (gain[v0] == gain[v1] && W[blk[0]] >= W[blk[1]])) {
                    v = v0;
                    s = 0;
                } else {
                    v = v1;
                    s = 1;
                }
            } else if (a0) {
                v = v0;
            } else if (a1) {
                v = v1;
                s = 1;
            }
            if (v < 0) break;
            heap[s].pop();
            const int from = blk[s], to = blk[1 - s];
            part[v] = to;
            locked[v] = 1;
            W[from] -= g.vwgt[v];
            W[to] += g.vwgt[v];
            cut_delta -= gain[v];
            moved.push_back(v);
            for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int x = g.adjncy[e];
                if (locked[x]) continue;
                const int px = part[x];
                if (px != from && px != to) continue;
                const int sx = px == blk[0] ? 0 : 1;
                if (pos[x] >= 0) {
                    // v left x's block (x in from) or joined it (x in to)
                    gain[x] += px == from ? 2LL * g.adjwgt[e] : -2LL * g.adjwgt[e];
                    enter(x, sx);
                } else {
                    bool bnd = false;
                    gain[x] = pair_gain(x, px, px == from ? to : from, &bnd);
                    if (bnd) enter(x, sx);
                }
            }
            const long long ov = overload(blk[0], blk[1]);
            if (ov < best_over || (ov == best_over && cut_delta < best_delta)) {
                best_over = ov;
                best_delta = cut_delta;
                best_len = moved.size();
                since = 0;
            } else if (++since > stall) {
                break;
            }
        }
        for (size_t i = moved.size(); i > best_len; --i) {  // roll back past the best prefix
            const int v = moved[i - 1];
            const int to = part[v], from = to == blk[0] ? blk[1] : blk[0];
            part[v] = from;
            W[to] -= g.vwgt[v];
            W[from] += g.vwgt[v];
        }
        for (int u : touched) {
            pos[u] = -1;
            locked[u] = 0;
        }
        for (int u : moved) locked[u] = 0;
        return -best_delta;
    }
};

}  // namespace

long long fm_refine_pairs(const HostGraph& g, const FmRefineParams& p, std::vector<int>& part) {
    const int n = g.n, k = p.k;
    if (n == 0 || k < 2) return 0;
    std::vector<long long> W(k, 0);
    for (int v = 0; v < n; ++v) W[part[v]] += g.vwgt[v];
    PairFm fm(g, part, W, p.block_bound);
    fm.slack = p.pass_slack;
    std::vector<int> stamp(k, -1);
    long long total = 0;
    for (int r = 0; r < p.rounds; ++r) {
        // (pair, node) for every boundary node and every other block it touches
        std::vector<std::pair<uint64_t, int>> ent;
        for (int u = 0; u < n; ++u) {
            const int a = part[u];
            for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
                const int b = part[g.adjncy[e]];
                if (b == a || stamp[b] == u) continue;
                stamp[b] = u;
                const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
                ent.emplace_back(key, u);
            }
        }
        std::fill(stamp.begin(), stamp.end(), -1);
        std::sort(ent.begin(), ent.end());
        // pairs by boundary size (descending), ties by pair
        std::vector<std::pair<int, size_t>> order;  // (-size, start)
        for (size_t i = 0; i < ent.size();) {
            size_t j = i;
            while (j < ent.size() && ent[j].first == ent[i].first) ++j;
            order.emplace_back(-(int)(j - i), i);
            i = j;
        }
        std::sort(order.begin(), order.end());
        long long round_gain = 0;
        std::vector<int> cand;
        for (size_t oi = 0; oi < order.size(); ++oi) {
            const size_t i = order[oi].second;
            const uint64_t pk = ent[i].first;
            cand.clear();
            for (size_t j = i; j < ent.size() && ent[j].first == pk; ++j) cand.push_back(ent[j].second);
            const int blk[2] = {(int)(pk >> 32), (int)(pk & 0xFFFFFFFFu)};
            fm.salt = mix64(p.seed ^ mix64(((uint64_t)r << 48) ^ pk));
            fm.step = 0;
            round_gain += fm.run(blk, cand, p.stall_pct);
        }
        total += round_gain;
        if (round_gain == 0) break;
    }
    return total;
}

}  // namespace kahip
