// Synchronous label propagation on gfx950: one kernel family serves both the size-constrained
// clustering of the coarsening phase (labels = clusters, bound B; reference:
// lib/partition/coarsening/clustering/size_constraint_label_propagation.cpp:134-194) and the
// LP refinement of the uncoarsening phase (labels = blocks, bound U; reference:
// lib/partition/uncoarsening/refinement/label_propagation_refinement/
// label_propagation_refinement.cpp:21-155).
//
// A round is S sub-rounds; a node belongs to sub-round hash3(seed ^ kSubSalt, 0, v) % S. One
// sub-round is four launches, all order-independent, so the result is bit-identical to the CPU
// mirror of the same schedule whatever the thread interleaving:
//   propose  every node of the sub-round picks its best eligible neighbouring label from the
//            label / weight snapshot; proposers append a mover record and add their weight and
//            priority key to the target's inc / top (integer atomics: order-free);
//   filter   (refinement) departures from overloaded blocks;
//   admit    a proposal is admitted if all proposals into its target fit, or if it is the
//            target's top-priority proposal and fits alone;
//   apply    admitted moves update labels and label weights; scratch is reset.
// Nodes are binned by degree (per sub-round lists built once per call):
//   deg <= 8/16/32/64   G lanes per node, one edge per lane, shuffle aggregation (no LDS);
//   deg <= 512          one wave per node, LDS hash table (open addressing, LDS atomics);
//   larger              one workgroup per node, hash table in a pre-cleared global region.

#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "lp.hpp"

namespace kahip {

namespace {

constexpr int kWaveCap = 1024;    // LDS hash slots per wave (degree <= 512 => load <= 1/2)
static_assert(kWaveCap == 1024, "propose_wave loads up to kWaveCap / 2 edges per wave");
constexpr int kWaveDense = 2048;  // dense LDS counters per wave (refinement, k <= 2048)
constexpr int kWaveDenseSmall = 512;  // the frontier kernel's counters for k <= 512 (less LDS)
// Clustering splits the wave list: nodes of degree <= BinConfig::wave_small (64, 128 or 256)
// come first and use a kernel of 2 x wave_small slots per wave (8 KB of LDS per workgroup at 128
// instead of 32 KB, so more waves per CU).
constexpr int kBlock = 256;
constexpr int kFrontierGrid = 2048;  // workgroups of a refinement list kernel (grid-stride)
// nodes per lane group in flight in propose_group: BinConfig::group_items (1, 2 or 4)
constexpr int kHubChunk = 4096;   // hub edges per workgroup
constexpr int kHubCap = 8192;     // LDS hash slots per hub chunk (load <= 1/2)
constexpr int kHubThreads = 1024;      // clustering hub chunks (4 edges per thread)
constexpr int kMidThreads = 256;       // hubs of kMidChunk edges or fewer (4 edges per thread)
constexpr int kMidChunk = 1024;
constexpr int kHubDenseThreads = 512;  // refinement hub chunks (8 edges per thread)
constexpr int kHubDense = 4096;   // dense LDS counters per hub (refinement, k <= 4096)
constexpr long long kAllHubCounters = 1LL << 28;  // counters (ints) for incremental hubs
constexpr int kHubDenseSlots = 12288;  // dynamic LDS counters (48 KB): min(8, slots / k) copies
constexpr int kFinThreads = 1024;  // hub_bucket_merge workgroup
constexpr int kFinCap = 16384;     // LDS hash slots of hub_bucket_merge (128 KB)
constexpr int kFinLds = 8192;      // buckets up to this many entries merge in LDS
constexpr int kMaxBuckets = 2048;  // buckets per hub: min(kBucketsPerChunk x chunks, kMaxBuckets)
constexpr int kBucketsPerChunk = 4;  // ~1000 entries per (hub, bucket): the small merge kernel takes most
constexpr int kTinyThreads = 256;    // tiny bucket merge: <= kTinyMax entries in 16 KB of LDS
constexpr int kTinyCap = 2048;
constexpr int kTinyMax = 1024;
constexpr int kSmallThreads = 512;   // small bucket merge: <= kSmallMax entries in 32 KB of LDS
constexpr int kSmallCap = 4096;
constexpr int kSmallMax = 2048;
constexpr int kHubStage = 256;       // merge tasks of hubs of at most this many chunks stage the
                                     // chunks' segment offsets in LDS
constexpr int kBigDegree = 64;    // moved nodes above this degree mark neighbours per workgroup

struct Cand {
    int conn;
    uint32_t tie;
    int lab;
    int size;
};
static_assert(sizeof(Cand) == sizeof(int4), "bucket results are stored as int4");

// Bucket of a label among P buckets of a hub (independent of the LDS slot hash).
__device__ inline int bucket_of(int lab, int P) {
    return (int)(mix32((uint32_t)lab ^ 0x5bd1e995u) % (uint32_t)P);
}

__device__ inline int bound_of(const LpArgs& a, int lab) {
    return a.lbound ? a.lbound[lab] : a.bound;
}

__device__ inline bool better(int ac, uint32_t at, int al, int bc, uint32_t bt, int bl) {
    return ac > bc || (ac == bc && (at > bt || (at == bt && al < bl)));
}

// Candidate (conn, lab) of node v (weight nw) against the best so far: the label's size (and
// bound) is read only when the candidate would beat c, so a thread reads the sizes of its running
// maxima only. The result, the maximum of better() over the eligible labels, is the same as with
// every size read: c only ever holds eligible labels, and the best eligible one beats whatever c
// holds when it comes up.
__device__ inline void consider(const LpArgs& a, int v, int nw, int conn, int lab, Cand& c) {
    const uint32_t t = hash3(a.tie_seed, (uint32_t)v, (uint32_t)lab);
    if (!better(conn, t, lab, c.conn, c.tie, c.lab)) return;
    const int sz = a.size[lab];
    if ((long long)sz + nw >= (long long)bound_of(a, lab)) return;
    c = Cand{conn, t, lab, sz};
}

template <int W>
__device__ inline void reduce_cand(Cand& c) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        int oc = __shfl_xor(c.conn, o, W);
        uint32_t ot = (uint32_t)__shfl_xor((int)c.tie, o, W);
        int ol = __shfl_xor(c.lab, o, W);
        int os = __shfl_xor(c.size, o, W);
        if (better(oc, ot, ol, c.conn, c.tie, c.lab)) {
            c.conn = oc;
            c.tie = ot;
            c.lab = ol;
            c.size = os;
        }
    }
}

// Rotation within each row of 16 lanes (DPP row_ror: a full-rate VALU move, no LDS crossbar).
template <int R>
__device__ inline int row_ror(int x) {
    if constexpr (R == 0) return x;
    else return __builtin_amdgcn_update_dpp(0, x, 0x120 + R, 0xF, 0xF, false);
}

// Adds, for rotations R..15 of (l0, w0, i0) within the lane's row, the weight w0 of every lane
// whose label l0 equals lab; a lane of lower id i0 with the same label clears leader. The ids are
// rotated with the values, so the result does not depend on the rotation's direction.
template <int R>
__device__ inline void agg_rot(int lab, int gl, int l0, int w0, int i0, int& conn, bool& leader) {
    if constexpr (R < 16) {
        const int lr = row_ror<R>(l0), wr = row_ror<R>(w0), ir = row_ror<R>(i0);
        if (lr == lab) {
            conn += wr;
            if (ir < gl) leader = false;
        }
        agg_rot<R + 1>(lab, gl, l0, w0, i0, conn, leader);
    }
}

template <int W>
__device__ inline int reduce_max(int x) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, W));
    return x;
}

// A proposal's contribution to the admission scratch: weight and priority into the target label
// (inc / top) or, for a departure from an overloaded block, out of the own block (out / dtop).
struct Contrib {
    int label;  // -1: none
    int w;
    unsigned long long key;
    bool departure;
};

// The proposal rule of one proposing node (called by one lane): writes the node's proposal slot
// and returns the scratch contribution, which the caller adds with atomics.
template <LpMode M>
__device__ inline Contrib propose(const LpArgs& a, int v, int own, int w, int conn_own,
                                  const Cand& c) {
    Contrib r{-1, 0, 0ull, false};
    if (c.conn < 0) return r;  // no eligible candidate
    const int size_own = a.size[own];
    bool go;
    bool overloaded = false;
    if (M == LpMode::Cluster) {
        go = c.conn > conn_own || (c.conn == conn_own && c.size >= size_own);
    } else {
        overloaded = size_own > bound_of(a, own);
        go = c.conn > conn_own || overloaded ||
             (c.conn == conn_own && (long long)c.size + 2LL * w <= (long long)size_own);
    }
    if (!go) return r;
    const int gain = c.conn - conn_own;
    const unsigned long long key = prio_key(gain, v);
    a.prop_t[v] = c.lab;
    a.prop_gain[v] = gain;
    a.prop_ok[v] = 1;
    const bool dep = M == LpMode::Refine && overloaded;
    return Contrib{dep ? own : c.lab, w, key, dep};
}

__device__ inline void add_contrib(const LpArgs& a, const Contrib& r) {
    if (r.label < 0) return;
    if (r.departure) {
        atomicAdd(&a.out[r.label], r.w);
        atomicMax(&a.dtop[r.label], r.key);
    } else {
        atomicAdd(&a.inc[r.label], r.w);
        atomicMax(&a.top[r.label], r.key);
    }
}

template <LpMode M>
__device__ inline void decide(const LpArgs& a, int v, int own, int w, int conn_own,
                              const Cand& c) {
    add_contrib(a, propose<M>(a, v, own, w, conn_own, c));
}

// Decision scan of a wave's LDS label table: the best eligible label among the slots lane,
// lane + 64, ... below cap (keys K, summed weights V; empty slots -1) and the weight towards the
// own label. A pass takes SLOTS slots per lane: it reads their keys and values, then issues the
// label-weight (and per-label bound) gathers of all of them, then evaluates, so a lane waits for
// one gather latency per pass instead of one per slot. Candidates compare by the total order of
// better(), so the result does not depend on the scan order. (The workgroup-per-hub kernels keep
// the slot-by-slot scan: there the batched form measured slower.)
template <int SLOTS>
__device__ __attribute__((always_inline)) inline void scan_wave_table(
    const LpArgs& a, const int* K, const int* V, int cap, int lane, int v, int own, int nw,
    int& conn_own, Cand& c) {
    for (int base = lane; base < cap; base += SLOTS * 64) {
        int lk[SLOTS], lv[SLOTS], ls[SLOTS], lb[SLOTS];
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) {
            const int i = base + j * 64;
            lk[j] = i < cap ? K[i] : -1;
            lv[j] = lk[j] >= 0 ? V[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) ls[j] = lk[j] >= 0 && lk[j] != own ? a.size[lk[j]] : 0;
        if (a.lbound) {  // uniform (a select between the two would put a.bound in scratch)
#pragma unroll
            for (int j = 0; j < SLOTS; ++j)
                lb[j] = lk[j] >= 0 && lk[j] != own ? a.lbound[lk[j]] : 0;
        } else {
#pragma unroll
            for (int j = 0; j < SLOTS; ++j) lb[j] = a.bound;
        }
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) {
            const int lab = lk[j];
            if (lab < 0) continue;
            if (lab == own) {
                conn_own = lv[j];
                continue;
            }
            if ((long long)ls[j] + nw >= (long long)lb[j] || lv[j] < c.conn) continue;
            const uint32_t t = hash3(a.tie_seed, (uint32_t)v, (uint32_t)lab);
            if (better(lv[j], t, lab, c.conn, c.tie, c.lab)) c = Cand{lv[j], t, lab, ls[j]};
        }
    }
}

// ---------------------------------------------------------------- G lanes per node (deg <= G)
// dcount: the list's length on the device (the compacted frontier of a refinement round);
// workgroups then stride over the list's virtual blocks, so a small frontier costs few of them.
// A virtual block holds IT x (kBlock / G) nodes: every lane group serves IT nodes at once, whose
// dependent loads (list, offsets, targets, labels, label weights) are issued together, so each
// wave has that many chains in flight.
// One virtual block (IT x kBlock / G nodes of the list) of the G-lane kernel; every thread of
// the workgroup calls it (it synchronises the workgroup). `salt` picks the traversal counter.
template <LpMode M, int G, int IT>
__device__ __attribute__((always_inline)) inline void group_vb(const LpArgs& a,
                                                               const int* __restrict__ list,
                                                               int count, int vb, int salt) {
    constexpr int NB = kBlock / G;  // nodes per item and workgroup
    const int gl = threadIdx.x & (G - 1);
    int v[IT], own[IT], beg[IT], deg[IT], lab[IT], w[IT];
    bool valid[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const long long gi = ((long long)vb * IT + i) * NB + threadIdx.x / G;
        valid[i] = gi < count;
        v[i] = valid[i] ? list[gi] : 0;
    }
    if (M == LpMode::Refine) {
#pragma unroll
        for (int i = 0; i < IT; ++i) valid[i] = valid[i] && a.active[v[i]] != 0;
    }
    // every lane of the group takes part in the shuffles below: no early exit
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        own[i] = valid[i] ? a.label[v[i]] : -1;
        beg[i] = valid[i] ? a.xadj[v[i]] : 0;
        deg[i] = valid[i] ? a.xadj[v[i] + 1] - beg[i] : 0;
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        if constexpr (G >= 16) {  // rows of 9..64 entries: streamed (non-temporal)
            lab[i] = gl < deg[i] ? __builtin_nontemporal_load(a.adjncy + beg[i] + gl) : -1;
            w[i] = gl < deg[i] ? __builtin_nontemporal_load(a.adjwgt + beg[i] + gl) : 0;
        } else {
            lab[i] = gl < deg[i] ? a.adjncy[beg[i] + gl] : -1;
            w[i] = gl < deg[i] ? a.adjwgt[beg[i] + gl] : 0;
        }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i)
        if (lab[i] >= 0) lab[i] = a.label[lab[i]];
    // conn = summed weight of the group's lanes with this lane's label; leader = the lowest of
    // them. Groups of 16, 32 and 64 lanes are whole DPP rows: rotations within the row, then
    // within each other row of the group after one swap (xor 16, 32, 48), visit every lane of
    // the group once.
    int conn[IT];
    bool leader[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        conn[i] = 0;
        leader[i] = lab[i] >= 0;
        if constexpr (G >= 16) {
            conn[i] = w[i];
            agg_rot<1>(lab[i], gl, lab[i], w[i], gl, conn[i], leader[i]);
#pragma unroll
            for (int x = 16; x < G; x += 16)
                agg_rot<0>(lab[i], gl, __shfl_xor(lab[i], x, G), __shfl_xor(w[i], x, G), gl ^ x,
                           conn[i], leader[i]);
        } else {
#pragma unroll
            for (int j = 0; j < G; ++j) {
                const int lj = __shfl(lab[i], j, G);
                const int wj = __shfl(w[i], j, G);
                if (lj == lab[i]) {
                    conn[i] += wj;
                    if (j < gl) leader[i] = false;
                }
            }
        }
    }
    int conn_own[IT], nw[IT], sz[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        conn_own[i] = reduce_max<G>((lab[i] >= 0 && lab[i] == own[i]) ? conn[i] : 0);
        nw[i] = valid[i] ? a.vwgt[v[i]] : 0;
        sz[i] = (leader[i] && lab[i] != own[i]) ? a.size[lab[i]] : 0;
    }
    Contrib r[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        Cand c{-1, 0u, 0x7fffffff, 0};
        if (leader[i] && lab[i] != own[i] &&
            (long long)sz[i] + nw[i] < (long long)bound_of(a, lab[i])) {
            c.conn = conn[i];
            c.tie = hash3(a.tie_seed, (uint32_t)v[i], (uint32_t)lab[i]);
            c.lab = lab[i];
            c.size = sz[i];
        }
        reduce_cand<G>(c);
        r[i] = (valid[i] && gl == 0) ? propose<M>(a, v[i], own[i], nw[i], conn_own[i], c)
                                     : Contrib{-1, 0, 0ull, false};
    }
    // The block's proposals are combined per (target, kind) in an LDS hash before the atomics:
    // many low-degree nodes propose the same (hub) label, and one atomic per distinct target
    // keeps them from serialising on it.
    constexpr int HC = 2 * IT * NB;  // slots (load <= 1/2)
    // probed with a mask: a size that is not a power of two leaves slots unreachable and the
    // probe loop below would never end
    static_assert((HC & (HC - 1)) == 0, "IT x kBlock / G must be a power of two");
    __shared__ int h_tag[HC];        // (label << 1) | departure, -1: empty
    __shared__ int h_w[HC];
    __shared__ unsigned long long h_key[HC];
    for (int i = threadIdx.x; i < HC; i += kBlock) {
        h_tag[i] = -1;
        h_w[i] = 0;
        h_key[i] = 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        if (r[i].label < 0) continue;
        const int tag = (r[i].label << 1) | (r[i].departure ? 1 : 0);
        const int h = probe_claim(h_tag, HC, tag);  // at most IT x NB tags in HC slots
        if (h < 0) {
            *a.err = kErrProbe;
            continue;
        }
        atomicAdd(&h_w[h], r[i].w);
        atomicMax(&h_key[h], r[i].key);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HC; i += kBlock) {
        const int tag = h_tag[i];
        if (tag >= 0) add_contrib(a, Contrib{tag >> 1, h_w[i], h_key[i], (tag & 1) != 0});
    }
    if (M == LpMode::Refine) {
        __shared__ unsigned long long tsum;
        if (threadIdx.x == 0) tsum = 0;
        __syncthreads();
        unsigned long long mine = 0;
#pragma unroll
        for (int i = 0; i < IT; ++i)
            if (valid[i] && gl == 0) mine += (unsigned long long)deg[i];
        if (mine) atomicAdd(&tsum, mine);
        __syncthreads();
        if (threadIdx.x == 0 && tsum) atomicAdd(&a.trav[salt & 63], tsum);
    }
    __syncthreads();  // the LDS of this virtual block is read before the next one writes it
}

template <LpMode M, int G, int IT>
__global__ __launch_bounds__(kBlock) void propose_group(LpArgs a, const int* __restrict__ list,
                                                        int count, const int* __restrict__ dcount) {
    if (dcount) count = *dcount;
    const int nvb = (int)(((long long)count + IT * (kBlock / G) - 1) / (IT * (kBlock / G)));
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) group_vb<M, G, IT>(a, list, count, vb, vb);
}

// ---------------------------------------------------------------- one wave per node
// Labels are aggregated in LDS: a hash table (open addressing, LDS atomics) in general, or a
// dense counter per label when the labels are the k blocks of the refinement and k is small.
// One virtual block (kBlock / 64 nodes of the list, a wave each) of the wave kernel; every thread
// of the workgroup calls it.
template <LpMode M, bool DENSE, int WCAP, int DCAP = kWaveDense>
__device__ __attribute__((always_inline)) inline void wave_vb(const LpArgs& a,
                                                              const int* __restrict__ list,
                                                              int count, int num_labels, int vb,
                                                              int salt) {
    static_assert(!DENSE || WCAP == kWaveCap, "the dense variant serves every wave degree");
    __shared__ int skey[kBlock / 64][DENSE ? 1 : WCAP];
    __shared__ int sval[kBlock / 64][DENSE ? DCAP : WCAP];  // dense: num_labels <= DCAP
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int gi = vb * (kBlock / 64) + wave;
    bool valid = gi < count;
    const int v = valid ? list[gi] : 0;
    if (M == LpMode::Refine && valid) valid = a.active[v] != 0;
    const int own = valid ? a.label[v] : -1;
    const int beg = valid ? a.xadj[v] : 0;
    const int deg = valid ? a.xadj[v + 1] - beg : 0;
    int cap = 64;
    if (DENSE) {
        cap = num_labels;
    } else {
        while (cap < 2 * deg) cap <<= 1;  // deg <= WCAP / 2 => cap <= WCAP
    }
    int* K = skey[wave];
    int* V = sval[wave];
    for (int i = lane; i < cap; i += 64) {
        if (!DENSE) K[i] = -1;
        V[i] = 0;
    }
    __syncthreads();
    // All of a lane's (<= 8) edges are loaded first so that their label gathers are in flight
    // together; the LDS insertions follow.
    constexpr int EPL = WCAP / 2 / 64;  // edges per lane at the maximum degree (WCAP / 2)
    int labs[EPL], ws[EPL];
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
        const int e = beg + lane + 64 * j;
        // a row of up to WCAP / 2 edges, read once per sub-round: streamed (non-temporal)
        labs[j] = e < beg + deg ? __builtin_nontemporal_load(a.adjncy + e) : -1;
        ws[j] = e < beg + deg ? __builtin_nontemporal_load(a.adjwgt + e) : 0;
    }
#pragma unroll
    for (int j = 0; j < EPL; ++j)
        if (labs[j] >= 0) labs[j] = a.label[labs[j]];
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
        const int lab = labs[j];
        if (lab < 0) continue;
        const int w = ws[j];
        if (DENSE) {
            atomicAdd(&V[lab], w);
            continue;
        }
        const int h = probe_claim(K, cap, lab);  // cap >= 2 x degree
        if (h < 0) {
            *a.err = kErrProbe;
            continue;
        }
        atomicAdd(&V[h], w);
    }
    __syncthreads();
    const int nw = valid ? a.vwgt[v] : 0;
    int conn_own = 0;
    Cand c{-1, 0u, 0x7fffffff, 0};
    if (valid && !DENSE) {
        constexpr int SL = WCAP / 64 < 8 ? WCAP / 64 : 8;
        scan_wave_table<SL>(a, K, V, cap, lane, v, own, nw, conn_own, c);
    } else if (valid) {
        for (int i = lane; i < cap; i += 64) {
            const int lab = V[i] ? i : -1;
            if (lab < 0) continue;
            const int conn = V[i];
            if (lab == own) {
                conn_own = conn;
                continue;
            }
            const int sz = a.size[lab];
            if ((long long)sz + nw >= (long long)bound_of(a, lab)) continue;
            const uint32_t t = hash3(a.tie_seed, (uint32_t)v, (uint32_t)lab);
            if (better(conn, t, lab, c.conn, c.tie, c.lab)) c = Cand{conn, t, lab, sz};
        }
    }
    conn_own = reduce_max<64>(conn_own);
    reduce_cand<64>(c);
    if (valid && lane == 0) {
        decide<M>(a, v, own, nw, conn_own, c);
        if (M == LpMode::Refine) atomicAdd(&a.trav[salt & 63], (unsigned long long)deg);
    }
    __syncthreads();  // the LDS of this virtual block is read before the next one clears it
}

template <LpMode M, bool DENSE, int WCAP = kWaveCap>
__global__ __launch_bounds__(kBlock) void propose_wave(LpArgs a, const int* __restrict__ list,
                                                       int count, int num_labels,
                                                       const int* __restrict__ dcount) {
    if (dcount) count = *dcount;
    const int nvb = (count + kBlock / 64 - 1) / (kBlock / 64);
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x)
        wave_vb<M, DENSE, WCAP>(a, list, count, num_labels, vb, vb);
}

// Refinement: the proposals of all degree lists of a sub-round in one launch. The lists are the
// round's compacted frontier (lengths on the device); the workgroups stride over the virtual
// blocks of the five lists in order (G = 8, 16, 32, 64 lanes per node, then a wave per node), so a
// sub-round costs one launch instead of five. Same per-node code as the per-list kernels.
// Refinement hubs with kept per-block counters (every hub of the level has them when their
// counters fit): one wave per active hub decides from its k counters (as hub_dense_decide does
// with keep = 1; no row is read).
__device__ __attribute__((always_inline)) inline void hubcnt_vb(const LpArgs& a,
                                                                const int* __restrict__ list,
                                                                int count, int num_labels,
                                                                int vb) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gi = vb * (kBlock / 64) + wave;
    if (gi >= count) return;  // uniform per wave; no workgroup barrier below
    const int v = list[gi];
    if (a.active[v] == 0) return;
    const int* cnt = a.hub_cnt + a.hub_slot[v];
    const int own = a.label[v];
    const int nw = a.vwgt[v];
    int conn_own = 0;
    Cand c{-1, 0u, 0x7fffffff, 0};
    for (int b = lane; b < num_labels; b += 64) {
        const int conn = cnt[b];
        if (conn == 0) continue;
        if (b == own) {
            conn_own = conn;
            continue;
        }
        const int sz = a.size[b];
        if ((long long)sz + nw >= (long long)bound_of(a, b)) continue;
        const uint32_t t = hash3(a.tie_seed, (uint32_t)v, (uint32_t)b);
        if (better(conn, t, b, c.conn, c.tie, c.lab)) c = Cand{conn, t, b, sz};
    }
    conn_own = reduce_max<64>(conn_own);
    reduce_cand<64>(c);
    if (lane == 0) decide<LpMode::Refine>(a, v, own, nw, conn_own, c);
}

struct FrontierLists {
    const int* list[kNumBins];
    const int* dcount[kNumBins];  // dcount[kNumBins - 1] == nullptr: hubs decide elsewhere
};
template <int IT, bool DENSE, int DCAP = kWaveDense>
__global__ __launch_bounds__(kBlock) void propose_frontier(LpArgs a, FrontierLists F,
                                                           int num_labels) {
    constexpr int NPB[kNumBins] = {IT * (kBlock / 8), IT * (kBlock / 16), IT * (kBlock / 32),
                                   IT * (kBlock / 64), kBlock / 64, kBlock / 64};
    int cnt[kNumBins], nvb[kNumBins], total = 0;
#pragma unroll
    for (int b = 0; b < kNumBins; ++b) {
        cnt[b] = F.dcount[b] ? *F.dcount[b] : 0;
        nvb[b] = (cnt[b] + NPB[b] - 1) / NPB[b];
        total += nvb[b];
    }
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int vb = t, b = 0;  // uniform per workgroup
#pragma unroll
        for (int k = 0; k < kNumBins - 1; ++k)
            if (b == k && vb >= nvb[k]) {
                vb -= nvb[k];
                b = k + 1;
            }
        switch (b) {
            case 0: group_vb<LpMode::Refine, 8, IT>(a, F.list[0], cnt[0], vb, t); break;
            case 1: group_vb<LpMode::Refine, 16, IT>(a, F.list[1], cnt[1], vb, t); break;
            case 2: group_vb<LpMode::Refine, 32, IT>(a, F.list[2], cnt[2], vb, t); break;
            case 3: group_vb<LpMode::Refine, 64, IT>(a, F.list[3], cnt[3], vb, t); break;
            case 4:
                wave_vb<LpMode::Refine, DENSE, kWaveCap, DCAP>(a, F.list[4], cnt[4], num_labels,
                                                                vb, t);
                break;
            default: hubcnt_vb(a, F.list[5], cnt[5], num_labels, vb);
        }
    }
}

// Workgroup-wide reduction of (conn_own, best candidate); the result is valid in thread 0.
template <int THREADS>
__device__ inline void block_reduce(int& conn_own, Cand& c) {
    __shared__ Cand red[THREADS / 64];
    __shared__ int red_own[THREADS / 64];
    conn_own = reduce_max<64>(conn_own);
    reduce_cand<64>(c);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave] = c;
        red_own[wave] = conn_own;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < THREADS / 64; ++i) {
            conn_own = max(conn_own, red_own[i]);
            if (better(red[i].conn, red[i].tie, red[i].lab, c.conn, c.tie, c.lab)) c = red[i];
        }
    }
}

// ---------------------------------------------------------------- hubs, refinement: dense LDS
// Per-wave copies of the per-block counters (the k blocks are few, so a single shared array
// would serialise on LDS atomics), summed before the decision.
// Chunked variant for refinement hubs of more than kHubChunk edges: one workgroup per chunk sums
// its edges into per-wave LDS counters and adds the non-zero block sums to the hub's global
// counters (dense_off[hub] + b); hub_dense_decide then decides. Single-chunk hubs decide here.
// Hubs with global counters (dense_off >= 0) add their chunk sums there and are decided by
// hub_dense_decide. build: only those hubs, regardless of the frontier (the initial counters of
// the incremental scheme); incremental: those hubs are skipped (the moves keep their counters up
// to date).
__global__ __launch_bounds__(kHubDenseThreads) void propose_hub_dense_chunk(
    LpArgs a, const int4* __restrict__ cmap, int num_labels, const int* __restrict__ dense_off,
    int* __restrict__ dcnt, int build, int incremental) {
    extern __shared__ int cnt[];  // copies x num_labels counters (dynamic: small k, more blocks)
    const int4 cm = cmap[blockIdx.x];  // {hub index, chunk of the hub, chunks of the hub, node}
    const int gi = cm.x;
    const int v = cm.w;
    const bool counted = dense_off[gi] >= 0;  // the hub keeps global counters
    if (build ? !counted : (a.active[v] == 0 || (incremental && counted)))
        return;  // uniform per block
    const int c_idx = cm.y;
    const int beg = a.xadj[v];
    const int end = a.xadj[v + 1];
    const int e0 = beg + c_idx * kHubChunk;
    const int e1 = min(end, e0 + kHubChunk);
    const int copies = min(kHubDenseThreads / 64, kHubDenseSlots / num_labels);
    int* mine = cnt + ((threadIdx.x >> 6) % copies) * num_labels;
    for (int i = threadIdx.x; i < copies * num_labels; i += kHubDenseThreads) cnt[i] = 0;
    __syncthreads();
    constexpr int EPT = kHubChunk / kHubDenseThreads;
    int labs[EPT], ws[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        const int e = e0 + threadIdx.x + kHubDenseThreads * j;
        // the hub's row is read once per sub-round: streamed (non-temporal), so the caches stay
        // with the label gathers
        labs[j] = e < e1 ? __builtin_nontemporal_load(a.adjncy + e) : -1;
        ws[j] = e < e1 ? __builtin_nontemporal_load(a.adjwgt + e) : 0;
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j)
        if (labs[j] >= 0) labs[j] = a.label[labs[j]];
#pragma unroll
    for (int j = 0; j < EPT; ++j)
        if (labs[j] >= 0) atomicAdd(&mine[labs[j]], ws[j]);
    __syncthreads();
    // traversed edges: the edges this chunk read (hubs deciding from kept counters read none)
    if (threadIdx.x == 0) atomicAdd(&a.trav[blockIdx.x & 63], (unsigned long long)(e1 - e0));
    if (counted) {
        int* out = dcnt + dense_off[gi];
        for (int b = threadIdx.x; b < num_labels; b += kHubDenseThreads) {
            int conn = 0;
            for (int j = 0; j < copies; ++j) conn += cnt[j * num_labels + b];
            if (conn) atomicAdd(&out[b], conn);
        }
        return;
    }
    const int own = a.label[v];
    const int nw = a.vwgt[v];
    int conn_own = 0;
    Cand c{-1, 0u, 0x7fffffff, 0};
    for (int b = threadIdx.x; b < num_labels; b += kHubDenseThreads) {
        int conn = 0;
        for (int j = 0; j < copies; ++j) conn += cnt[j * num_labels + b];
        if (conn == 0) continue;
        if (b == own) {
            conn_own = conn;
            continue;
        }
        const int sz = a.size[b];
        if ((long long)sz + nw >= (long long)bound_of(a, b)) continue;
        const uint32_t t = hash3(a.tie_seed, (uint32_t)v, (uint32_t)b);
        if (better(conn, t, b, c.conn, c.tie, c.lab)) c = Cand{conn, t, b, sz};
    }
    block_reduce<kHubDenseThreads>(conn_own, c);
    if (threadIdx.x == 0) decide<LpMode::Refine>(a, v, own, nw, conn_own, c);
}

// One workgroup per multi-chunk refinement hub: decides from the summed counters and clears them
// (keep: the incremental counters stay).
__global__ __launch_bounds__(256) void hub_dense_decide(
    LpArgs a, const int* __restrict__ list, const int* __restrict__ multi, int n_multi,
    int num_labels, const int* __restrict__ dense_off, int* __restrict__ dcnt, int keep) {
    const int mi = blockIdx.x;
    if (mi >= n_multi) return;
    const int gi = multi[mi];
    const int v = list[gi];
    if (a.active[v] == 0) return;  // uniform per block
    int* cnt = dcnt + dense_off[gi];
    const int own = a.label[v];
    const int nw = a.vwgt[v];
    int conn_own = 0;
    Cand c{-1, 0u, 0x7fffffff, 0};
    for (int b = threadIdx.x; b < num_labels; b += 256) {
        const int conn = cnt[b];
        if (conn == 0) continue;
        if (!keep) cnt[b] = 0;
        if (b == own) {
            conn_own = conn;
            continue;
        }
        const int sz = a.size[b];
        if ((long long)sz + nw >= (long long)bound_of(a, b)) continue;
        const uint32_t t = hash3(a.tie_seed, (uint32_t)v, (uint32_t)b);
        if (better(conn, t, b, c.conn, c.tie, c.lab)) c = Cand{conn, t, b, sz};
    }
    block_reduce<256>(conn_own, c);
    // no traversed edges here: the chunks that read the hub's edges counted them
    if (threadIdx.x == 0) decide<LpMode::Refine>(a, v, own, nw, conn_own, c);
}

// ---------------------------------------------------------------- hubs: chunked LDS hash
// One workgroup per chunk of CHUNK edges of a hub. The chunk's labels are aggregated in an LDS
// hash table; a hub of one chunk decides right away, otherwise the chunk's (label, partial sum)
// entries are written out grouped by bucket(label), one bucket per later workgroup. The mid-size
// instance (kMidThreads, kMidChunk) serves hubs of one chunk only, in small workgroups that
// share a CU eight at a time.
template <LpMode M, int THREADS, int CHUNK>
__global__ __launch_bounds__(THREADS) void propose_hub_chunk(
    LpArgs a, const int4* __restrict__ cmap, int2* __restrict__ partials,
    const long long* __restrict__ offm_base, int* __restrict__ offm,
    const int* __restrict__ ct_base, int* __restrict__ cntT) {
    constexpr int CAP = 2 * CHUNK;
    constexpr bool MULTI = CHUNK == kHubChunk;  // hubs of several chunks come here
    static_assert(!MULTI || THREADS == kHubThreads, "the multi-chunk path assumes kHubThreads");
    __shared__ int K[CAP];
    __shared__ int V[CAP];
    const int chunk = blockIdx.x;
    const int4 cm = cmap[chunk];  // {hub index, chunk of the hub, chunks of the hub, node}
    const int gi = cm.x;
    const int v = cm.w;
    if (M == LpMode::Refine && a.active[v] == 0) return;  // uniform per block
    const int nchunks = cm.z;
    const int c_idx = cm.y;
    const int beg = a.xadj[v];
    const int end = a.xadj[v + 1];
    const int e0 = beg + c_idx * CHUNK;
    const int e1 = min(end, e0 + CHUNK);
    // a hub of one chunk uses a table of the next power of two >= 2 x its degree
    int cap = CAP;
    if (nchunks == 1)
        while (cap > 256 && cap >= 4 * (e1 - e0)) cap >>= 1;
    for (int i = threadIdx.x; i < cap; i += THREADS) {
        K[i] = -1;
        V[i] = 0;
    }
    __syncthreads();
    // load every edge of the thread first (independent label gathers in flight), then insert
    constexpr int EPT = CHUNK / THREADS;
    int labs[EPT], ws[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        const int e = e0 + threadIdx.x + THREADS * j;
        // the hub's row is read once per sub-round: streamed (non-temporal), so the caches stay
        // with the label gathers
        labs[j] = e < e1 ? __builtin_nontemporal_load(a.adjncy + e) : -1;
        ws[j] = e < e1 ? __builtin_nontemporal_load(a.adjwgt + e) : 0;
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j)
        if (labs[j] >= 0) labs[j] = a.label[labs[j]];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        const int lab = labs[j];
        if (lab < 0) continue;
        const int h = probe_claim(K, cap, lab);  // cap >= 2 x the chunk's edges
        if (h < 0) {
            *a.err = kErrProbe;
            continue;
        }
        atomicAdd(&V[h], ws[j]);
    }
    __syncthreads();
    if constexpr (MULTI) if (nchunks > 1) {
        // The chunk's (label, partial sum) entries, grouped by bucket(label): P buckets per hub,
        // bucket b of chunk c at [offm[c*(P+1)+b], offm[c*(P+1)+b+1]) of the chunk's region.
        __shared__ int bcur[kMaxBuckets + 1];
        __shared__ int wsum[kHubThreads / 64];
        const int P = min(kBucketsPerChunk * nchunks, kMaxBuckets);
        for (int i = threadIdx.x; i <= P; i += kHubThreads) bcur[i] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < kHubCap; i += kHubThreads)
            if (K[i] >= 0) atomicAdd(&bcur[bucket_of(K[i], P)], 1);
        __syncthreads();
        // exclusive scan of bcur[0..P): each thread owns a run of buckets
        const int per = (P + kHubThreads - 1) / kHubThreads;
        const int b0 = threadIdx.x * per, b1 = min(P, b0 + per);
        int run = 0;
        for (int b = b0; b < b1; ++b) run += bcur[b];
        int incl = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if ((int)(threadIdx.x & 63) >= o) incl += y;
        }
        if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
        __syncthreads();
        int wbase = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) wbase += wsum[w];
        int acc = wbase + incl - run;
        for (int b = b0; b < b1; ++b) {
            const int cnt = bcur[b];
            bcur[b] = acc;
            acc += cnt;
        }
        __syncthreads();
        int* offs = offm + offm_base[gi] + (long long)c_idx * (P + 1);
        int* cnt_t = cntT + ct_base[gi] + c_idx;  // bucket-major: [b * nchunks + c_idx]
        for (int b = threadIdx.x; b < P; b += kHubThreads) {
            offs[b] = bcur[b];
            if (b + 1 < P) cnt_t[(long long)b * nchunks] = bcur[b + 1] - bcur[b];
        }
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int w = 0; w < kHubThreads / 64; ++w) tot += wsum[w];
            offs[P] = tot;
            cnt_t[(long long)(P - 1) * nchunks] = tot - bcur[P - 1];
        }
        __syncthreads();
        int2* out = partials + (long long)chunk * kHubChunk;
        for (int i = threadIdx.x; i < kHubCap; i += kHubThreads) {
            const int lab = K[i];
            if (lab < 0) continue;
            out[atomicAdd(&bcur[bucket_of(lab, P)], 1)] = make_int2(lab, V[i]);
        }
        return;
    }
    const int own = a.label[v];
    const int nw = a.vwgt[v];
    int conn_own = 0;
    Cand c{-1, 0u, 0x7fffffff, 0};
    for (int i = threadIdx.x; i < cap; i += THREADS) {
        const int lab = K[i];
        if (lab < 0) continue;
        const int conn = V[i];
        if (lab == own) {
            conn_own = conn;
            continue;
        }
        consider(a, v, nw, conn, lab, c);
    }
    block_reduce<THREADS>(conn_own, c);
    if (threadIdx.x == 0) {
        decide<M>(a, v, own, nw, conn_own, c);
        if (M == LpMode::Refine) atomicAdd(&a.trav[blockIdx.x & 63], (unsigned long long)(end - beg));
    }
}

// Merge of one (multi-chunk hub, bucket) task: the bucket's entries of all chunks of the hub are
// summed in an LDS hash (or, past kFinLds entries, in the hub's global table, filtering by
// bucket); the bucket's best candidate and, for the bucket of the own label, conn(own) are left
// in bres / bown. Three sizes: LEVEL 0 (256 threads, 16 KB of LDS, eight per CU) takes every
// task of the sub-round and queues those of more than kTinyMax entries for LEVEL 1 (512 threads,
// 32 KB), which queues those of more than kSmallMax for LEVEL 2 (1024 threads, 128 KB); the
// queued levels grid-stride over their queue.
template <LpMode M, int LEVEL>
__global__ __launch_bounds__(LEVEL == 0 ? kTinyThreads : (LEVEL == 1 ? kSmallThreads : kFinThreads))
void hub_bucket_merge(
    LpArgs a, int n_tasks, const int4* __restrict__ tmap, const int2* __restrict__ tsub,
    const int* __restrict__ posT, const int* __restrict__ chunk_prefix,
    const long long* __restrict__ offm_base, const int* __restrict__ offm,
    const int2* __restrict__ partials, const long long* __restrict__ tab_off,
    int* __restrict__ gkey, int* __restrict__ gval, Cand* __restrict__ bres,
    int* __restrict__ bown, uint8_t* __restrict__ fb, const int* __restrict__ in_q,
    const int* __restrict__ in_n, int* __restrict__ out_q, int* __restrict__ out_n) {
    constexpr int THREADS = LEVEL == 0 ? kTinyThreads : (LEVEL == 1 ? kSmallThreads : kFinThreads);
    constexpr int CAP = LEVEL == 0 ? kTinyCap : (LEVEL == 1 ? kSmallCap : kFinCap);
    constexpr int MAXE = LEVEL == 0 ? kTinyMax : kSmallMax;  // queued beyond (levels 0, 1)
    static_assert(2 * MAXE <= CAP && 2 * kFinLds <= kFinCap, "merge tables fill at most half");
    __shared__ int K[CAP];
    __shared__ int V[CAP];
    __shared__ int spre[kHubStage + 1];  // task offsets of the chunks' bucket segments
    __shared__ int sstart[kHubStage];    // the segments' starts in their chunks' partials
    const int n_iter = LEVEL == 0 ? n_tasks : *in_n;
    for (int it = blockIdx.x; it < n_iter; it += gridDim.x) {
        const int task = LEVEL == 0 ? it : in_q[it];
        const int4 tm = tmap[task];  // {first count index, one past the last, hub index, node}
        const int2 ts = tsub[task];  // {multi-chunk hub index, bucket | buckets << 16}
        const int mi = ts.x;
        const int gi = tm.z;
        const int v = tm.w;
        if (M == LpMode::Refine && a.active[v] == 0) {  // uniform per block
            __syncthreads();
            continue;
        }
        const int P = ts.y >> 16;
        const int b = ts.y & 0xffff;
        const int r0 = posT[tm.x];
        const int r1 = posT[tm.y];
        const int total = r1 - r0;
        if (LEVEL < 2 && total > MAXE) {
            if (threadIdx.x == 0) out_q[atomicAdd(out_n, 1)] = task;
            __syncthreads();
            continue;
        }
        const bool in_lds = LEVEL < 2 || total <= kFinLds;
        const int deg = a.xadj[v + 1] - a.xadj[v];
        // The task's entries are read where the hub's chunks wrote them: chunk c holds entries
        // [pre(c), pre(c + 1)) of the task, pre(c) = posT[tm.x + c] - r0, in its partials at
        // offm[offm_base[gi] + c (P + 1) + b]; the chunk of entry i is the largest c with
        // pre(c) <= i (empty segments are skipped).
        const int nch = tm.y - tm.x;
        const long long c0 = chunk_prefix[gi];
        const int* offs = offm + offm_base[gi] + b;
        const bool staged = nch <= kHubStage;  // uniform
        if (staged)
            for (int c = threadIdx.x; c <= nch; c += THREADS) {
                spre[c] = posT[tm.x + c] - r0;
                if (c < nch) sstart[c] = offs[(long long)c * (P + 1)];
            }
        long long cap = CAP;
        int* TK = K;
        int* TV = V;
        if (in_lds) {
            cap = 64;
            while (cap < 2 * total) cap <<= 1;
            for (int i = threadIdx.x; i < cap; i += THREADS) {
                K[i] = -1;
                V[i] = 0;
            }
        } else {
            cap = 64;
            while (cap < 2LL * deg) cap <<= 1;
            TK = gkey + tab_off[gi];
            TV = gval + tab_off[gi];
            if (threadIdx.x == 0) fb[mi] = 1;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < total; i += THREADS) {
            int lo = 0, hi = nch - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                const int pm = staged ? spre[mid] : posT[tm.x + mid] - r0;
                if (pm <= i) lo = mid; else hi = mid - 1;
            }
            const int pre = staged ? spre[lo] : posT[tm.x + lo] - r0;
            const int start = staged ? sstart[lo] : offs[(long long)lo * (P + 1)];
            const int2 kv = partials[(c0 + lo) * kHubChunk + start + (i - pre)];
            // LDS: cap >= 2 x the bucket's entries; global: cap >= 2 x the hub's degree
            const long long h = probe_claim(TK, cap, kv.x);
            if (h < 0) {
                *a.err = kErrProbe;
                continue;
            }
            atomicAdd(&TV[h], kv.y);
        }
        // Only the global table needs the device-scope fence (it waits for the atomics and
        // invalidates the vector L1); it is costly, so the LDS path, the common one, skips it.
        if (!in_lds) __threadfence();  // uniform per block
        __syncthreads();
        const int own = a.label[v];
        const int nw = a.vwgt[v];
        int conn_own = 0;
        Cand c{-1, 0u, 0x7fffffff, 0};
        for (long long i = threadIdx.x; i < cap; i += THREADS) {
            const int lab = TK[i];
            if (lab < 0) continue;
            if (!in_lds && bucket_of(lab, P) != b) continue;
            const int conn = TV[i];
            if (lab == own) {
                conn_own = conn;
                continue;
            }
            consider(a, v, nw, conn, lab, c);
        }
        block_reduce<THREADS>(conn_own, c);
        if (threadIdx.x == 0) {
            bres[task] = c;
            bown[task] = conn_own;
        }
        __syncthreads();
    }
}

// One workgroup per multi-chunk hub: the best of its bucket winners decides; a global table used
// by an oversized bucket is cleared here.
template <LpMode M>
__global__ __launch_bounds__(256) void hub_decide(
    LpArgs a, const int* __restrict__ list, const int* __restrict__ multi, int n_multi,
    const int* __restrict__ task_prefix, const long long* __restrict__ tab_off,
    int* __restrict__ gkey, int* __restrict__ gval, const Cand* __restrict__ bres,
    const int* __restrict__ bown, uint8_t* __restrict__ fb) {
    const int mi = blockIdx.x;
    if (mi >= n_multi) return;
    const int gi = multi[mi];
    const int v = list[gi];
    if (M == LpMode::Refine && a.active[v] == 0) return;  // uniform per block
    const int t0 = task_prefix[mi], t1 = task_prefix[mi + 1];
    int conn_own = 0;
    Cand c{-1, 0u, 0x7fffffff, 0};
    for (int t = t0 + threadIdx.x; t < t1; t += 256) {
        conn_own = max(conn_own, bown[t]);
        const Cand o = bres[t];
        if (better(o.conn, o.tie, o.lab, c.conn, c.tie, c.lab)) c = o;
    }
    const int deg = a.xadj[v + 1] - a.xadj[v];
    block_reduce<256>(conn_own, c);
    if (threadIdx.x == 0) {
        decide<M>(a, v, a.label[v], a.vwgt[v], conn_own, c);
        if (M == LpMode::Refine) atomicAdd(&a.trav[blockIdx.x & 63], (unsigned long long)deg);
    }
    if (fb[mi]) {  // uniform per block
        long long cap = 64;
        while (cap < 2LL * deg) cap <<= 1;
        int* TK = gkey + tab_off[gi];
        int* TV = gval + tab_off[gi];
        for (long long i = threadIdx.x; i < cap; i += 256) {
            TK[i] = -1;
            TV[i] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0) fb[mi] = 0;
    }
}

// ---------------------------------------------------------------- filter / admit / apply
// The filter / admit / apply kernels run over the node list of the sub-round; nodes without a
// proposal (prop_t < 0) return at once.
__device__ inline void filter_one(const LpArgs& a, int v) {
    const int t = a.prop_t[v];
    if (t < 0) return;
    const int own = a.label[v];
    const int so = a.size[own];
    const int ub = bound_of(a, own);
    if (so <= ub) return;  // not a departure from an overloaded block
    const int w = a.vwgt[v];
    const unsigned long long key = prio_key(a.prop_gain[v], v);
    const bool pass = (long long)so - (long long)a.out[own] >= (long long)ub ||
                      a.dtop[own] == key;
    if (pass) {
        atomicAdd(&a.inc[t], w);
        atomicMax(&a.top[t], key);
    } else {
        a.prop_ok[v] = 0;
    }
}

__device__ inline void admit_one(const LpArgs& a, int v) {
    const int t = a.prop_t[v];
    if (t < 0 || !a.prop_ok[v]) return;
    const int w = a.vwgt[v];
    const long long st = a.size[t];
    const long long bt = bound_of(a, t);
    const bool ok = st + (long long)a.inc[t] < bt ||
                    (a.top[t] == prio_key(a.prop_gain[v], v) && st + w < bt);
    if (!ok) a.prop_ok[v] = 0;
}

// Applies v's proposal if admitted and clears its scratch. Returns 0 if v did not move, 1 if it
// moved, 2 if it moved and (refinement) its row of more than kBigDegree entries is left to the
// caller's workgroup (old label in *big_from).
template <LpMode M>
__device__ inline int apply_one(const LpArgs& a, int v, int salt, int* big_from) {
    const int t = a.prop_t[v];
    if (t < 0) return 0;
    const int own = a.label[v];
    a.inc[t] = 0;
    a.top[t] = 0ull;
    if (M == LpMode::Refine) {
        a.out[own] = 0;
        a.dtop[own] = 0ull;
    }
    a.prop_t[v] = -1;
    if (!a.prop_ok[v]) return 0;
    const int w = a.vwgt[v];
    atomicSub(&a.size[own], w);
    atomicAdd(&a.size[t], w);
    a.label[v] = t;
    if (M == LpMode::Refine) {
        const int beg = a.xadj[v], end = a.xadj[v + 1];
        if (end - beg > kBigDegree) {
            *big_from = own;
            return 2;
        } else {
            atomicAdd(&a.trav[salt & 63], (unsigned long long)(end - beg));
            for (int e = beg; e < end; ++e) {
                const int u = a.adjncy[e];
                a.next_active[u] = 1;
                if (a.hub_cnt) {
                    const int slot = a.hub_slot[u];
                    if (slot >= 0) {
                        atomicSub(&a.hub_cnt[slot + own], a.adjwgt[e]);
                        atomicAdd(&a.hub_cnt[slot + t], a.adjwgt[e]);
                    }
                }
            }
        }
    }
    return 1;
}

// Frontier marking (and hub counter updates) of moved node v's row by the whole workgroup.
__device__ inline void mark_row(const LpArgs& a, int v, int from, int to) {
    const int beg = a.xadj[v], end = a.xadj[v + 1];
    if (threadIdx.x == 0) atomicAdd(&a.trav[(v + blockIdx.x) & 63], (unsigned long long)(end - beg));
    for (int e = beg + threadIdx.x; e < end; e += blockDim.x) {
        const int u = a.adjncy[e];
        a.next_active[u] = 1;
        if (a.hub_cnt) {
            const int slot = a.hub_slot[u];
            if (slot >= 0) {
                atomicSub(&a.hub_cnt[slot + from], a.adjwgt[e]);
                atomicAdd(&a.hub_cnt[slot + to], a.adjwgt[e]);
            }
        }
    }
}

__global__ void filter_departures(LpArgs a, const int* __restrict__ list, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) filter_one(a, list[i]);
}

__global__ void admit_moves(LpArgs a, const int* __restrict__ list, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) admit_one(a, list[i]);
}

template <LpMode M>
__global__ __launch_bounds__(256) void apply_moves(LpArgs a, const int* __restrict__ list,
                                                   int count) {
    __shared__ int block_moves;
    if (threadIdx.x == 0) block_moves = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int from;
    if (i < count && apply_one<M>(a, list[i], blockIdx.x + threadIdx.x, &from))
        atomicAdd(&block_moves, 1);
    __syncthreads();
    if (threadIdx.x == 0 && block_moves) atomicAdd(a.moves, block_moves);
}

// The refinement's filter / admit / apply run over the sub-round's compacted frontier of every
// list, the hub list included (lengths on the device; every proposer is active, and nodes off
// the frontier have no proposal, on which the three steps do nothing). Thread i of the
// grid-stride loop takes element i of the lists' concatenation.
struct SubLists {
    const int* list[kNumBins];
    const int* dcount[kNumBins];  // device-side length, or nullptr: hcount
    int hcount[kNumBins];
};
__device__ inline int sub_lengths(const SubLists& L, int* cnt) {
    int total = 0;
#pragma unroll
    for (int b = 0; b < kNumBins; ++b) {
        cnt[b] = L.dcount[b] ? *L.dcount[b] : L.hcount[b];
        total += cnt[b];
    }
    return total;
}
__device__ inline int sub_node(const SubLists& L, const int* cnt, int i) {
    // unrolled select (no dynamically indexed register array)
    const int* lp = L.list[0];
    bool found = false;
#pragma unroll
    for (int b = 0; b < kNumBins; ++b) {
        const bool here = !found && i < cnt[b];
        if (here) lp = L.list[b];
        found |= here;
        if (!found) i -= cnt[b];
    }
    return lp[i];
}

__global__ __launch_bounds__(256) void filter_departures_sub(LpArgs a, SubLists L) {
    int cnt[kNumBins];
    const int total = sub_lengths(L, cnt);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256)
        filter_one(a, sub_node(L, cnt, i));
}

__global__ __launch_bounds__(256) void admit_moves_sub(LpArgs a, SubLists L) {
    int cnt[kNumBins];
    const int total = sub_lengths(L, cnt);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256)
        admit_one(a, sub_node(L, cnt, i));
}

// Moved nodes of more than kBigDegree entries have their rows marked by the whole workgroup
// right after each 256-node step (at most one such node per thread and step).
__global__ __launch_bounds__(256) void apply_moves_sub(LpArgs a, SubLists L) {
    __shared__ int block_moves;
    __shared__ int nbig;
    __shared__ int bigv[256], bigfrom[256], bigto[256];
    if (threadIdx.x == 0) block_moves = 0;
    int cnt[kNumBins];
    const int total = sub_lengths(L, cnt);
    int mine = 0;
    for (int base = blockIdx.x * 256; base < total; base += gridDim.x * 256) {  // uniform
        if (threadIdx.x == 0) nbig = 0;
        __syncthreads();
        const int i = base + threadIdx.x;
        if (i < total) {
            const int v = sub_node(L, cnt, i);
            int from;
            const int r = apply_one<LpMode::Refine>(a, v, blockIdx.x + threadIdx.x, &from);
            mine += r != 0;
            if (r == 2) {
                const int k = atomicAdd(&nbig, 1);
                bigv[k] = v;
                bigfrom[k] = from;
                bigto[k] = a.label[v];
            }
        }
        __syncthreads();
        const int nb = nbig;
        for (int k = 0; k < nb; ++k) mark_row(a, bigv[k], bigfrom[k], bigto[k]);
        __syncthreads();  // the step's big movers are done before nbig is reset
    }
    if (mine) atomicAdd(&block_moves, mine);
    __syncthreads();
    if (threadIdx.x == 0 && block_moves) atomicAdd(a.moves, block_moves);
}

__global__ void set_hub_slots(const int* __restrict__ list, const int* __restrict__ multi,
                              int n_multi, const int* __restrict__ dense_off,
                              int* __restrict__ hub_slot) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_multi) return;
    const int i = multi[t];
    hub_slot[list[i]] = dense_off[i];
}

// ---------------------------------------------------------------- binning
__device__ inline int bin_of(const BinConfig& cfg, int deg) {
#pragma unroll
    for (int b = 0; b < kNumBins - 1; ++b)
        if (deg <= cfg.hi[b]) return b;
    return kNumBins - 1;
}

// counts[l] / edges[l]: nodes and edges of list l = sub-round * kNumBins + bin; with split,
// counts[kMaxSub * kNumBins + sr] also receives the wave-bin nodes of degree <= cfg.wave_small.
__global__ void bin_count(int n, const int* __restrict__ xadj, uint32_t seed, int S,
                          BinConfig cfg, const uint8_t* __restrict__ include, int split,
                          int* __restrict__ counts, unsigned long long* __restrict__ edges) {
    constexpr int L = kMaxSub * kNumBins;
    __shared__ int local[L + kMaxSub];
    __shared__ unsigned long long ledges[L];
    for (int i = threadIdx.x; i < L + kMaxSub; i += blockDim.x) {
        local[i] = 0;
        if (i < L) ledges[i] = 0;
    }
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && (!include || include[v])) {
        const int deg = xadj[v + 1] - xadj[v];
        if (deg > 0) {
            const int sr = subround_of(seed, v, S);
            const int b = bin_of(cfg, deg);
            const int l = sr * kNumBins + b;
            atomicAdd(&local[l], 1);
            atomicAdd(&ledges[l], (unsigned long long)deg);
            if (split && b == kNumBins - 2 && deg <= cfg.wave_small) atomicAdd(&local[L + sr], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < S * kNumBins; i += blockDim.x) {
        if (local[i]) {
            atomicAdd(&counts[i], local[i]);
            atomicAdd(&edges[i], ledges[i]);
        }
    }
    for (int i = threadIdx.x; i < S; i += blockDim.x)
        if (local[L + i]) atomicAdd(&counts[L + i], local[L + i]);
}

// With split, a wave list holds its nodes of degree <= cfg.wave_small first (counts[L + sr] of
// them, cursor slot L + sr), then the others.
__global__ void bin_scatter(int n, const int* __restrict__ xadj, uint32_t seed, int S,
                            BinConfig cfg, const uint8_t* __restrict__ include, int split,
                            const int* __restrict__ counts, const int* __restrict__ offsets,
                            int* __restrict__ cursor, int* __restrict__ lists) {
    constexpr int L = kMaxSub * kNumBins;
    __shared__ int local[L + kMaxSub];
    __shared__ int base[L + kMaxSub];
    for (int i = threadIdx.x; i < L + kMaxSub; i += blockDim.x) local[i] = 0;
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    int l = -1, slot = -1, r = 0, skip = 0;
    if (v < n && (!include || include[v])) {
        const int deg = xadj[v + 1] - xadj[v];
        if (deg > 0) {
            const int sr = subround_of(seed, v, S);
            const int b = bin_of(cfg, deg);
            l = sr * kNumBins + b;
            slot = l;
            if (split && b == kNumBins - 2) {
                if (deg <= cfg.wave_small) slot = L + sr;  // front part of the wave list
                else skip = counts[L + sr];            // after the small ones
            }
            r = atomicAdd(&local[slot], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L + kMaxSub; i += blockDim.x)
        base[i] = local[i] ? atomicAdd(&cursor[i], local[i]) : 0;
    __syncthreads();
    if (l >= 0) lists[offsets[l] + skip + base[slot] + r] = v;
}

// Frontier of a refinement round: the active nodes of every non-hub list, compacted per list
// into clists (same offsets as lists); ccount[l] receives list l's length. The order inside a
// list is free: the synchronous schedule does not depend on it.
constexpr int kCompactItems = 8;  // list entries per thread of compact_active
__global__ __launch_bounds__(256) void compact_active(const int* __restrict__ lists,
                                                      const int* __restrict__ offsets, int nseg,
                                                      int total, const uint8_t* __restrict__ active,
                                                      int* __restrict__ ccount,
                                                      int* __restrict__ clists) {
    __shared__ int off[kMaxSub * kNumBins];
    __shared__ int local[kMaxSub * kNumBins];
    __shared__ int base[kMaxSub * kNumBins];
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) {
        off[i] = offsets[i];
        local[i] = 0;
    }
    __syncthreads();
    // kCompactItems entries per thread, 256 apart (coalesced), their flag gathers in flight
    // together; one global count update per list and workgroup
    const int i0 = blockIdx.x * 256 * kCompactItems + threadIdx.x;
    int v[kCompactItems], l[kCompactItems], r[kCompactItems];
#pragma unroll
    for (int j = 0; j < kCompactItems; ++j) {
        const int i = i0 + 256 * j;
        v[j] = i < total ? lists[i] : -1;
    }
#pragma unroll
    for (int j = 0; j < kCompactItems; ++j) {
        const int i = i0 + 256 * j;
        l[j] = -1;
        if (v[j] >= 0 && active[v[j]]) {  // every list, the hub list included (its frontier
                                          // feeds the counter-based hub decisions)
            int lo = 0, hi = nseg - 1;  // last list with off <= i
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (off[mid] <= i) lo = mid; else hi = mid - 1;
            }
            l[j] = lo;
            r[j] = atomicAdd(&local[lo], 1);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nseg; j += blockDim.x)
        base[j] = local[j] ? atomicAdd(&ccount[j], local[j]) : 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCompactItems; ++j)
        if (l[j] >= 0) clists[off[l[j]] + base[l[j]] + r[j]] = v[j];
}

// Block descriptors of the hub kernels, built once per level: chunk c of the sub-round's hub
// list -> {hub index, chunk of the hub, chunks of the hub, node}.
__global__ void build_chunk_map(const int* __restrict__ list, int count,
                                const int* __restrict__ chunk_prefix, int4* __restrict__ cmap) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= count) return;
    const int c0 = chunk_prefix[gi], c1 = chunk_prefix[gi + 1];
    const int v = list[gi];
    for (int c = c0; c < c1; ++c) cmap[c] = make_int4(gi, c - c0, c1 - c0, v);
}

// Hub gi of the mid-size kernel -> mmap[midpos[gi]] = {hub index, 0, 1, node}.
__global__ void build_mid_map(const int* __restrict__ list, int count,
                              const int* __restrict__ midpos, int4* __restrict__ mmap) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= count) return;
    const int m = midpos[gi];
    if (m >= 0) mmap[m] = make_int4(gi, 0, 1, list[gi]);
}

// (hub, bucket) task t of a sub-round -> {multi-chunk hub index, bucket, buckets, hub index}.
// tmap[t] = {index of the bucket's first count in the bucket-major table, index one past its
// last, hub index, node}; tsub[t] = {multi-chunk hub index, bucket | buckets << 16}.
__global__ void build_task_map(const int* __restrict__ list, const int* __restrict__ multi,
                               int n_multi, const int* __restrict__ task_prefix,
                               const int* __restrict__ chunk_prefix,
                               const int* __restrict__ ct_base, int4* __restrict__ tmap,
                               int2* __restrict__ tsub) {
    const int mi = blockIdx.x * blockDim.x + threadIdx.x;
    if (mi >= n_multi) return;
    const int t0 = task_prefix[mi], t1 = task_prefix[mi + 1];
    const int gi = multi[mi];
    const int nch = chunk_prefix[gi + 1] - chunk_prefix[gi];
    const int cb = ct_base[gi];
    const int v = list[gi];
    for (int t = t0; t < t1; ++t) {
        const int b = t - t0;
        tmap[t] = make_int4(cb + b * nch, cb + (b + 1) * nch, gi, v);
        tsub[t] = make_int2(mi, b | ((t1 - t0) << 16));
    }
}

__global__ void hub_degrees(const int* __restrict__ list, int count, const int* __restrict__ xadj,
                            int* __restrict__ deg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) deg[i] = xadj[list[i] + 1] - xadj[list[i]];
}

__global__ void fill_i32(int* p, long long n, int value) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void fill_u64(unsigned long long* p, long long n, unsigned long long value) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void fill_i64(long long* p, long long n, long long value) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void fill_u8(uint8_t* p, long long n, uint8_t value) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void init_labels(int n, int* __restrict__ label, int* __restrict__ size,
                            const int* __restrict__ vwgt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) {
        label[v] = v;
        size[v] = vwgt[v];
    }
}

__global__ void block_weights_kernel(int n, const int* __restrict__ part,
                                     const int* __restrict__ vwgt, int k, int* __restrict__ W) {
    extern __shared__ int lw[];
    for (int i = threadIdx.x; i < k; i += blockDim.x) lw[i] = 0;
    __syncthreads();
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        atomicAdd(&lw[part[v]], vwgt[v]);
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x)
        if (lw[i]) atomicAdd(&W[i], lw[i]);
}

__global__ void block_weights_global(int n, const int* __restrict__ part,
                                     const int* __restrict__ vwgt, int* __restrict__ W) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicAdd(&W[part[v]], vwgt[v]);
}

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

// The G-lane proposal kernel of one list with `items` nodes per lane group: clustering launches
// one workgroup per virtual block, refinement strides over its device-side frontier count.
template <int G>
void launch_group(bool cl, int items, const LpArgs& a, const int* lst, int cnt, const int* dc,
                  hipStream_t s) {
    const int g = grid_for((long long)cnt * G, kBlock * items);
    const dim3 grid(cl ? g : std::min(g, kFrontierGrid));
    switch (items) {
        case 1:
            if (cl) hipLaunchKernelGGL((propose_group<LpMode::Cluster, G, 1>), grid, dim3(kBlock), 0, s, a, lst, cnt, nullptr);
            else hipLaunchKernelGGL((propose_group<LpMode::Refine, G, 1>), grid, dim3(kBlock), 0, s, a, lst, cnt, dc);
            break;
        case 4:
            if (cl) hipLaunchKernelGGL((propose_group<LpMode::Cluster, G, 4>), grid, dim3(kBlock), 0, s, a, lst, cnt, nullptr);
            else hipLaunchKernelGGL((propose_group<LpMode::Refine, G, 4>), grid, dim3(kBlock), 0, s, a, lst, cnt, dc);
            break;
        default:
            if (cl) hipLaunchKernelGGL((propose_group<LpMode::Cluster, G, 2>), grid, dim3(kBlock), 0, s, a, lst, cnt, nullptr);
            else hipLaunchKernelGGL((propose_group<LpMode::Refine, G, 2>), grid, dim3(kBlock), 0, s, a, lst, cnt, dc);
    }
}

void launch_frontier(int items, bool dense, const LpArgs& a, const FrontierLists& F, int nl,
                     int listed, hipStream_t s) {
    // virtual blocks of at least 4 nodes: a listed count bounds the grid
    const dim3 grid(std::min(grid_for(listed, kBlock / 64), kFrontierGrid));
    // dense counters of k <= kWaveDenseSmall labels fit a quarter of the LDS
    const bool small = dense && nl <= kWaveDenseSmall;
    switch (items * 4 + (dense ? 1 : 0) + (small ? 2 : 0)) {
        case 4: hipLaunchKernelGGL((propose_frontier<1, false>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 5: hipLaunchKernelGGL((propose_frontier<1, true>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 7: hipLaunchKernelGGL((propose_frontier<1, true, kWaveDenseSmall>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 16: hipLaunchKernelGGL((propose_frontier<4, false>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 17: hipLaunchKernelGGL((propose_frontier<4, true>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 19: hipLaunchKernelGGL((propose_frontier<4, true, kWaveDenseSmall>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 9: hipLaunchKernelGGL((propose_frontier<2, true>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        case 11: hipLaunchKernelGGL((propose_frontier<2, true, kWaveDenseSmall>), grid, dim3(kBlock), 0, s, a, F, nl); break;
        default: hipLaunchKernelGGL((propose_frontier<2, false>), grid, dim3(kBlock), 0, s, a, F, nl);
    }
}

}  // namespace

// ================================================================= host side

void fill_async(int* p, long long n, int value, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, value);
}
void fill_async(unsigned long long* p, long long n, unsigned long long value, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, value);
}
void fill_async(uint8_t* p, long long n, uint8_t value, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_u8, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, value);
}
void fill_async(long long* p, long long n, long long value, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_i64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, value);
}

void LpWorkspace::prepare(LpMode mode, const DGraph& g, int num_labels, int S, uint32_t seed,
                          const BinConfig& bins, hipStream_t s, const uint8_t* include) {
    KH_REQUIRE(S >= 1 && S <= kMaxSub, "sub-rounds must be in [1, 8]");
    KH_REQUIRE(bins.hi[kNumBins - 2] <= 512, "the wave bin holds degrees <= 512");
    KH_REQUIRE(bins.group_items == 1 || bins.group_items == 2 || bins.group_items == 4,
               "group_items must be 1, 2 or 4");
    KH_REQUIRE(bins.wave_small == 64 || bins.wave_small == 128 || bins.wave_small == 256,
               "wave_small must be 64, 128 or 256");
    bins_ = bins;
    KH_REQUIRE(g.n < (1 << 30) && num_labels < (1 << 30), "labels must be below 2^30");
    subrounds_ = S;
    num_labels_ = num_labels;
    dense_wave_ = mode == LpMode::Refine && num_labels <= kWaveDense;
    dense_hub_ = mode == LpMode::Refine && num_labels <= kHubDense;
    const int L = kMaxSub * kNumBins;
    lists_.reserve(g.n);
    if (mode == LpMode::Refine) {
        clists_.reserve(g.n);
        ccount_.reserve(L);
    }
    counts_.reserve(L + kMaxSub);
    cursor_.reserve(L + kMaxSub);
    offsets_.reserve(L);
    edges_.reserve(L);
    prop_t_.reserve(g.n);
    prop_gain_.reserve(g.n);
    prop_ok_.reserve(g.n);
    moves_.reserve(kMaxRounds);
    inc_.reserve(num_labels);
    out_.reserve(num_labels);
    top_.reserve(num_labels);
    dtop_.reserve(num_labels);
    active_.reserve(g.n);
    next_active_.reserve(g.n);
    n_big_tasks_.reserve((long long)kMaxSub * kMaxRounds);
    n_mid_tasks_.reserve((long long)kMaxSub * kMaxRounds);
    trav_.reserve(64);
    err_.reserve(1);

    fill_async(counts_.p, L + kMaxSub, 0, s);
    fill_async(cursor_.p, L + kMaxSub, 0, s);
    const int split = mode == LpMode::Cluster ? 1 : 0;  // small-degree wave kernel (clustering)
    fill_async(edges_.p, L, 0ull, s);
    fill_async(prop_t_.p, g.n, -1, s);
    fill_async(inc_.p, num_labels, 0, s);
    fill_async(out_.p, num_labels, 0, s);
    fill_async(top_.p, num_labels, 0ull, s);
    fill_async(dtop_.p, num_labels, 0ull, s);
    if (g.n > 0)
        hipLaunchKernelGGL(bin_count, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n, g.xadj,
                           seed, S, bins, include, split, counts_.p, edges_.p);
    KH_CHECK(hipMemcpyAsync(host_counts, counts_.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
    KH_CHECK(hipMemcpyAsync(host_wave_small_, counts_.p + L, sizeof(int) * kMaxSub,
                            hipMemcpyDeviceToHost, s));
    KH_CHECK(hipMemcpyAsync(host_edges, edges_.p, sizeof(long long) * L, hipMemcpyDeviceToHost,
                            s));
    KH_CHECK(hipStreamSynchronize(s));
    int acc = 0;
    for (int i = 0; i < L; ++i) {
        host_offsets[i] = acc;
        acc += (i < S * kNumBins) ? host_counts[i] : 0;
    }
    KH_CHECK(hipMemcpyAsync(offsets_.p, host_offsets, sizeof(int) * L, hipMemcpyHostToDevice, s));
    if (g.n > 0)
        hipLaunchKernelGGL(bin_scatter, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n, g.xadj,
                           seed, S, bins, include, split, counts_.p, offsets_.p, cursor_.p,
                           lists_.p);

    // hubs (last bin): chunk prefix per sub-round list, global tables for multi-chunk hubs
    int hubs = 0;
    for (int sr = 0; sr < S; ++sr) {
        hub_base_[sr] = hubs;
        hub_n_[sr] = host_counts[sr * kNumBins + kNumBins - 1];
        hubs += hub_n_[sr];
    }
    hub_total_ = 0;
    dense_total_ = 0;
    for (int sr = 0; sr < S; ++sr) {
        hub_chunks_[sr] = hub_multi_n_[sr] = mid_n_[sr] = 0;
        mid_edges_[sr] = 0;
    }
    if (hubs == 0) {
        KH_CHECK(hipGetLastError());
        return;
    }
    hub_deg_.reserve(hubs);
    for (int sr = 0; sr < S; ++sr) {
        const int l = sr * kNumBins + kNumBins - 1;
        if (hub_n_[sr] == 0) continue;
        hipLaunchKernelGGL(hub_degrees, dim3(grid_for(hub_n_[sr], 256)), dim3(256), 0, s,
                           lists_.p + host_offsets[l], hub_n_[sr], g.xadj,
                           hub_deg_.p + hub_base_[sr]);
    }
    std::vector<int> deg(hubs);
    KH_CHECK(hipMemcpyAsync(deg.data(), hub_deg_.p, sizeof(int) * hubs, hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    if (dense_hub_) {
        // chunked dense refinement: chunk prefix per sub-round and global counters for the hubs
        // that keep them (every hub when hubs x labels counters fit in kAllHubCounters, else the
        // hubs of more than one chunk)
        std::vector<int> prefix, multi, doff(hubs, -1);
        long long dtotal = 0;
        dense_all_ = (long long)hubs * num_labels <= kAllHubCounters;
        for (int sr = 0; sr < S; ++sr) {
            chunk_base_[sr] = (int)prefix.size();
            multi_base_[sr] = (int)multi.size();
            int c = 0;
            for (int i = 0; i < hub_n_[sr]; ++i) {
                const int nch = (deg[hub_base_[sr] + i] + kHubChunk - 1) / kHubChunk;
                prefix.push_back(c);
                c += nch;
                if (nch > 1 || dense_all_) {
                    doff[hub_base_[sr] + i] = (int)dtotal;
                    dtotal += num_labels;
                    multi.push_back(i);
                }
            }
            prefix.push_back(c);
            hub_chunks_[sr] = c;
            hub_multi_n_[sr] = (int)multi.size() - multi_base_[sr];
        }
        KH_REQUIRE(dtotal < (1LL << 31), "too many refinement hub counters");
        dense_total_ = dtotal;
        chunk_prefix_.reserve((long long)prefix.size());
        multi_.reserve(std::max<long long>((long long)multi.size(), 1));
        dense_off_.reserve(hubs);
        dense_cnt_.reserve(std::max<long long>(dtotal, 1));
        fill_async(dense_cnt_.p, std::max<long long>(dtotal, 1), 0, s);
        hub_slot_.reserve(g.n);
        fill_async(hub_slot_.p, g.n, -1, s);
        KH_CHECK(hipMemcpyAsync(chunk_prefix_.p, prefix.data(), sizeof(int) * prefix.size(),
                                hipMemcpyHostToDevice, s));
        if (!multi.empty())
            KH_CHECK(hipMemcpyAsync(multi_.p, multi.data(), sizeof(int) * multi.size(),
                                    hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(dense_off_.p, doff.data(), sizeof(int) * hubs,
                                hipMemcpyHostToDevice, s));
        build_chunk_maps(s);
        for (int sr = 0; sr < S; ++sr)
            if (hub_multi_n_[sr] > 0)
                hipLaunchKernelGGL(set_hub_slots, dim3(grid_for(hub_multi_n_[sr], 256)),
                                   dim3(256), 0, s,
                                   lists_.p + host_offsets[sr * kNumBins + kNumBins - 1],
                                   multi_.p + multi_base_[sr], hub_multi_n_[sr],
                                   dense_off_.p + hub_base_[sr], hub_slot_.p);
        KH_CHECK(hipStreamSynchronize(s));  // the host vectors must outlive the copies
        KH_CHECK(hipGetLastError());
        return;
    }
    std::vector<int> prefix, multi, tasks, midpos(hubs, -1);
    std::vector<long long> toff(hubs, 0), tcap(hubs, 0), obase(hubs, 0);
    std::vector<int> ctb(hubs, 0);
    int mids = 0;
    long long offm_total = 0, ct_total = 0, ct_sr_start = 0;
    int max_tasks = 0, max_multi = 0;
    for (int sr = 0; sr < S; ++sr) {
        chunk_base_[sr] = (int)prefix.size();
        multi_base_[sr] = (int)multi.size();
        task_base_[sr] = (int)tasks.size();
        int c = 0, t = 0;
        (void)0;
        mid_base_[sr] = mids;
        mid_edges_[sr] = 0;
        for (int i = 0; i < hub_n_[sr]; ++i) {
            const int d = deg[hub_base_[sr] + i];
            prefix.push_back(c);
            if (d <= kMidChunk) {  // the mid-size kernel, no chunk of the big one
                midpos[hub_base_[sr] + i] = mids++ - mid_base_[sr];
                mid_edges_[sr] += d;
                continue;
            }
            const int nch = (d + kHubChunk - 1) / kHubChunk;
            c += nch;
            if (nch > 1) {
                const int P = std::min(kBucketsPerChunk * nch, kMaxBuckets);
                long long cap = 64;
                while (cap < 2LL * d) cap <<= 1;
                toff[hub_base_[sr] + i] = hub_total_;
                tcap[hub_base_[sr] + i] = cap;
                hub_total_ += cap;
                obase[hub_base_[sr] + i] = offm_total;
                offm_total += (long long)nch * (P + 1);
                ctb[hub_base_[sr] + i] = (int)ct_total;
                ct_total += (long long)nch * P;
                multi.push_back(i);
                tasks.push_back(t);
                t += P;
            }
        }
        prefix.push_back(c);
        tasks.push_back(t);
        ct_sr_base_[sr] = (int)ct_sr_start;
        ct_sr_size_[sr] = (int)(ct_total - ct_sr_start);
        ct_total += 1;  // zero sentinel after the sub-round's counts
        ct_sr_start = ct_total;
        hub_chunks_[sr] = c;
        mid_n_[sr] = mids - mid_base_[sr];
        hub_multi_n_[sr] = (int)multi.size() - multi_base_[sr];
        hub_tasks_[sr] = t;
        max_tasks = std::max(max_tasks, t);
        max_multi = std::max(max_multi, hub_multi_n_[sr]);
    }
    chunk_prefix_.reserve((long long)prefix.size());
    tab_off_.reserve(hubs);
    tab_cap_.reserve(hubs);
    offm_base_.reserve(hubs);
    task_prefix_.reserve((long long)tasks.size());
    multi_.reserve((long long)multi.size());
    offm_.reserve(offm_total);
    KH_REQUIRE(ct_total < (1LL << 31), "hub bucket table exceeds 2^31 entries");
    ct_base_.reserve(hubs);
    cnt_t_.reserve(ct_total);
    pos_t_.reserve(ct_total);
    fill_async(cnt_t_.p, ct_total, 0, s);  // sentinels stay zero
    KH_CHECK(hipMemcpyAsync(ct_base_.p, ctb.data(), sizeof(int) * hubs, hipMemcpyHostToDevice,
                            s));
    {
        size_t bytes = 0, need = 0;
        for (int sr = 0; sr < S; ++sr) {
            KH_CHECK(rocprim::exclusive_scan(nullptr, bytes, cnt_t_.p, pos_t_.p, 0,
                                             (size_t)ct_sr_size_[sr] + 1, rocprim::plus<int>(), s));
            need = std::max(need, bytes);
        }
        scan_tmp_.reserve((long long)need + 1);
        scan_tmp_bytes_ = need;
    }
    bres_.reserve(max_tasks);
    bown_.reserve(max_tasks);
    big_tasks_.reserve(max_tasks);
    mid_tasks_.reserve(max_tasks);
    fb_.reserve(max_multi);
    fill_async(fb_.p, max_multi, (uint8_t)0, s);
    KH_CHECK(hipMemcpyAsync(chunk_prefix_.p, prefix.data(), sizeof(int) * prefix.size(),
                            hipMemcpyHostToDevice, s));
    KH_CHECK(hipMemcpyAsync(tab_off_.p, toff.data(), sizeof(long long) * hubs,
                            hipMemcpyHostToDevice, s));
    KH_CHECK(hipMemcpyAsync(tab_cap_.p, tcap.data(), sizeof(long long) * hubs,
                            hipMemcpyHostToDevice, s));
    KH_CHECK(hipMemcpyAsync(offm_base_.p, obase.data(), sizeof(long long) * hubs,
                            hipMemcpyHostToDevice, s));
    KH_CHECK(hipMemcpyAsync(task_prefix_.p, tasks.data(), sizeof(int) * tasks.size(),
                            hipMemcpyHostToDevice, s));
    if (!multi.empty())
        KH_CHECK(hipMemcpyAsync(multi_.p, multi.data(), sizeof(int) * multi.size(),
                                hipMemcpyHostToDevice, s));
    build_chunk_maps(s);
    mmap_.reserve(std::max(mids, 1));
    midpos_.reserve(hubs);
    KH_CHECK(hipMemcpyAsync(midpos_.p, midpos.data(), sizeof(int) * hubs, hipMemcpyHostToDevice,
                            s));
    for (int sr = 0; sr < S; ++sr) {
        KH_REQUIRE(mid_base_[sr] + mid_n_[sr] <= std::max(mids, 1), "mid hub map overflow");
        if (mid_n_[sr] > 0)
            hipLaunchKernelGGL(build_mid_map, dim3(grid_for(hub_n_[sr], 256)), dim3(256), 0, s,
                               lists_.p + host_offsets[sr * kNumBins + kNumBins - 1], hub_n_[sr],
                               midpos_.p + hub_base_[sr], mmap_.p + mid_base_[sr]);
    }
    {
        long long all_tasks = 0;  // one descriptor per (hub, bucket) task of every sub-round
        for (int sr = 0; sr < S; ++sr) all_tasks += hub_tasks_[sr];
        tmap_.reserve(std::max<long long>(all_tasks, 1));
        tsub_.reserve(std::max<long long>(all_tasks, 1));
        long long tb = 0;
        for (int sr = 0; sr < S; ++sr) {
            KH_REQUIRE(tb + hub_tasks_[sr] <= all_tasks, "task map overflow");
            if (hub_multi_n_[sr] > 0)
                hipLaunchKernelGGL(build_task_map, dim3(grid_for(hub_multi_n_[sr], 256)), dim3(256),
                                   0, s, lists_.p + host_offsets[sr * kNumBins + kNumBins - 1],
                                   multi_.p + multi_base_[sr], hub_multi_n_[sr],
                                   task_prefix_.p + task_base_[sr],
                                   chunk_prefix_.p + chunk_base_[sr], ct_base_.p + hub_base_[sr],
                                   tmap_.p + tb, tsub_.p + tb);
            tmap_base_[sr] = (int)tb;
            tb += hub_tasks_[sr];
        }
    }
    if (hub_total_ > 0) {
        hub_key_.reserve(hub_total_);
        hub_val_.reserve(hub_total_);
        fill_async(hub_key_.p, hub_total_, -1, s);
        fill_async(hub_val_.p, hub_total_, 0, s);
    }
    int max_chunks = 0;
    for (int sr = 0; sr < S; ++sr) max_chunks = std::max(max_chunks, hub_chunks_[sr]);
    partials_.reserve((long long)max_chunks * kHubChunk);
    KH_CHECK(hipStreamSynchronize(s));  // the host vectors above must outlive the copies
    KH_CHECK(hipGetLastError());
}

void LpWorkspace::build_chunk_maps(hipStream_t s) {
    long long total = 0;
    for (int sr = 0; sr < subrounds_; ++sr) total += hub_chunks_[sr];
    cmap_.reserve(std::max<long long>(total, 1));
    long long cb = 0;
    for (int sr = 0; sr < subrounds_; ++sr) {
        KH_REQUIRE(cb + hub_chunks_[sr] <= total, "chunk map overflow");
        cmap_base_[sr] = (int)cb;
        if (hub_n_[sr] > 0)
            hipLaunchKernelGGL(build_chunk_map, dim3(grid_for(hub_n_[sr], 256)), dim3(256), 0, s,
                               lists_.p + host_offsets[sr * kNumBins + kNumBins - 1], hub_n_[sr],
                               chunk_prefix_.p + chunk_base_[sr], cmap_.p + cb);
        cb += hub_chunks_[sr];
    }
}

static const char* kBinName[kNumBins] = {"g8", "g16", "g32", "g64", "wave", "hub"};

void LpWorkspace::run(LpMode mode, const DGraph& g, int* label, int* size, int bound, int rounds,
                      uint32_t seed, hipStream_t s, Profiler* prof, const char* phase,
                      const int* lbound) {
    KH_REQUIRE(rounds >= 0 && rounds <= kMaxRounds, "LP rounds must be in [0, 64]");
    const int S = subrounds_;
    const bool cl = mode == LpMode::Cluster;
    KH_REQUIRE(cl || !(dense_wave_ || dense_hub_) || num_labels_ > 0, "prepare() not called");
    // no allocation here: prepare() sized everything
    fill_async(n_big_tasks_.p, (long long)kMaxSub * kMaxRounds, 0, s);
    fill_async(n_mid_tasks_.p, (long long)kMaxSub * kMaxRounds, 0, s);
    fill_async(moves_.p, kMaxRounds, 0, s);
    fill_async(trav_.p, 64, 0ull, s);
    fill_async(err_.p, 1, 0, s);
    if (!cl) {
        fill_async(active_.p, g.n, (uint8_t)1, s);
        fill_async(next_active_.p, g.n, (uint8_t)0, s);
    }
    LpArgs a{};
    a.xadj = g.xadj;
    a.adjncy = g.adjncy;
    a.adjwgt = g.adjwgt;
    a.vwgt = g.vwgt;
    a.label = label;
    a.size = size;
    a.bound = bound;
    a.lbound = lbound;
    a.inc = inc_.p;
    a.top = top_.p;
    a.out = out_.p;
    a.dtop = dtop_.p;
    a.prop_t = prop_t_.p;
    a.prop_gain = prop_gain_.p;
    a.prop_ok = prop_ok_.p;
    a.trav = trav_.p;
    a.err = err_.p;
    // Refinement hubs of more than one chunk keep per-block connection counters across the
    // rounds: one scan builds them, every applied move updates its hub neighbours' counters, and
    // the hubs then decide from the counters instead of rescanning their edges each round.
    const bool incremental = !cl && dense_hub_ && dense_total_ > 0;
    a.hub_slot = incremental ? hub_slot_.p : nullptr;
    a.hub_cnt = incremental ? dense_cnt_.p : nullptr;
    uint8_t* act = active_.p;
    uint8_t* nxt = next_active_.p;
    const std::string base = std::string(phase ? phase : (cl ? "lp_cluster" : "lp_refine"));
    if (incremental && rounds > 0) {
        ProfScope ps(prof, base + "_hub_counters", 0.0, 0, s);
        fill_async(dense_cnt_.p, dense_total_, 0, s);
        const int copies = std::min(kHubDenseThreads / 64, kHubDenseSlots / num_labels_);
        const size_t lds = sizeof(int) * (size_t)copies * num_labels_;
        for (int sr = 0; sr < S; ++sr) {
            if (hub_multi_n_[sr] == 0) continue;
            hipLaunchKernelGGL(propose_hub_dense_chunk, dim3(hub_chunks_[sr]), dim3(kHubDenseThreads), lds, s, a, cmap_.p + cmap_base_[sr], num_labels_, dense_off_.p + hub_base_[sr], dense_cnt_.p, 1, 0);
        }
    }
    const int nl = num_labels_;
    // every refinement hub decides from kept counters: those decisions join the frontier launch
    const bool fused_hubs = incremental && dense_all_;
    const int nseg = S * kNumBins;
    const int listed = host_offsets[nseg - 1] + host_counts[nseg - 1];
    for (int r = 0; r < rounds; ++r) {
        a.tie_seed = tie_seed_of(seed, r);
        a.moves = moves_.p + r;
        a.active = cl ? nullptr : act;
        a.next_active = cl ? nullptr : nxt;
        if (!cl && listed > 0) {
            ProfScope ps(prof, base + "_frontier", 0.0, 0, s);
            fill_async(ccount_.p, nseg, 0, s);
            hipLaunchKernelGGL(compact_active, dim3(grid_for(listed, 256 * kCompactItems)), dim3(256), 0, s,
                               lists_.p, offsets_.p, nseg, listed, act, ccount_.p, clists_.p);
        }
        for (int sr = 0; sr < S; ++sr) {
            int sub_total = 0;
            if (!cl) {
                // refinement: the five degree lists' proposals in one launch over the frontier,
                // and the hubs' decisions from their kept counters when every hub has them
                FrontierLists F{};
                int listed_sr = 0;
                for (int b = 0; b < kNumBins; ++b) {
                    const int l = sr * kNumBins + b;
                    const bool take = b < kNumBins - 1 || fused_hubs;
                    F.list[b] = clists_.p + host_offsets[l];
                    F.dcount[b] = take ? ccount_.p + l : nullptr;
                    listed_sr += take ? host_counts[l] : 0;
                }
                if (listed_sr > 0)
                    launch_frontier(bins_.group_items, dense_wave_, a, F, nl, listed_sr, s);
            }
            for (int b = 0; b < kNumBins; ++b) {
                const int l = sr * kNumBins + b;
                const int cnt = host_counts[l];
                sub_total += cnt;
                if (cnt == 0) continue;
                if (!cl && (b < kNumBins - 1 || fused_hubs)) continue;  // in the frontier launch
                // refinement: the round's frontier of the list (length on the device)
                const bool fr = !cl && b < kNumBins - 1;
                const int* lst = (fr ? clists_.p : lists_.p) + host_offsets[l];
                const int* dc = fr ? ccount_.p + l : nullptr;
                // algorithmic bytes: 12 B per traversed edge (target, weight, label gather) and
                // 16 B per node (two offsets, weight, own label); label weights are O(moved).
                // (the hub list's mid-size hubs are timed in a scope of their own)
                const bool hub = b == kNumBins - 1;
                const long long ledges = host_edges[l] - (hub ? mid_edges_[sr] : 0);
                const int lnodes = cnt - (hub ? mid_n_[sr] : 0);
                const double bytes = cl ? 12.0 * (double)ledges + 16.0 * lnodes : 0.0;
                ProfScope ps(prof, base + "_propose_" + kBinName[b], bytes, cl ? ledges : 0, s);
                switch (b) {
                    case 0: launch_group<8>(cl, bins_.group_items, a, lst, cnt, dc, s); break;
                    case 1: launch_group<16>(cl, bins_.group_items, a, lst, cnt, dc, s); break;
                    case 2: launch_group<32>(cl, bins_.group_items, a, lst, cnt, dc, s); break;
                    case 3: launch_group<64>(cl, bins_.group_items, a, lst, cnt, dc, s); break;
                    case 4:
                        if (cl) {
                            const int ns = host_wave_small_[sr];  // degree <= wave_small, first
                            if (ns > 0) {
                                const dim3 gs(grid_for(ns, kBlock / 64));
                                if (bins_.wave_small == 64) hipLaunchKernelGGL((propose_wave<LpMode::Cluster, false, 128>), gs, dim3(kBlock), 0, s, a, lst, ns, nl, nullptr);
                                else if (bins_.wave_small == 256) hipLaunchKernelGGL((propose_wave<LpMode::Cluster, false, 512>), gs, dim3(kBlock), 0, s, a, lst, ns, nl, nullptr);
                                else hipLaunchKernelGGL((propose_wave<LpMode::Cluster, false, 256>), gs, dim3(kBlock), 0, s, a, lst, ns, nl, nullptr);
                            }
                            if (cnt > ns) hipLaunchKernelGGL((propose_wave<LpMode::Cluster, false>), dim3(grid_for(cnt - ns, kBlock / 64)), dim3(kBlock), 0, s, a, lst + ns, cnt - ns, nl, nullptr);
                        }
                        else if (dense_wave_) hipLaunchKernelGGL((propose_wave<LpMode::Refine, true>), dim3(std::min(grid_for(cnt, kBlock / 64), kFrontierGrid)), dim3(kBlock), 0, s, a, lst, cnt, nl, dc);
                        else hipLaunchKernelGGL((propose_wave<LpMode::Refine, false>), dim3(std::min(grid_for(cnt, kBlock / 64), kFrontierGrid)), dim3(kBlock), 0, s, a, lst, cnt, nl, dc);
                        break;
                    default: {
                        if (!cl && dense_hub_) {
                            const int4* cm = cmap_.p + cmap_base_[sr];
                            const int* dof = dense_off_.p + hub_base_[sr];
                            const int copies = std::min(kHubDenseThreads / 64, kHubDenseSlots / nl);
                            const size_t lds = sizeof(int) * (size_t)copies * nl;
                            if (!(incremental && dense_all_))  // otherwise every hub decides from counters
                                hipLaunchKernelGGL(propose_hub_dense_chunk, dim3(hub_chunks_[sr]), dim3(kHubDenseThreads), lds, s, a, cm, nl, dof, dense_cnt_.p, 0, incremental ? 1 : 0);
                            ps.close();
                            if (hub_multi_n_[sr] > 0) {
                                ProfScope pf(prof, base + "_propose_hub_merge", 0.0, hub_multi_n_[sr], s);
                                hipLaunchKernelGGL(hub_dense_decide, dim3(hub_multi_n_[sr]), dim3(256), 0, s, a, lst, multi_.p + multi_base_[sr], hub_multi_n_[sr], nl, dof, dense_cnt_.p, incremental ? 1 : 0);
                            }
                            break;
                        }
                        const int4* cm = cmap_.p + cmap_base_[sr];
                        const long long* to = tab_off_.p + hub_base_[sr];
                        const long long* ob = offm_base_.p + hub_base_[sr];
                        const int* ctb = ct_base_.p + hub_base_[sr];
                        if (hub_chunks_[sr] > 0) {
                            if (cl) hipLaunchKernelGGL((propose_hub_chunk<LpMode::Cluster, kHubThreads, kHubChunk>), dim3(hub_chunks_[sr]), dim3(kHubThreads), 0, s, a, cm, partials_.p, ob, offm_.p, ctb, cnt_t_.p);
                            else hipLaunchKernelGGL((propose_hub_chunk<LpMode::Refine, kHubThreads, kHubChunk>), dim3(hub_chunks_[sr]), dim3(kHubThreads), 0, s, a, cm, partials_.p, ob, offm_.p, ctb, cnt_t_.p);
                        }
                        ps.close();  // the hub scope times the chunked kernel alone
                        if (mid_n_[sr] > 0) {
                            const double mb = cl ? 12.0 * (double)mid_edges_[sr] + 16.0 * mid_n_[sr] : 0.0;
                            ProfScope pm(prof, base + "_propose_hub_mid", mb, cl ? mid_edges_[sr] : 0, s);
                            const int4* mm = mmap_.p + mid_base_[sr];
                            if (cl) hipLaunchKernelGGL((propose_hub_chunk<LpMode::Cluster, kMidThreads, kMidChunk>), dim3(mid_n_[sr]), dim3(kMidThreads), 0, s, a, mm, partials_.p, ob, offm_.p, ctb, cnt_t_.p);
                            else hipLaunchKernelGGL((propose_hub_chunk<LpMode::Refine, kMidThreads, kMidChunk>), dim3(mid_n_[sr]), dim3(kMidThreads), 0, s, a, mm, partials_.p, ob, offm_.p, ctb, cnt_t_.p);
                        }
                        if (hub_multi_n_[sr] > 0) {
                            ProfScope pf(prof, base + "_propose_hub_merge", 0.0, hub_tasks_[sr], s);
                            const int* ml = multi_.p + multi_base_[sr];
                            const int* tp = task_prefix_.p + task_base_[sr];
                            const int4* tm = tmap_.p + tmap_base_[sr];
                            const int2* tsb = tsub_.p + tmap_base_[sr];
                            const int nm = hub_multi_n_[sr];
                            Cand* br = reinterpret_cast<Cand*>(bres_.p);
                            const int nt = hub_tasks_[sr];
                            // bucket-major positions, relative to this sub-round's block of counts
                            size_t sb = scan_tmp_bytes_;
                            KH_CHECK(rocprim::exclusive_scan(scan_tmp_.p, sb, cnt_t_.p + ct_sr_base_[sr], pos_t_.p + ct_sr_base_[sr], 0, (size_t)ct_sr_size_[sr] + 1, rocprim::plus<int>(), s));
                            const int* pt = pos_t_.p;
                            const int* cp = chunk_prefix_.p + chunk_base_[sr];
                            int* bt = big_tasks_.p;
                            int* nbt = n_big_tasks_.p + (r * S + sr);
                            int* mt = mid_tasks_.p;
                            int* nmt = n_mid_tasks_.p + (r * S + sr);
                            if (cl) {
                                hipLaunchKernelGGL((hub_bucket_merge<LpMode::Cluster, 0>), dim3(nt), dim3(kTinyThreads), 0, s, a, nt, tm, tsb, pt, cp, ob, offm_.p, partials_.p, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p, nullptr, nullptr, mt, nmt);
                                hipLaunchKernelGGL((hub_bucket_merge<LpMode::Cluster, 1>), dim3(std::min(nt, 1024)), dim3(kSmallThreads), 0, s, a, nt, tm, tsb, pt, cp, ob, offm_.p, partials_.p, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p, mt, nmt, bt, nbt);
                                hipLaunchKernelGGL((hub_bucket_merge<LpMode::Cluster, 2>), dim3(std::min(nt, 256)), dim3(kFinThreads), 0, s, a, nt, tm, tsb, pt, cp, ob, offm_.p, partials_.p, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p, bt, nbt, nullptr, nullptr);
                                hipLaunchKernelGGL((hub_decide<LpMode::Cluster>), dim3(nm), dim3(256), 0, s, a, lst, ml, nm, tp, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p);
                            } else {
                                hipLaunchKernelGGL((hub_bucket_merge<LpMode::Refine, 0>), dim3(nt), dim3(kTinyThreads), 0, s, a, nt, tm, tsb, pt, cp, ob, offm_.p, partials_.p, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p, nullptr, nullptr, mt, nmt);
                                hipLaunchKernelGGL((hub_bucket_merge<LpMode::Refine, 1>), dim3(std::min(nt, 1024)), dim3(kSmallThreads), 0, s, a, nt, tm, tsb, pt, cp, ob, offm_.p, partials_.p, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p, mt, nmt, bt, nbt);
                                hipLaunchKernelGGL((hub_bucket_merge<LpMode::Refine, 2>), dim3(std::min(nt, 256)), dim3(kFinThreads), 0, s, a, nt, tm, tsb, pt, cp, ob, offm_.p, partials_.p, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p, bt, nbt, nullptr, nullptr);
                                hipLaunchKernelGGL((hub_decide<LpMode::Refine>), dim3(nm), dim3(256), 0, s, a, lst, ml, nm, tp, to, hub_key_.p, hub_val_.p, br, bown_.p, fb_.p);
                            }
                        }
                        break;
                    }
                }
            }
            if (sub_total == 0) continue;
            const int gm = grid_for(sub_total, 256);
            const int* sub_list = lists_.p + host_offsets[sr * kNumBins];  // bins are contiguous
            ProfScope ps(prof, base + "_admit_apply", 0.0, 0, s);
            if (cl) {
                hipLaunchKernelGGL(admit_moves, dim3(gm), dim3(256), 0, s, a, sub_list, sub_total);
                hipLaunchKernelGGL((apply_moves<LpMode::Cluster>), dim3(gm), dim3(256), 0, s, a, sub_list, sub_total);
            } else {
                SubLists L{};
                for (int b = 0; b < kNumBins; ++b) {
                    const int l = sr * kNumBins + b;
                    // every list's compacted frontier (an inactive hub proposes nothing)
                    L.list[b] = clists_.p + host_offsets[l];
                    L.dcount[b] = ccount_.p + l;
                    L.hcount[b] = host_counts[l];
                }
                const int gs = std::min(gm, 1024);
                hipLaunchKernelGGL(filter_departures_sub, dim3(gs), dim3(256), 0, s, a, L);
                hipLaunchKernelGGL(admit_moves_sub, dim3(gs), dim3(256), 0, s, a, L);
                hipLaunchKernelGGL(apply_moves_sub, dim3(gs), dim3(256), 0, s, a, L);
            }
        }
        if (!cl) {
            std::swap(act, nxt);
            fill_async(nxt, g.n, (uint8_t)0, s);
        }
    }
    KH_CHECK(hipGetLastError());
}

void LpWorkspace::read_moves(int rounds, int* out, hipStream_t s) {
    if (rounds <= 0) return;
    KH_CHECK(hipMemcpyAsync(out, moves_.p, sizeof(int) * rounds, hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
}

void LpWorkspace::check(hipStream_t s) {
    int e = 0;
    KH_CHECK(hipMemcpyAsync(&e, err_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    KH_REQUIRE(e == 0, "label propagation: a hash table probe found no free slot (kernel table "
                       "sizing invariant violated); the result is invalid");
}

long long LpWorkspace::traversed(hipStream_t s) {
    unsigned long long h[64];
    KH_CHECK(hipMemcpyAsync(h, trav_.p, sizeof(h), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    long long t = 0;
    for (int i = 0; i < 64; ++i) t += (long long)h[i];
    return t;
}

void LpWorkspace::init_cluster_labels(const DGraph& g, int* label, int* size, hipStream_t s) {
    if (g.n > 0)
        hipLaunchKernelGGL(init_labels, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n, label,
                           size, g.vwgt);
}

void compute_block_weights(const DGraph& g, const int* part, int k, int* W, hipStream_t s) {
    KH_REQUIRE(k >= 1, "k must be >= 1");
    fill_async(W, k, 0, s);
    if (g.n == 0) return;
    if (k > 8192) {  // per-workgroup LDS histograms only up to 32 KiB
        hipLaunchKernelGGL(block_weights_global, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n,
                           part, g.vwgt, W);
        return;
    }
    const int grid = std::min(grid_for(g.n, 256), 2048);
    hipLaunchKernelGGL(block_weights_kernel, dim3(grid), dim3(256), sizeof(int) * k, s, g.n,
                       part, g.vwgt, k, W);
}

}  // namespace kahip
