// Cluster contraction on gfx950: first-appearance relabel (reference:
// lib/partition/coarsening/clustering/size_constraint_label_propagation.cpp:207-232) and the
// quotient graph of a clustering (reference: contraction.cpp:94-133 ->
// lib/partition/uncoarsening/refinement/quotient_graph_refinement/complete_boundary.h:511-562).
//
// The reference scans the members of coarse node p in ascending id and their edges in CSR order
// and emits each neighbouring cluster q at its first appearance. Since the members' CSR rows are
// contiguous and increasing in node id, that order is the order of min{e : src(e) in p,
// cluster(adjncy[e]) = q}. The kernels aggregate (sum of weights, first appearance) per (p, q) in
// any thread order and emit each row in first-appearance order: bit-exact by construction.
//
//   remap       leader[c] = atomicMin over members; flags of leaders; exclusive scan -> new ids.
//   count       coarse weights, member counts, summed member degrees fd (row capacity bound);
//               the members of every coarse node in ascending id (a stable radix sort by
//               cluster), so that the row's edge t = 0, 1, ... grows with the fine edge index.
//   rows_thread one thread per row of fd <= 8 (rows of fd = 0 need no work);
//   rows_wave   one wave per row of fd <= 128 / <= 512: LDS hash (q -> w, min t), a stage
//               indexed by t that marks each entry at its first appearance, one ordered pass;
//   rows_block  the same with one workgroup per row of fd <= 1024 (256 threads) / <= 4096 (1024);
//   rows_global_* rows of larger fd: their edges in chunks of 4096 over many workgroups: LDS
//               pre-aggregation, a merge per (row, bucket) that stages each merged entry at the
//               flattened index of its first edge, and a compaction of the staged entries in
//               that index order (= ascending min e, no sort).
//   compact     rows are gathered into the final CSR after a scan of the coarse degrees.
// Rows are binned by fd, an upper bound of their distinct neighbours, so no tier overflows.

#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "contract.hpp"
#include "lp.hpp"

namespace kahip {

namespace {

constexpr int kWaveTierCap[2] = {2 * 128, 2 * 512};  // LDS slots of the one-wave tiers
constexpr int kMidThreads = 256;     // one-workgroup tiers: rows of fd <= 1024 (24 KB of hash,
constexpr int kMidCap = 2 * 1024;    // 8 KB of member staging, 5 per CU) ...
constexpr int kBigThreads = 1024;    // ... and of fd <= 4096 (96 KB + 32 KB)
constexpr int kBigCap = 2 * 4096;
constexpr int kThreadTierFd = 8;                    // largest fd of the thread tier
constexpr int kTierFd[4] = {128, 512, 1024, 4096};  // largest fd of each LDS tier
static_assert(2 * kTierFd[0] == kWaveTierCap[0] && 2 * kTierFd[1] == kWaveTierCap[1] &&
                  2 * kTierFd[2] == kMidCap && 2 * kTierFd[3] == kBigCap,
              "every LDS tier holds rows of up to half its table in member edges");
constexpr int kGChunk = 4096;      // member edges per workgroup in the global tier
constexpr int kGCap = 8192;        // LDS hash slots of a global-tier chunk
constexpr int kGThreads = 1024;  // 4 member edges per thread
constexpr int kGMergeThreads = 1024;
static_assert(2 * kGChunk <= kGCap, "a global-tier chunk fills at most half its table");

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

__global__ void leader_min(int n, const int* __restrict__ label, int* __restrict__ leader) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicMin(&leader[label[v]], v);
}

__global__ void leader_flag(int n, const int* __restrict__ label, const int* __restrict__ leader,
                            int* __restrict__ flag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = leader[label[v]] == v ? 1 : 0;
}

__global__ void apply_newid(int n, const int* __restrict__ label, const int* __restrict__ leader,
                            const int* __restrict__ newid, int* __restrict__ map) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) map[v] = newid[leader[label[v]]];
}

__global__ void count_members(int n, const int* __restrict__ xadj, const int* __restrict__ vwgt,
                              const int* __restrict__ map, int* __restrict__ cnt,
                              int* __restrict__ fdeg, int* __restrict__ cvwgt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int p = map[v];
    atomicAdd(&cnt[p], 1);
    atomicAdd(&fdeg[p], xadj[v + 1] - xadj[v]);
    atomicAdd(&cvwgt[p], vwgt[v]);
}

__global__ void iota_kernel(int n, int* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = v;
}

// Rows by summed member degree fd (an upper bound of the row's distinct neighbours), so that no
// tier can overflow: fd = 0 needs no work (cdeg stays 0); fd <= 8 goes to the thread tier,
// fd <= 128 and fd <= 512 to the one-wave tiers, fd <= 1024 and fd <= 4096 to the
// one-workgroup tiers, larger rows to the chunked global tier. Workgroup-aggregated list appends.
__global__ __launch_bounds__(256) void row_bins(int nc, const int* __restrict__ fdeg,
                                                int* __restrict__ l0, int* __restrict__ l1,
                                                int* __restrict__ l2, int* __restrict__ l3,
                                                int* __restrict__ l4, int* __restrict__ counts,
                                                int* __restrict__ global_list,
                                                int* __restrict__ global_count) {
    __shared__ int lc[6], bc[6];
    if (threadIdx.x < 6) lc[threadIdx.x] = 0;
    __syncthreads();
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int cls = -1, rank = 0;
    if (p < nc) {
        const int fd = fdeg[p];
        if (fd > 0) {
            cls = fd <= kThreadTierFd ? 0 : fd <= kTierFd[0] ? 1 : fd <= kTierFd[1] ? 2
                : fd <= kTierFd[2] ? 3 : fd <= kTierFd[3] ? 4 : 5;
            rank = atomicAdd(&lc[cls], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = lc[threadIdx.x];
        bc[threadIdx.x] = c ? atomicAdd(threadIdx.x < 5 ? &counts[threadIdx.x] : global_count, c) : 0;
    }
    __syncthreads();
    if (cls < 0) return;
    int* const lists[6] = {l0, l1, l2, l3, l4, global_list};
    lists[cls][bc[cls] + rank] = p;
}

// One thread per row of fd <= kThreadTierFd: the members (ascending id) and their edges are
// scanned in order, so a neighbour cluster is appended at its first appearance and later
// appearances add their weight (a linear search of at most kThreadTierFd entries, kept in LDS
// columns per thread).
__global__ __launch_bounds__(256) void rows_thread(
    const int* __restrict__ rows, int nrows, const int* __restrict__ xadj,
    const int* __restrict__ adjncy, const int* __restrict__ adjwgt, const int* __restrict__ map,
    const int* __restrict__ moff, const int* __restrict__ members, const int* __restrict__ tpos,
    int* __restrict__ cdeg, int* __restrict__ tq, int* __restrict__ tw) {
    __shared__ int Q[kThreadTierFd][256];
    __shared__ int W[kThreadTierFd][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nrows) return;  // no barrier below
    const int p = rows[i];
    const int t = threadIdx.x;
    int cnt = 0;  // distinct <= fd <= kThreadTierFd
    const int me = moff[p + 1];
    for (int m = moff[p]; m < me; ++m) {
        const int u = members[m];
        const int e1 = xadj[u + 1];
        for (int e = xadj[u]; e < e1; ++e) {
            const int q = map[adjncy[e]];
            if (q == p) continue;
            const int w = adjwgt[e];
            int j = 0;
            while (j < cnt && Q[j][t] != q) ++j;
            if (j == cnt) {
                Q[j][t] = q;
                W[j][t] = w;
                ++cnt;
            } else {
                W[j][t] += w;
            }
        }
    }
    const int off = tpos[p];
    for (int j = 0; j < cnt; ++j) {
        tq[off + j] = Q[j][t];
        tw[off + j] = W[j][t];
    }
    cdeg[p] = cnt;
}

// The two row kernels below aggregate a coarse row p in LDS: its members (ascending node id)
// are expanded with a load-balanced scan, and edge t of the row (t = 0, 1, ... over the members'
// CSR rows in order, so t grows with the fine edge index e) adds (q -> sum w, min t) to an LDS
// hash, q = cluster of the target (q != p). A row of fd member edges has at most fd distinct
// entries, and every tier takes only rows whose fd fits half its table. Then stage[min t] = slot
// marks each entry at its first appearance, and one ordered pass over t writes the row in the
// reference's order (complete_boundary.h:511-562) at its capacity offset tpos[p].

// One wave (= one workgroup) per row of fd <= CAP / 2.
template <int CAP>
__global__ __launch_bounds__(64) void rows_wave(
    const int* __restrict__ rows, int nrows, const int* __restrict__ xadj,
    const int* __restrict__ adjncy, const int* __restrict__ adjwgt, const int* __restrict__ map,
    const int* __restrict__ moff, const int* __restrict__ members, const int* __restrict__ fdeg,
    const int* __restrict__ tpos, int* __restrict__ cdeg, int* __restrict__ tq,
    int* __restrict__ tw, int* __restrict__ err) {
    constexpr int MAXD = CAP / 2;
    constexpr int kEpl = 4;
    __shared__ int K[CAP];
    __shared__ int V[CAP];
    __shared__ int M[CAP];
    __shared__ int stage[MAXD];
    __shared__ int mpre[64];
    __shared__ int mstart[64];
    if ((int)blockIdx.x >= nrows) return;  // uniform
    const int p = rows[blockIdx.x];
    const int lane = threadIdx.x;
    const int fd = fdeg[p];  // <= MAXD (row_bins)
    int cap = 64;
    while (cap < 2 * fd) cap <<= 1;
    for (int i = lane; i < cap; i += 64) {
        K[i] = -1;
        V[i] = 0;
        M[i] = 0x7fffffff;
    }
    for (int t = lane; t < fd; t += 64) stage[t] = -1;
    __syncthreads();
    const int mb = moff[p], me = moff[p + 1];
    int tbase = 0;
    for (int base = mb; base < me; base += 64) {
        const int mi = base + lane;
        int b = 0, d = 0;
        if (mi < me) {
            const int u = members[mi];
            b = xadj[u];
            d = xadj[u + 1] - b;
        }
        int incl = d;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const int total = __shfl(incl, 63, 64);
        mpre[lane] = incl;
        mstart[lane] = b;
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += 64 * kEpl) {
            int es[kEpl], qs[kEpl], ws[kEpl];
#pragma unroll
            for (int j = 0; j < kEpl; ++j) {
                const int t = t0 + lane + 64 * j;
                es[j] = -1;
                if (t < total) {
                    int lo = 0, hi = 63;  // first member whose inclusive prefix exceeds t
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (mpre[mid] > t) hi = mid; else lo = mid + 1;
                    }
                    es[j] = mstart[lo] + (t - (lo ? mpre[lo - 1] : 0));
                }
            }
#pragma unroll
            for (int j = 0; j < kEpl; ++j) {
                qs[j] = es[j] >= 0 ? __builtin_nontemporal_load(adjncy + es[j]) : 0;
                ws[j] = es[j] >= 0 ? __builtin_nontemporal_load(adjwgt + es[j]) : 0;
            }
#pragma unroll
            for (int j = 0; j < kEpl; ++j)
                if (es[j] >= 0) qs[j] = map[qs[j]];
#pragma unroll
            for (int j = 0; j < kEpl; ++j) {
                const int q = qs[j];
                if (es[j] < 0 || q == p) continue;
                const int h = probe_claim(K, cap, q);  // distinct <= fd <= cap / 2
                if (h < 0) {
                    *err = kErrProbe;
                    continue;
                }
                atomicAdd(&V[h], ws[j]);
                atomicMin(&M[h], tbase + t0 + lane + 64 * j);
            }
        }
        tbase += total;
        __syncthreads();  // mpre / mstart are rewritten by the next batch
    }
    for (int i = lane; i < cap; i += 64)
        if (K[i] >= 0) stage[M[i]] = i;  // M[i] < fd: a first appearance is unique
    __syncthreads();
    const int off = tpos[p];
    int cnt = 0;
    for (int c = 0; c < fd; c += 64) {
        const int t = c + lane;
        const int slot = t < fd ? stage[t] : -1;
        const unsigned long long m = __ballot(slot >= 0);
        if (slot >= 0) {
            const int at = off + cnt + __popcll(m & ((1ull << lane) - 1));
            tq[at] = K[slot];
            tw[at] = V[slot];
        }
        cnt += __popcll(m);
    }
    if (lane == 0) cdeg[p] = cnt;
}

// One workgroup of THREADS per row of fd <= CAP / 2 (grid-stride over the list). The members
// are expanded MAXD at a time: S holds each member's (inclusive degree prefix, row start), every
// thread takes every THREADS-th edge of the batch and gathers kEpl targets at a time, so the
// target and cluster loads of many edges are in flight together. After the hash phase S holds
// the stage (fd <= MAXD ints).
template <int THREADS, int CAP>
__global__ __launch_bounds__(THREADS) void rows_block(
    const int* __restrict__ list, const int* __restrict__ list_count, const int* __restrict__ xadj,
    const int* __restrict__ adjncy, const int* __restrict__ adjwgt, const int* __restrict__ map,
    const int* __restrict__ moff, const int* __restrict__ members, const int* __restrict__ fdeg,
    const int* __restrict__ tpos, int* __restrict__ cdeg, int* __restrict__ tq,
    int* __restrict__ tw, int* __restrict__ err) {
    constexpr int MAXD = CAP / 2;
    constexpr int kEpl = 4;
    constexpr int R = MAXD / THREADS;  // members per thread in a batch's prefix scan
    static_assert(MAXD % THREADS == 0, "table shape");
    __shared__ int K[CAP];
    __shared__ int V[CAP];
    __shared__ int M[CAP];
    __shared__ unsigned long long S[MAXD];
    __shared__ int wsum[THREADS / 64];
    int* const stage = reinterpret_cast<int*>(S);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int count = *list_count;
    for (int li = blockIdx.x; li < count; li += gridDim.x) {
        const int p = list[li];
        const int fd = fdeg[p];  // <= MAXD (row_bins)
        int cap = 64;
        while (cap < 2 * fd) cap <<= 1;
        for (int i = threadIdx.x; i < cap; i += THREADS) {
            K[i] = -1;
            V[i] = 0;
            M[i] = 0x7fffffff;
        }
        __syncthreads();
        const int mb = moff[p], me = moff[p + 1];
        int tbase = 0;
        for (int base = mb; base < me; base += MAXD) {
            // (inclusive degree prefix, row start) of the batch's members
            const int nb = min(MAXD, me - base);
            int st[R], dg[R], run = 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int i = threadIdx.x * R + j;
                st[j] = 0;
                dg[j] = 0;
                if (i < nb) {
                    const int u = members[base + i];
                    st[j] = xadj[u];
                    dg[j] = xadj[u + 1] - st[j];
                }
                run += dg[j];
            }
            int incl = run;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            if (lane == 63) wsum[wave] = incl;
            __syncthreads();
            int acc = incl - run, total = 0;
            for (int w = 0; w < THREADS / 64; ++w) {
                if (w < wave) acc += wsum[w];
                total += wsum[w];
            }
#pragma unroll
            for (int j = 0; j < R; ++j) {
                acc += dg[j];
                S[threadIdx.x * R + j] = ((unsigned long long)(uint32_t)acc << 32) | (uint32_t)st[j];
            }
            __syncthreads();
            for (int t0 = 0; t0 < total; t0 += THREADS * kEpl) {
                int es[kEpl], qs[kEpl], ws[kEpl];
#pragma unroll
                for (int j = 0; j < kEpl; ++j) {
                    const int t = t0 + threadIdx.x + THREADS * j;
                    es[j] = -1;
                    if (t < total) {
                        int lo = 0, hi = nb - 1;  // first member whose inclusive prefix exceeds t
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if ((int)(S[mid] >> 32) > t) hi = mid; else lo = mid + 1;
                        }
                        const int before = lo ? (int)(S[lo - 1] >> 32) : 0;
                        es[j] = (int)(uint32_t)S[lo] + (t - before);
                    }
                }
#pragma unroll
                for (int j = 0; j < kEpl; ++j) {
                    qs[j] = es[j] >= 0 ? __builtin_nontemporal_load(adjncy + es[j]) : 0;
                    ws[j] = es[j] >= 0 ? __builtin_nontemporal_load(adjwgt + es[j]) : 0;
                }
#pragma unroll
                for (int j = 0; j < kEpl; ++j)
                    if (es[j] >= 0) qs[j] = map[qs[j]];
#pragma unroll
                for (int j = 0; j < kEpl; ++j) {
                    const int q = qs[j];
                    if (es[j] < 0 || q == p) continue;
                    const int h = probe_claim(K, cap, q);  // distinct <= fd <= cap / 2
                    if (h < 0) {
                        *err = kErrProbe;
                        continue;
                    }
                    atomicAdd(&V[h], ws[j]);
                    atomicMin(&M[h], tbase + t0 + (int)threadIdx.x + THREADS * j);
                }
            }
            tbase += total;
            __syncthreads();  // S is rewritten by the next batch (or becomes the stage)
        }
        for (int t = threadIdx.x; t < fd; t += THREADS) stage[t] = -1;
        __syncthreads();
        for (int i = threadIdx.x; i < cap; i += THREADS)
            if (K[i] >= 0) stage[M[i]] = i;  // M[i] < fd: a first appearance is unique
        __syncthreads();
        // ordered compaction: thread k owns t in [k PER, k PER + PER), then a workgroup scan
        constexpr int PMAX = MAXD / THREADS;
        const int per = (fd + THREADS - 1) / THREADS;  // <= PMAX
        int sl[PMAX];
        int mine = 0;
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            const int t = (int)threadIdx.x * per + j;
            sl[j] = j < per && t < fd ? stage[t] : -1;
            mine += sl[j] >= 0;
        }
        int inc2 = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc2, o, 64);
            if (lane >= o) inc2 += y;
        }
        if (lane == 63) wsum[wave] = inc2;
        __syncthreads();
        int at = inc2 - mine, cnt = 0;
        for (int w = 0; w < THREADS / 64; ++w) {
            if (w < wave) at += wsum[w];
            cnt += wsum[w];
        }
        at += tpos[p];
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            if (sl[j] < 0) continue;
            tq[at] = K[sl[j]];
            tw[at] = V[sl[j]];
            ++at;
        }
        if (threadIdx.x == 0) cdeg[p] = cnt;
        __syncthreads();  // the tables, the stage and wsum are reused by the next row
    }
}

__global__ void member_degrees(int n, const int* __restrict__ members,
                               const int* __restrict__ xadj, int* __restrict__ mdeg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mdeg[i] = xadj[members[i] + 1] - xadj[members[i]];
}

// Rows beyond the LDS tiers, split into chunks of kGChunk member edges: the row's members are
// sorted by node id and their edges flattened through the member-degree prefix mpre, so the
// flattened index f of an edge grows with its CSR index e within a row. Row i has
// P_i = min(chunks, kGMaxBuckets) buckets of neighbour clusters q (bucket = gbucket(q, P_i)).
//   gchunk   each chunk aggregates (q -> sum w, min f) in LDS and writes its entries grouped by
//            bucket (per-chunk bucket offsets offm, bucket-major counts cntT); it also clears
//            the stage slots of its f range;
//   gmerge   one workgroup per (row, bucket) reads the bucket's segment of every chunk of the
//            row (located through the scan of cntT), sums them in LDS and stores every merged
//            entry (q, w) in the stage at f = its min f; tasks of more than kGMergeMax entries
//            are queued for gmerge_global (a private global table per task);
//   gemit    counts the staged entries of each chunk, then (after a scan over the chunks)
//            writes them in f order, which is the order of the first appearance of q in the
//            reference's scan of the row (complete_boundary.h:511-562).
// No global atomics on shared tables: every merge is local to one workgroup.
constexpr int kGMaxBuckets = 1024;
constexpr int kGMergeCap = 8192;  // LDS slots of gmerge
constexpr int kGMergeMax = kGMergeCap / 2;

__device__ inline int gbucket(int q, int P) {
    return (int)(mix32((uint32_t)q ^ 0x27d4eb2fu) % (uint32_t)P);
}

// Per-chunk descriptors of the global tier: da = {row index ri, chunk of the row, first member
// overlapping the chunk m0, members overlapping it nm}, db = {coarse node p, chunks of the row,
// f0, f1} (the chunk's edges [f0, f1) in the flattened member-edge space). One thread per chunk,
// so the binary searches run side by side instead of serially at the start of every workgroup.
__global__ void gchunk_desc(const int* __restrict__ crow, int nchunks, const int* __restrict__ list,
                            const int* __restrict__ chunk_prefix, const int* __restrict__ moff,
                            const int* __restrict__ mpre, int4* __restrict__ da,
                            int4* __restrict__ db) {
    const int chunk = blockIdx.x * blockDim.x + threadIdx.x;
    if (chunk >= nchunks) return;
    const int ri = crow[chunk];
    const int p = list[ri];
    const int c0 = chunk_prefix[ri];
    const int nch = chunk_prefix[ri + 1] - c0;
    const int mb = moff[p], me = moff[p + 1];
    const long long f_row = mpre[mb];
    const long long f_end = mpre[me];
    const long long f0 = f_row + (long long)(chunk - c0) * kGChunk;
    const long long f1 = min(f_end, f0 + kGChunk);
    int lo = mb, hi = me;  // largest member index with mpre <= f0
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((long long)mpre[mid] <= f0) lo = mid; else hi = mid;
    }
    int lo2 = lo, hi2 = me;  // largest member index with mpre < f1
    while (hi2 - lo2 > 1) {
        const int mid = (lo2 + hi2) >> 1;
        if ((long long)mpre[mid] < f1) lo2 = mid; else hi2 = mid;
    }
    da[chunk] = make_int4(ri, chunk - c0, lo, lo2 - lo + 1);
    db[chunk] = make_int4(p, nch, (int)f0, (int)f1);  // f < m2 < 2^31
}

__global__ __launch_bounds__(kGThreads) void gchunk(
    const int4* __restrict__ da, const int4* __restrict__ db, const int* __restrict__ xadj,
    const int* __restrict__ adjncy, const int* __restrict__ adjwgt, const int* __restrict__ map,
    const int* __restrict__ moff, const int* __restrict__ members, const int* __restrict__ mpre,
    const long long* __restrict__ offm_base, int* __restrict__ offm,
    const long long* __restrict__ ct_base, int* __restrict__ cntT, int4* __restrict__ partials,
    int* __restrict__ stage_q, int* __restrict__ err) {
    __shared__ int K[kGCap];
    __shared__ int V[kGCap];
    __shared__ int M[kGCap];
    __shared__ int bcur[kGMaxBuckets + 1];
    __shared__ int wsum[kGThreads / 64];
    __shared__ int sbeg[kGChunk];  // edge-space start (mpre) of the chunk's members
    __shared__ int sst[kGChunk];   // their CSR row start
    const int chunk = blockIdx.x;
    const int4 A = da[chunk], B = db[chunk];
    for (int i = threadIdx.x; i < kGCap; i += kGThreads) {
        K[i] = -1;
        V[i] = 0;
        M[i] = 0x7fffffff;
    }
    const int ri = A.x, c_idx = A.y;
    const int p = B.x, nch = B.y;
    const int P = min(nch, kGMaxBuckets);
    const int mb = moff[p], me = moff[p + 1];
    const long long f0 = B.z, f1 = B.w;
    // The members overlapping [f0, f1) are staged in LDS; thread t then takes the chunk's edges
    // t, t + kGThreads, ... and finds each one's member by a binary search in LDS, so
    // neighbouring threads load neighbouring edges. (Members of degree 0 share their start with
    // the next member; more than kGChunk members in a chunk can only come from those, and such a
    // chunk searches the member prefix in global memory instead.)
    constexpr int PER = kGChunk / kGThreads;
    const int m0 = A.z, nm = A.w;
    const bool staged = nm <= kGChunk;  // uniform
    // clear the stage of [f0, f1) (the merge kernels fill it after this one); a sequential
    // 4-byte store per edge measured faster than a mark bitmap set with atomics by the merges
    for (long long f = f0 + threadIdx.x; f < f1; f += kGThreads) stage_q[f] = -1;
    if (staged)
        for (int i = threadIdx.x; i < nm; i += kGThreads) {
            sbeg[i] = mpre[m0 + i];
            sst[i] = xadj[members[m0 + i]];
        }
    __syncthreads();
    int es[PER], qs[PER], ws[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const long long f = f0 + threadIdx.x + (long long)j * kGThreads;
        es[j] = -1;
        if (f >= f1) continue;
        if (staged) {
            int lo = 0, hi = nm - 1;  // last staged member starting at or before f
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((long long)sbeg[mid] <= f) lo = mid; else hi = mid - 1;
            }
            es[j] = sst[lo] + (int)(f - sbeg[lo]);
        } else {
            int lo = mb, hi = me;  // largest member index with mpre <= f
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if ((long long)mpre[mid] <= f) lo = mid; else hi = mid;
            }
            es[j] = xadj[members[lo]] + (int)(f - mpre[lo]);
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        qs[j] = es[j] >= 0 ? __builtin_nontemporal_load(adjncy + es[j]) : 0;
        ws[j] = es[j] >= 0 ? __builtin_nontemporal_load(adjwgt + es[j]) : 0;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j)
        if (es[j] >= 0) qs[j] = map[qs[j]];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int q = qs[j];
        if (es[j] < 0 || q == p) continue;
        const int h = probe_claim(K, kGCap, q);  // <= kGChunk entries in kGCap slots
        if (h < 0) {
            *err = kErrProbe;
            continue;
        }
        atomicAdd(&V[h], ws[j]);
        atomicMin(&M[h], (int)(f0 + threadIdx.x + (long long)j * kGThreads));  // this edge's f
    }
    for (int i = threadIdx.x; i <= P; i += kGThreads) bcur[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < kGCap; i += kGThreads)
        if (K[i] >= 0) atomicAdd(&bcur[gbucket(K[i], P)], 1);
    __syncthreads();
    // exclusive scan of bcur[0..P): each thread owns a run of buckets
    const int per = (P + kGThreads - 1) / kGThreads;
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
    int* offs = offm + offm_base[ri] + (long long)c_idx * (P + 1);
    int* cnt_t = cntT + ct_base[ri] + c_idx;  // bucket-major: [b * nch + c_idx]
    int tot = 0;
    for (int w = 0; w < kGThreads / 64; ++w) tot += wsum[w];
    for (int b = threadIdx.x; b < P; b += kGThreads) {
        offs[b] = bcur[b];
        cnt_t[(long long)b * nch] = (b + 1 < P ? bcur[b + 1] : tot) - bcur[b];
    }
    if (threadIdx.x == 0) offs[P] = tot;
    __syncthreads();
    int4* out = partials + (long long)chunk * kGChunk;
    for (int i = threadIdx.x; i < kGCap; i += kGThreads) {
        const int q = K[i];
        if (q < 0) continue;
        out[atomicAdd(&bcur[gbucket(q, P)], 1)] = make_int4(q, V[i], M[i], 0);
    }
}

// Entry i of merge task (row ri, bucket b), read where its chunk wrote it (no scatter pass):
// chunk c of the row holds entries [pre(c), pre(c + 1)) of the task, pre(c) = posT[base + c] -
// r0 with base = ct_base[ri] + b * nch (posT: exclusive scan of the bucket-major counts), at
// partials[(c0 + c) * kGChunk + start(c) + ...], start(c) = offm[offm_base[ri] + c (P + 1) + b].
// The chunk of entry i is the largest c with pre(c) <= i (empty segments are skipped).
constexpr int kGStage = 1024;  // rows of at most this many chunks stage pre / start in LDS
struct TaskSegs {
    const int* posT;  // + base
    const int* offs;  // offm + offm_base[ri] + b, stride P + 1
    long long stride;
    int r0, nch;
    long long c0;
};
template <bool STAGED>
__device__ inline int4 task_entry(const TaskSegs& t, const int* spre, const int* sstart,
                                  const int4* __restrict__ partials, int i) {
    int lo = 0, hi = t.nch - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int pm = STAGED ? spre[mid] : t.posT[mid] - t.r0;
        if (pm <= i) lo = mid; else hi = mid - 1;
    }
    const int pre = STAGED ? spre[lo] : t.posT[lo] - t.r0;
    const int start = STAGED ? sstart[lo] : t.offs[(long long)lo * t.stride];
    return partials[(t.c0 + lo) * kGChunk + start + (i - pre)];
}

__global__ __launch_bounds__(kGMergeThreads) void gmerge(
    const int* __restrict__ chunk_prefix, const int* __restrict__ task_prefix,
    const int* __restrict__ trow, int n_tasks, const long long* __restrict__ ct_base,
    const int* __restrict__ posT, const long long* __restrict__ offm_base,
    const int* __restrict__ offm, const int4* __restrict__ partials, int* __restrict__ stage_q,
    int* __restrict__ stage_w, int* __restrict__ big_tasks, int* __restrict__ n_big,
    int* __restrict__ err) {
    __shared__ int K[kGMergeCap];
    __shared__ int V[kGMergeCap];
    __shared__ int M[kGMergeCap];
    __shared__ int spre[kGStage + 1];
    __shared__ int sstart[kGStage];
    const int task = blockIdx.x;
    if (task >= n_tasks) return;  // uniform
    const int ri = trow[task];
    TaskSegs ts;
    ts.c0 = chunk_prefix[ri];
    ts.nch = chunk_prefix[ri + 1] - (int)ts.c0;
    const int b = task - task_prefix[ri];
    const int P = min(ts.nch, kGMaxBuckets);
    ts.posT = posT + ct_base[ri] + (long long)b * ts.nch;
    ts.offs = offm + offm_base[ri] + b;
    ts.stride = P + 1;
    ts.r0 = ts.posT[0];
    const int total = ts.posT[ts.nch] - ts.r0;
    if (total > kGMergeMax) {
        if (threadIdx.x == 0) big_tasks[atomicAdd(n_big, 1)] = task;
        return;  // uniform
    }
    const bool staged = ts.nch <= kGStage;  // uniform
    if (staged)
        for (int c = threadIdx.x; c <= ts.nch; c += kGMergeThreads) {
            spre[c] = ts.posT[c] - ts.r0;
            if (c < ts.nch) sstart[c] = ts.offs[(long long)c * ts.stride];
        }
    int cap = 64;
    while (cap < 2 * total) cap <<= 1;
    for (int i = threadIdx.x; i < cap; i += kGMergeThreads) {
        K[i] = -1;
        V[i] = 0;
        M[i] = 0x7fffffff;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += kGMergeThreads) {
        const int4 e = staged ? task_entry<true>(ts, spre, sstart, partials, i)
                              : task_entry<false>(ts, spre, sstart, partials, i);
        const int h = probe_claim(K, cap, e.x);  // cap >= 2 x the task's entries
        if (h < 0) {
            *err = kErrProbe;
            continue;
        }
        atomicAdd(&V[h], e.y);
        atomicMin(&M[h], e.z);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cap; i += kGMergeThreads) {
        const int q = K[i];
        if (q < 0) continue;
        const int f = M[i];  // one merged entry per (row, q): its min f is unique
        stage_q[f] = q;
        stage_w[f] = V[i];
    }
}

// Oversized (row, bucket) ranges: one workgroup each with a private global hash table of
// tcap[t] slots at toff[t] (cleared beforehand), then the same append as gmerge.
__global__ __launch_bounds__(1024) void gmerge_global(
    const int* __restrict__ big_tasks, int n_big, const int* __restrict__ chunk_prefix,
    const int* __restrict__ task_prefix, const int* __restrict__ trow,
    const long long* __restrict__ ct_base, const int* __restrict__ posT,
    const long long* __restrict__ offm_base, const int* __restrict__ offm,
    const int4* __restrict__ partials, const long long* __restrict__ toff,
    const long long* __restrict__ tcap, int* __restrict__ TK, int* __restrict__ TV,
    int* __restrict__ TM, int* __restrict__ stage_q, int* __restrict__ stage_w,
    int* __restrict__ err) {
    const int t = blockIdx.x;
    if (t >= n_big) return;
    const int task = big_tasks[t];
    const int ri = trow[task];
    TaskSegs ts;
    ts.c0 = chunk_prefix[ri];
    ts.nch = chunk_prefix[ri + 1] - (int)ts.c0;
    const int b = task - task_prefix[ri];
    const int P = min(ts.nch, kGMaxBuckets);
    ts.posT = posT + ct_base[ri] + (long long)b * ts.nch;
    ts.offs = offm + offm_base[ri] + b;
    ts.stride = P + 1;
    ts.r0 = ts.posT[0];
    const int total = ts.posT[ts.nch] - ts.r0;
    const long long cap = tcap[t];
    int* K = TK + toff[t];
    int* V = TV + toff[t];
    int* Mn = TM + toff[t];
    for (int i = threadIdx.x; i < total; i += 1024) {
        const int4 e = task_entry<false>(ts, nullptr, nullptr, partials, i);
        const long long h = probe_claim(K, cap, e.x);  // cap >= 2 x the task's entries
        if (h < 0) {
            *err = kErrProbe;
            continue;
        }
        atomicAdd(&V[h], e.y);
        atomicMin(&Mn[h], e.z);
    }
    __threadfence();  // the table's atomics complete before it is read back (rare path)
    __syncthreads();
    for (long long i = threadIdx.x; i < cap; i += 1024) {
        const int q = K[i];
        if (q < 0) continue;
        const int f = Mn[i];
        stage_q[f] = q;
        stage_w[f] = V[i];
    }
}

__global__ void big_task_totals(const int* __restrict__ big_tasks, int n_big,
                                const int* __restrict__ chunk_prefix,
                                const int* __restrict__ task_prefix,
                                const int* __restrict__ trow,
                                const long long* __restrict__ ct_base,
                                const int* __restrict__ posT, int* __restrict__ totals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_big) return;
    const int task = big_tasks[t];
    const int ri = trow[task];
    const int nch = chunk_prefix[ri + 1] - chunk_prefix[ri];
    const int b = task - task_prefix[ri];
    totals[t] = posT[ct_base[ri] + (long long)(b + 1) * nch] - posT[ct_base[ri] + (long long)b * nch];
}

constexpr int kEmitThreads = 256;  // kGChunk / kEmitThreads = 16 staged slots per thread
static_assert(kGChunk % kEmitThreads == 0, "every emission thread takes whole slots");

// Staged entries of each chunk (one workgroup per chunk).
__global__ __launch_bounds__(kEmitThreads) void gemit_count(const int4* __restrict__ db,
                                                             const int* __restrict__ stage_q,
                                                             int* __restrict__ ecnt) {
    __shared__ int wsum[kEmitThreads / 64];
    const int chunk = blockIdx.x;
    const long long f0 = db[chunk].z, f1 = db[chunk].w;
    int c = 0;
    for (long long f = f0 + threadIdx.x; f < f1; f += kEmitThreads) c += stage_q[f] >= 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kEmitThreads / 64; ++w) t += wsum[w];
        ecnt[chunk] = t;
    }
}

// Writes each chunk's staged entries in f order at tpos[p] + (entries of the row's earlier
// chunks) + (rank in the chunk); the row's last chunk also writes its coarse degree. Thread t
// takes the chunk's slots [16 t, 16 t + 16), so ranks follow f.
__global__ __launch_bounds__(kEmitThreads) void gemit_write(
    const int4* __restrict__ da, const int4* __restrict__ db, const int* __restrict__ stage_q,
    const int* __restrict__ stage_w, const int* __restrict__ epos,
    const int* __restrict__ tpos, int* __restrict__ cdeg, int* __restrict__ tq,
    int* __restrict__ tw) {
    constexpr int PER = kGChunk / kEmitThreads;
    static_assert(PER <= 32, "a thread's staged slots fit one bit mask");
    __shared__ int wsum[kEmitThreads / 64];
    const int chunk = blockIdx.x;
    const int4 A = da[chunk], B = db[chunk];
    const int p = B.x;
    const long long f0 = B.z, f1 = B.w;
    const long long a0 = f0 + (long long)threadIdx.x * PER;
    // staged slots among the thread's [a0, min(a0 + PER, f1)), bit j = slot a0 + j
    unsigned bits = 0u;
#pragma unroll
    for (int j = 0; j < PER; ++j)
        if (a0 + j < f1 && stage_q[a0 + j] >= 0) bits |= 1u << j;
    const int mine = __popc(bits);
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if ((int)(threadIdx.x & 63) >= o) incl += y;
    }
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    int before = incl - mine;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += wsum[w];
    const int first = chunk - A.y;  // the row's first chunk
    int at = tpos[p] + (epos[chunk] - epos[first]) + before;
    while (bits) {
        const int j = __ffs(bits) - 1;
        bits &= bits - 1u;
        tq[at] = stage_q[a0 + j];
        tw[at] = stage_w[a0 + j];
        ++at;
    }
    if (threadIdx.x == 0 && A.y + 1 == B.y)  // the row's last chunk
        cdeg[p] = epos[chunk + 1] - epos[first];
}

__global__ void compact_rows(int nc, const int* __restrict__ tpos, const int* __restrict__ cxadj,
                             const int* __restrict__ tq, const int* __restrict__ tw,
                             int* __restrict__ cadj, int* __restrict__ cw) {
    const int p = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= nc) return;
    const int src = tpos[p], dst = cxadj[p], c = cxadj[p + 1] - dst;
    for (int j = lane; j < c; j += 64) {
        cadj[dst + j] = tq[src + j];
        cw[dst + j] = tw[src + j];
    }
}

__global__ void project_kernel(int n, const int* __restrict__ map,
                               const int* __restrict__ coarse_part, int* __restrict__ fine_part) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) fine_part[v] = coarse_part[map[v]];
}

__global__ void row_fdeg(const int* __restrict__ list, int count, const int* __restrict__ fdeg,
                         int* __restrict__ fd_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) fd_out[i] = fdeg[list[i]];
}

}  // namespace

// ----------------------------------------------------------------------------- host side

void ContractWorkspace::scan_i32(const int* in, int* out, int n, hipStream_t s) {
    size_t bytes = 0;
    KH_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(),
                                     s));
    char* t = scan_tmp_.reserve((long long)bytes + 1);
    KH_CHECK(rocprim::exclusive_scan(t, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s));
}

int ContractWorkspace::remap(const DGraph& g, const int* label, int* map, hipStream_t s,
                             Profiler* prof) {
    const int n = g.n;
    if (n == 0) return 0;
    ProfScope ps(prof, "remap", 16.0 * n, 0, s);
    leader_.reserve(n);
    flag_.reserve(n + 1);
    newid_.reserve(n + 1);
    fill_async(leader_.p, n, 0x7fffffff, s);
    hipLaunchKernelGGL(leader_min, dim3(grid_for(n, 256)), dim3(256), 0, s, n, label, leader_.p);
    hipLaunchKernelGGL(leader_flag, dim3(grid_for(n, 256)), dim3(256), 0, s, n, label, leader_.p,
                       flag_.p);
    fill_async(flag_.p + n, 1, 0, s);
    scan_i32(flag_.p, newid_.p, n + 1, s);
    hipLaunchKernelGGL(apply_newid, dim3(grid_for(n, 256)), dim3(256), 0, s, n, label, leader_.p,
                       newid_.p, map);
    int nc = 0;
    KH_CHECK(hipMemcpyAsync(&nc, newid_.p + n, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    return nc;
}

void ContractWorkspace::contract(const DGraph& g, const int* map, int nc, OwnedGraph& out,
                                 hipStream_t s, Profiler* prof) {
    const int n = g.n;
    out.alloc_nodes(nc);
    ProfScope ps_all(prof, "contract", 0.0, g.m2, s);
    cnt_.reserve(nc + 1);
    fdeg_.reserve(nc + 1);
    moff_.reserve(nc + 1);
    tpos_.reserve(nc + 1);
    cdeg_.reserve(nc + 1);
    members_.reserve(n);
    fill_async(cnt_.p, nc + 1, 0, s);
    fill_async(fdeg_.p, nc + 1, 0, s);
    fill_async(cdeg_.p, nc + 1, 0, s);
    fill_async(out.g.vwgt, nc, 0, s);
    if (n > 0) {
        hipLaunchKernelGGL(count_members, dim3(grid_for(n, 256)), dim3(256), 0, s, n, g.xadj,
                           g.vwgt, map, cnt_.p, fdeg_.p, out.g.vwgt);
        // members grouped by cluster, ascending node id inside a group (a stable radix sort of
        // the node ids by cluster)
        iota_.reserve(n);
        mkeys_.reserve(n);
        hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, iota_.p);
        int bits = 1;
        while (bits < 31 && (1LL << bits) < (long long)nc) ++bits;
        const unsigned* keys = reinterpret_cast<const unsigned*>(map);
        unsigned* keys_out = reinterpret_cast<unsigned*>(mkeys_.p);
        size_t bytes = 0;
        KH_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys_out, iota_.p, members_.p,
                                           (unsigned)n, 0, bits, s));
        char* t = scan_tmp_.reserve((long long)bytes + 1);
        KH_CHECK(rocprim::radix_sort_pairs(t, bytes, keys, keys_out, iota_.p, members_.p,
                                           (unsigned)n, 0, bits, s));
    }
    scan_i32(cnt_.p, moff_.p, nc + 1, s);
    scan_i32(fdeg_.p, tpos_.p, nc + 1, s);
    const long long m2 = g.m2;
    tq_.reserve(m2);
    tw_.reserve(m2);
    ovl2_list_.reserve(nc);
    ovl_count_.reserve(1);
    fill_async(ovl_count_.p, 1, 0, s);
    err_.reserve(1);
    fill_async(err_.p, 1, 0, s);
    if (nc > 0) {
        // algorithmic bytes of the row pass: 12 B per fine edge (target, weight, cluster of
        // target), 12 B per fine node (offsets, member id), 8 B per coarse entry written
        const double bytes = 12.0 * (double)m2 + 12.0 * n;
        for (auto& l : rlist_) l.reserve(nc);
        rcount_.reserve(5);
        fill_async(rcount_.p, 5, 0, s);
        hipLaunchKernelGGL(row_bins, dim3(grid_for(nc, 256)), dim3(256), 0, s, nc, fdeg_.p,
                           rlist_[0].p, rlist_[1].p, rlist_[2].p, rlist_[3].p, rlist_[4].p,
                           rcount_.p, ovl2_list_.p, ovl_count_.p);
        int rc[3] = {0, 0, 0};
        KH_CHECK(hipMemcpyAsync(rc, rcount_.p, sizeof(rc), hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        {
            ProfScope ps(prof, "contract_rows_lds", bytes, m2, s);
            if (rc[0] > 0)
                hipLaunchKernelGGL(rows_thread, dim3(grid_for(rc[0], 256)), dim3(256), 0, s,
                                   rlist_[0].p, rc[0], g.xadj, g.adjncy, g.adjwgt, map, moff_.p,
                                   members_.p, tpos_.p, cdeg_.p, tq_.p, tw_.p);
            if (rc[1] > 0)
                hipLaunchKernelGGL((rows_wave<kWaveTierCap[0]>), dim3(rc[1]), dim3(64), 0, s,
                                   rlist_[1].p, rc[1], g.xadj, g.adjncy, g.adjwgt, map, moff_.p,
                                   members_.p, fdeg_.p, tpos_.p, cdeg_.p, tq_.p, tw_.p, err_.p);
            if (rc[2] > 0)
                hipLaunchKernelGGL((rows_wave<kWaveTierCap[1]>), dim3(rc[2]), dim3(64), 0, s,
                                   rlist_[2].p, rc[2], g.xadj, g.adjncy, g.adjwgt, map, moff_.p,
                                   members_.p, fdeg_.p, tpos_.p, cdeg_.p, tq_.p, tw_.p, err_.p);
        }
        {
            // grid-stride over the device-side counts: five 256-thread workgroups per CU, one
            // 1024-thread workgroup per CU
            ProfScope ps(prof, "contract_rows_lds_big", 0.0, 0, s);
            hipLaunchKernelGGL((rows_block<kMidThreads, kMidCap>), dim3(256 * 5),
                               dim3(kMidThreads), 0, s, rlist_[3].p, rcount_.p + 3, g.xadj,
                               g.adjncy, g.adjwgt, map, moff_.p, members_.p, fdeg_.p, tpos_.p,
                               cdeg_.p, tq_.p, tw_.p, err_.p);
            hipLaunchKernelGGL((rows_block<kBigThreads, kBigCap>), dim3(256), dim3(kBigThreads),
                               0, s, rlist_[4].p, rcount_.p + 4, g.xadj, g.adjncy, g.adjwgt, map,
                               moff_.p, members_.p, fdeg_.p, tpos_.p, cdeg_.p, tq_.p, tw_.p,
                               err_.p);
        }
    }
    int novl = 0;
    KH_CHECK(hipMemcpyAsync(&novl, ovl_count_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    if (novl > 0) {
        ProfScope ps(prof, "contract_rows_global", 0.0, novl, s);
        ofd_.reserve(novl);
        hipLaunchKernelGGL(row_fdeg, dim3(grid_for(novl, 256)), dim3(256), 0, s, ovl2_list_.p,
                           novl, fdeg_.p, ofd_.p);
        std::vector<int> fd(novl);
        KH_CHECK(hipMemcpyAsync(fd.data(), ofd_.p, sizeof(int) * novl, hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        // per row: chunks of its member edges, buckets, count / offset table bases, tasks
        std::vector<int> cpre(novl + 1), tpre(novl + 1), crow, trow;
        std::vector<long long> ofb(novl), ctb(novl);
        long long offm_total = 0, ct_total = 0;
        int nchunks = 0, ntasks = 0;
        for (int i = 0; i < novl; ++i) {
            const int nch = std::max(1, (fd[i] + kGChunk - 1) / kGChunk);
            const int P = std::min(nch, kGMaxBuckets);
            cpre[i] = nchunks;
            tpre[i] = ntasks;
            ofb[i] = offm_total;
            ctb[i] = ct_total;
            nchunks += nch;
            ntasks += P;
            offm_total += (long long)nch * (P + 1);
            ct_total += (long long)nch * P;
            crow.insert(crow.end(), nch, i);  // row of every chunk and of every (row, bucket) task
            trow.insert(trow.end(), P, i);
        }
        cpre[novl] = nchunks;
        tpre[novl] = ntasks;
        KH_REQUIRE(ct_total < (1LL << 31), "contraction scratch exceeds 2^31 entries");
        gchunk_prefix_.reserve(novl + 1);
        gtask_prefix_.reserve(novl + 1);
        goffm_base_.reserve(novl);
        gct_base_.reserve(novl);
        KH_CHECK(hipMemcpyAsync(gchunk_prefix_.p, cpre.data(), sizeof(int) * (novl + 1),
                                hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(gtask_prefix_.p, tpre.data(), sizeof(int) * (novl + 1),
                                hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(goffm_base_.p, ofb.data(), sizeof(long long) * novl,
                                hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(gct_base_.p, ctb.data(), sizeof(long long) * novl,
                                hipMemcpyHostToDevice, s));
        gcrow_.reserve(nchunks);
        gtrow_.reserve(ntasks);
        KH_CHECK(hipMemcpyAsync(gcrow_.p, crow.data(), sizeof(int) * nchunks,
                                hipMemcpyHostToDevice, s));
        KH_CHECK(hipMemcpyAsync(gtrow_.p, trow.data(), sizeof(int) * ntasks,
                                hipMemcpyHostToDevice, s));
        goffm_.reserve(offm_total);
        gcnt_.reserve(ct_total + 1);
        gpos_.reserve(ct_total + 1);
        gpart_.reserve((long long)nchunks * kGChunk);
        gbig_.reserve(std::max(ntasks, 1));
        gnbig_.reserve(1);
        fill_async(gcnt_.p + ct_total, 1, 0, s);  // sentinel
        fill_async(gnbig_.p, 1, 0, s);
        mdeg_.reserve((long long)n + 1);
        mpre_.reserve((long long)n + 1);
        hipLaunchKernelGGL(member_degrees, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                           members_.p, g.xadj, mdeg_.p);
        fill_async(mdeg_.p + n, 1, 0, s);
        scan_i32(mdeg_.p, mpre_.p, n + 1, s);
        stage_q_.reserve(std::max<long long>(m2, 1));
        stage_w_.reserve(std::max<long long>(m2, 1));
        gda_.reserve(nchunks);
        gdb_.reserve(nchunks);
        hipLaunchKernelGGL(gchunk_desc, dim3(grid_for(nchunks, 256)), dim3(256), 0, s, gcrow_.p,
                           nchunks, ovl2_list_.p, gchunk_prefix_.p, moff_.p, mpre_.p, gda_.p,
                           gdb_.p);
        hipLaunchKernelGGL(gchunk, dim3(nchunks), dim3(kGThreads), 0, s, gda_.p, gdb_.p, g.xadj,
                           g.adjncy, g.adjwgt, map, moff_.p, members_.p, mpre_.p, goffm_base_.p,
                           goffm_.p, gct_base_.p, gcnt_.p, gpart_.p, stage_q_.p, err_.p);
        scan_i32(gcnt_.p, gpos_.p, (int)ct_total + 1, s);
        hipLaunchKernelGGL(gmerge, dim3(ntasks), dim3(kGMergeThreads), 0, s, gchunk_prefix_.p,
                           gtask_prefix_.p, gtrow_.p, ntasks, gct_base_.p, gpos_.p,
                           goffm_base_.p, goffm_.p, gpart_.p, stage_q_.p, stage_w_.p, gbig_.p,
                           gnbig_.p, err_.p);
        int nbig = 0;
        KH_CHECK(hipMemcpyAsync(&nbig, gnbig_.p, sizeof(int), hipMemcpyDeviceToHost, s));
        KH_CHECK(hipStreamSynchronize(s));
        if (nbig > 0) {
            // rare: a (row, bucket) range too large for LDS gets a private global table
            DevBuf<int> totals;
            totals.reserve(nbig);
            hipLaunchKernelGGL(big_task_totals, dim3(grid_for(nbig, 256)), dim3(256), 0, s,
                               gbig_.p, nbig, gchunk_prefix_.p, gtask_prefix_.p, gtrow_.p,
                               gct_base_.p, gpos_.p, totals.p);
            std::vector<int> tot(nbig);
            KH_CHECK(hipMemcpyAsync(tot.data(), totals.p, sizeof(int) * nbig,
                                    hipMemcpyDeviceToHost, s));
            KH_CHECK(hipStreamSynchronize(s));
            std::vector<long long> toff(nbig), tcap(nbig);
            long long ttotal = 0;
            for (int t = 0; t < nbig; ++t) {
                long long c = 64;
                while (c < 2LL * tot[t]) c <<= 1;
                toff[t] = ttotal;
                tcap[t] = c;
                ttotal += c;
            }
            hcap_.reserve(nbig);
            hoff_.reserve(nbig);
            KH_CHECK(hipMemcpyAsync(hoff_.p, toff.data(), sizeof(long long) * nbig,
                                    hipMemcpyHostToDevice, s));
            KH_CHECK(hipMemcpyAsync(hcap_.p, tcap.data(), sizeof(long long) * nbig,
                                    hipMemcpyHostToDevice, s));
            hk_.reserve(ttotal);
            hv_.reserve(ttotal);
            hm_.reserve(ttotal);
            fill_async(hk_.p, ttotal, -1, s);
            fill_async(hv_.p, ttotal, 0, s);
            fill_async(hm_.p, ttotal, 0x7fffffff, s);
            hipLaunchKernelGGL(gmerge_global, dim3(nbig), dim3(1024), 0, s, gbig_.p, nbig,
                               gchunk_prefix_.p, gtask_prefix_.p, gtrow_.p, gct_base_.p, gpos_.p,
                               goffm_base_.p, goffm_.p, gpart_.p, hoff_.p, hcap_.p, hk_.p, hv_.p,
                               hm_.p, stage_q_.p, stage_w_.p, err_.p);
            KH_CHECK(hipStreamSynchronize(s));  // host vectors above must outlive the copies
        }
        // emission in f order: counts per chunk, a scan over the chunks, ordered writes
        ecnt_.reserve((long long)nchunks + 1);
        epos_.reserve((long long)nchunks + 1);
        fill_async(ecnt_.p + nchunks, 1, 0, s);
        hipLaunchKernelGGL(gemit_count, dim3(nchunks), dim3(kEmitThreads), 0, s, gdb_.p,
                           stage_q_.p, ecnt_.p);
        scan_i32(ecnt_.p, epos_.p, nchunks + 1, s);
        hipLaunchKernelGGL(gemit_write, dim3(nchunks), dim3(kEmitThreads), 0, s, gda_.p, gdb_.p,
                           stage_q_.p, stage_w_.p, epos_.p, tpos_.p, cdeg_.p, tq_.p, tw_.p);
    }
    scan_i32(cdeg_.p, out.xadj_dev(), nc + 1, s);
    int mc = 0, err = 0;
    KH_CHECK(hipMemcpyAsync(&mc, out.xadj_dev() + nc, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipMemcpyAsync(&err, err_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    KH_REQUIRE(err == 0, "contraction: a hash table probe found no free slot (kernel table sizing "
                         "invariant violated); the result is invalid");
    out.alloc_edges(mc);
    if (nc > 0) {
        ProfScope ps(prof, "contract_compact", 16.0 * mc, 0, s);
        hipLaunchKernelGGL(compact_rows, dim3(grid_for(nc, 4)), dim3(256), 0, s, nc, tpos_.p,
                           out.g.xadj, tq_.p, tw_.p, out.g.adjncy, out.g.adjwgt);
    }
    KH_CHECK(hipGetLastError());
}

void project_partition(int n_fine, const int* map, const int* coarse_part, int* fine_part,
                       hipStream_t s, Profiler* prof) {
    if (n_fine == 0) return;
    ProfScope ps(prof, "project", 8.0 * n_fine, 0, s);
    hipLaunchKernelGGL(project_kernel, dim3(grid_for(n_fine, 256)), dim3(256), 0, s, n_fine, map,
                       coarse_part, fine_part);
}

}  // namespace kahip
