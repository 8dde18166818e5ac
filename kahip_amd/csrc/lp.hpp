// Host interface of the synchronous LP kernels (clustering and refinement).
#pragma once

#include "common.hpp"
#include "devbuf.hpp"

namespace kahip {

constexpr int kMaxSub = 8;
constexpr int kMaxRounds = 64;

// Kernel arguments shared by the propose / filter / admit / apply launches of one sub-round.
struct LpArgs {
    const int* xadj;
    const int* adjncy;
    const int* adjwgt;
    const int* vwgt;
    int* label;        // label of each node (cluster id or block id)
    int* size;         // weight of each label
    int bound;         // admission bound: size + w < bound
    const int* lbound; // per-label bounds (a disjoint union of graphs), nullptr: `bound` for all
    uint32_t tie_seed;
    const uint8_t* active;  // refinement frontier (nullptr for clustering)
    uint8_t* next_active;
    int* inc;                  // per label: summed weight of proposals into it
    unsigned long long* top;   // per label: top proposal priority into it
    int* out;                  // per label: summed weight of departures (refinement)
    unsigned long long* dtop;  // per label: top departure priority (refinement)
    int* prop_t;               // per node: proposed label, -1 outside propose..apply
    int* prop_gain;
    uint8_t* prop_ok;
    int* moves;
    unsigned long long* trav;  // refinement: traversed directed edges, 64 spread counters
    // Refinement with incremental hub counters (nullptr otherwise): hub_slot[v] is the offset of
    // node v's per-block connection counters in hub_cnt, or -1; moves update the counters of
    // their hub neighbours.
    const int* hub_slot;
    int* hub_cnt;
    int* err;  // device error flag (kErrProbe), cleared by run(), read by check()
};

struct LpStats {
    int moves[kMaxRounds];
};

class LpWorkspace {
public:
    // Bins the nodes of g into per-(sub-round, degree bin) lists and sizes the scratch for
    // num_labels labels. One small host readback.
    // Nodes with include[v] == 0 (include may be null) are left out: they never propose.
    void prepare(LpMode mode, const DGraph& g, int num_labels, int subrounds, uint32_t seed,
                 const BinConfig& bins, hipStream_t s, const uint8_t* include = nullptr);
    // Runs `rounds` synchronous rounds. label/size are device arrays (node -> label,
    // label -> weight). Refinement starts from an all-active frontier.
    // lbound (optional): the bound of each label, replacing `bound`.
    void run(LpMode mode, const DGraph& g, int* label, int* size, int bound, int rounds,
             uint32_t seed, hipStream_t s, Profiler* prof, const char* phase,
             const int* lbound = nullptr);
    // label[v] = v, size[v] = vwgt[v]
    static void init_cluster_labels(const DGraph& g, int* label, int* size, hipStream_t s);
    // Moves per round of the last run (synchronises the stream).
    void read_moves(int rounds, int* out, hipStream_t s);
    // Directed edges read by the last refinement run (synchronises the stream): the rows scanned
    // by the proposal kernels (hubs deciding from their kept counters read none) and the rows of
    // moved nodes scanned to mark their neighbours and update hub counters.
    long long traversed(hipStream_t s);
    // Throws if a kernel of the last run() set the device error flag (synchronises the stream).
    void check(hipStream_t s);

    int host_counts[kMaxSub * kNumBins] = {};
    int host_wave_small_[kMaxSub] = {};  // clustering: wave-list nodes of the small-LDS kernel
    long long host_edges[kMaxSub * kNumBins] = {};
    int host_offsets[kMaxSub * kNumBins] = {};

private:
    void build_chunk_maps(hipStream_t s);  // cmap_ of every sub-round (after chunk_prefix_)
    int subrounds_ = 1;
    int num_labels_ = 0;
    BinConfig bins_;  // degree bins and tile shapes of the last prepare()
    bool dense_wave_ = false, dense_hub_ = false;
    // per sub-round hub bookkeeping (hub list = last degree bin)
    int hub_n_[kMaxSub] = {};
    int hub_chunks_[kMaxSub] = {};
    int hub_multi_n_[kMaxSub] = {};
    int hub_base_[kMaxSub] = {};    // offset into tab_off_/tab_cap_ (list order)
    int chunk_base_[kMaxSub] = {};  // offset into chunk_prefix_
    int cmap_base_[kMaxSub] = {};   // offset into cmap_ (chunks of the earlier sub-rounds)
    int tmap_base_[kMaxSub] = {};   // offset into tmap_ (tasks of the earlier sub-rounds)
    int mid_n_[kMaxSub] = {};       // hubs of the mid-size kernel per sub-round
    int mid_base_[kMaxSub] = {};    // offset into mmap_
    long long mid_edges_[kMaxSub] = {};  // edges of those hubs
    int multi_base_[kMaxSub] = {};  // offset into multi_
    int task_base_[kMaxSub] = {};   // offset into task_prefix_
    int hub_tasks_[kMaxSub] = {};   // (hub, bucket) tasks per sub-round
    int ct_sr_base_[kMaxSub] = {};  // bucket-major count block of each sub-round
    int ct_sr_size_[kMaxSub] = {};
    size_t scan_tmp_bytes_ = 0;
    DevBuf<int> hub_deg_, chunk_prefix_, multi_;
    DevBuf<long long> tab_off_, tab_cap_;
    DevBuf<int2> partials_;
    DevBuf<int> ct_base_, cnt_t_, pos_t_;
    DevBuf<char> scan_tmp_;
    DevBuf<long long> offm_base_;
    DevBuf<int> offm_, task_prefix_, bown_, big_tasks_, n_big_tasks_, mid_tasks_, n_mid_tasks_;
    DevBuf<int4> bres_;
    DevBuf<int4> cmap_, tmap_, mmap_;  // hub-chunk, (hub, bucket) task, mid-size hub descriptors
    DevBuf<int> midpos_;
    DevBuf<int2> tsub_;  // (hub, bucket) task: {multi-chunk hub index, bucket | buckets << 16}
    DevBuf<uint8_t> fb_;
    DevBuf<int> lists_, counts_, cursor_, offsets_;
    DevBuf<int> clists_, ccount_;  // refinement frontier per list (compacted each round)
    DevBuf<unsigned long long> edges_;
    DevBuf<int> prop_t_, prop_gain_, moves_;
    DevBuf<uint8_t> prop_ok_;
    DevBuf<int> inc_, out_;
    DevBuf<unsigned long long> trav_;
    DevBuf<unsigned long long> top_, dtop_;
    DevBuf<uint8_t> active_, next_active_;
    DevBuf<int> hub_key_, hub_val_;
    DevBuf<int> dense_off_, dense_cnt_;  // chunked dense refinement hubs: per-hub counters
    DevBuf<int> hub_slot_;
    DevBuf<int> err_;
    long long dense_total_ = 0;
    bool dense_all_ = false;
    long long hub_total_ = 0;
};

void fill_async(int* p, long long n, int value, hipStream_t s);
void fill_async(unsigned long long* p, long long n, unsigned long long value, hipStream_t s);
void fill_async(uint8_t* p, long long n, uint8_t value, hipStream_t s);
void fill_async(long long* p, long long n, long long value, hipStream_t s);

// W[b] = sum of vwgt over nodes in block b (k <= 65536).
void compute_block_weights(const DGraph& g, const int* part, int k, int* W, hipStream_t s);

}  // namespace kahip
