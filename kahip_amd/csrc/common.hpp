// Shared definitions for the MI355X (gfx950) LP coarsening / refinement library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace kahip {

// ------------------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define KH_CHECK(expr)                                                                         \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            throw ::kahip::Error(std::string("HIP error ") + hipGetErrorString(_e) + " at " + \
                                 __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr);      \
    } while (0)

#define KH_REQUIRE(cond, msg)                                    \
    do {                                                         \
        if (!(cond)) throw ::kahip::Error(std::string(msg));     \
    } while (0)

// ------------------------------------------------------------------------------ hashing
// Wellons' "lowbias32" finaliser; the synchronous LP schedule draws every pseudo-random choice
// from hash3 so that any thread order gives the same result.
__host__ __device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
__host__ __device__ inline uint32_t hash3(uint32_t s, uint32_t a, uint32_t b) {
    return mix32(mix32(mix32(s ^ 0x85ebca6bU) ^ a) ^ b);
}

// Open-addressing tables of the kernels: keys >= 0, empty slot -1, cap a power of two, linear
// probing from mix32(key). probe_claim returns the slot holding key, claiming the first empty
// slot of its probe path when key is absent (fresh = true then), or -1 once all cap slots hold
// other keys. Every table is sized so that -1 cannot occur; callers record it in a device error
// flag that the host turns into an error, so no probe loop can run without end.
template <class I>
__device__ inline I probe_claim(int* K, I cap, int key, bool& fresh) {
    I h = (I)(mix32((uint32_t)key) & (uint32_t)(cap - 1));
    for (I step = 0; step < cap; ++step) {
        const int old = atomicCAS(&K[h], -1, key);
        if (old == -1 || old == key) {
            fresh = old == -1;
            return h;
        }
        h = (h + 1) & (cap - 1);
    }
    fresh = false;
    return (I)-1;
}
template <class I>
__device__ inline I probe_claim(int* K, I cap, int key) {
    bool fresh;
    return probe_claim(K, cap, key, fresh);
}

// Device error flags (one int per workspace, cleared before its kernels run, read by the host
// after them).
constexpr int kErrProbe = 1;  // a hash table probe found no slot (sizing invariant broken)

constexpr uint32_t kSubSalt = 0x68bc21ebU;
constexpr uint32_t kTieSalt = 0x2f6b7d93U;
constexpr uint32_t kGolden = 0x9e3779b9U;

__host__ __device__ inline int subround_of(uint32_t seed, int v, int S) {
    return (int)(hash3(seed ^ kSubSalt, 0u, (uint32_t)v) % (uint32_t)S);
}
__host__ __device__ inline uint32_t tie_seed_of(uint32_t seed, int round) {
    return seed ^ kTieSalt ^ ((uint32_t)round * kGolden);
}

// Proposal priority packed for a 64-bit atomicMax: higher gain first, then lower node id.
// gain lies in [-2^31, 2^31) (all weights are int32 and their totals stay below 2^31).
__host__ __device__ inline unsigned long long prio_key(int gain, int v) {
    return ((unsigned long long)((uint32_t)gain ^ 0x80000000u) << 32) |
           (unsigned long long)(0xFFFFFFFFu - (uint32_t)v);
}

// ------------------------------------------------------------------------------ graph
// Device CSR graph, int32 structure-of-arrays. adjwgt / vwgt are never null on the device (unit
// weights are materialised at upload).
struct DGraph {
    int n = 0;
    int64_t m2 = 0;  // directed entries (xadj[n])
    int* xadj = nullptr;
    int* adjncy = nullptr;
    int* adjwgt = nullptr;
    int* vwgt = nullptr;
    int64_t total_vwgt = 0;
    int max_degree = 0;
};

// LP mode shared by clustering (labels = clusters, bound B) and refinement (labels = blocks,
// bound U).
enum class LpMode : int { Cluster = 0, Refine = 1 };

// Degree bins of the LP and contraction kernels. Bin b holds nodes with degree in
// (bin_hi[b-1], bin_hi[b]]; the last bin is unbounded. Thresholds are tunable.
constexpr int kNumBins = 6;
struct BinConfig {
    int hi[kNumBins] = {8, 16, 32, 64, 512, 0x7fffffff};
    // Launch (tile) shapes of the LP kernels; the synchronous schedule makes the result
    // independent of them.
    int group_items = 2;   // nodes per lane group in flight in the 8..64-lane kernels: 1, 2 or 4
    int wave_small = 128;  // clustering: wave-bin nodes up to this degree use the small-table wave
                           // kernel (2 x wave_small LDS slots per wave): 64, 128 or 256
};

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace kahip
