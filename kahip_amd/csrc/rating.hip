// Edge ratings for the LP clustering on gfx950.
//
// Plain: expansion*^2 = w(e)^2 / (c(u) c(v)) (edge_ratings.cpp:171-183), quantised linearly:
//   score = clamp(round(1024 * w^2 / (c(u) c(v))), 1, kScoreMax).
// Learned: features -> MLP -> logit -> score = clamp(round(128 * logit) + 2048, 1, kScoreMax).
// Features of a directed edge e = (u, v), symmetric in u and v, from integers only:
//   f0 log2(1 + min deg)   f1 log2(1 + max deg)    f2 log2(1 + w(e))
//   f3 log2(1 + min c)     f4 log2(1 + max c)      f5 popcount(sig(u) & sig(v)) / 8
//   f6 number of 3 seeded one-round, one-sub-round LP sweeps that put u and v in the same cluster
//   f7 log2(1 + expansion*^2)
// sig(x) is a 64-bit sketch of x's neighbourhood (bit mix32(y) & 63 set for every neighbour y),
// so f5 is a capped, collision-prone common-neighbour count computed in O(1) per edge.
//
// MLP on MFMA (v_mfma_f32_16x16x32_bf16): a wave evaluates 16 edges at a time with the edges
// along N. The layer-1 output tile of a lane (4 hidden units of one edge) is exactly what the
// B operand of layer 2 needs once the K order of each 32-wide K block is permuted (the packed
// W2 uses the same permutation), so activations never leave registers. Layer 3 (64 -> 1) is a
// per-lane dot product plus two cross-lane adds.

#include <cmath>
#include <cstring>
#include <vector>

#include "rating.hpp"

namespace kahip {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

inline int grid_for(long long work, int block) { return (int)((work + block - 1) / block); }

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__device__ inline float round_bf(float f) { return bf2f(f2bf(f)); }

// ------------------------------------------------------------------ features
// Sketch of a node with at most kSketchLight neighbours: one thread. Heavier nodes are appended
// to heavy[1..] (heavy[0] counts them, zero on entry; capacity m2 / (kSketchLight + 1)) and
// sketched by a whole workgroup each, so a hub's row is read by 256 lanes, not by one.
constexpr int kSketchLight = 64;
constexpr int kSketchHeavyGrid = 4096;

__device__ inline unsigned long long sketch_bit(int y) {
    return 1ull << (mix32((uint32_t)y) & 63u);
}

__global__ void sketch_light_kernel(int n, const int* __restrict__ xadj,
                                    const int* __restrict__ adjncy,
                                    unsigned long long* __restrict__ sig, int* __restrict__ heavy) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int b = xadj[u], end = xadj[u + 1];
    if (end - b > kSketchLight) {
        heavy[1 + atomicAdd(heavy, 1)] = u;
        return;
    }
    unsigned long long s = 0;
    for (int e = b; e < end; ++e) s |= sketch_bit(adjncy[e]);
    sig[u] = s;
}

__global__ __launch_bounds__(256) void sketch_heavy_kernel(const int* __restrict__ xadj,
                                                           const int* __restrict__ adjncy,
                                                           const int* __restrict__ heavy,
                                                           unsigned long long* __restrict__ sig) {
    __shared__ unsigned long long part[4];
    const int count = heavy[0];
    for (int i = blockIdx.x; i < count; i += gridDim.x) {
        const int u = heavy[1 + i];
        const int end = xadj[u + 1];
        unsigned long long s = 0;
#pragma unroll 4
        for (int e = xadj[u] + (int)threadIdx.x; e < end; e += 256) s |= sketch_bit(adjncy[e]);
        uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo |= __shfl_xor(lo, off);
            hi |= __shfl_xor(hi, off);
        }
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = ((unsigned long long)hi << 32) | lo;
        __syncthreads();
        if (threadIdx.x == 0) sig[u] = part[0] | part[1] | part[2] | part[3];
        __syncthreads();
    }
}

// Everything the features of an edge need from one endpoint, in one 32-byte record, so that the
// neighbour side of an edge costs one random sector instead of five.
struct __align__(32) NodeRec {
    int deg, c;
    unsigned long long sig;
    int l0, l1, l2, pad;
};

__global__ void node_rec_kernel(int n, const int* __restrict__ xadj, const int* __restrict__ vwgt,
                                const unsigned long long* __restrict__ sig,
                                const int* __restrict__ l0, const int* __restrict__ l1,
                                const int* __restrict__ l2, NodeRec* __restrict__ rec) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    rec[u] = NodeRec{xadj[u + 1] - xadj[u], vwgt[u], sig[u], l0[u], l1[u], l2[u], 0};
}

__device__ inline float lg(float x) { return __log2f(1.0f + x); }

// OUT = float: 8 floats per edge (the features API); OUT = bf16x8: the 8 features rounded to
// BF16 (round to nearest even, as the MLP rounds its inputs), 16 B per edge.
template <class OUT>
__global__ __launch_bounds__(256) void features_kernel(
    int n, long long m2, const int* __restrict__ xadj, const int* __restrict__ adjncy,
    const int* __restrict__ adjwgt, const NodeRec* __restrict__ rec, OUT* __restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long e0 = t * 8;
    if (e0 >= m2) return;
    int lo = 0, hi = n;  // source of e0
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((long long)xadj[mid] <= e0) lo = mid; else hi = mid;
    }
    int u = lo;
    long long next = xadj[u + 1];
    NodeRec ru = rec[u];
    const long long e1 = e0 + 8 < m2 ? e0 + 8 : m2;
    for (long long e = e0; e < e1; ++e) {
        if (e >= next) {
            while (e >= next) {
                ++u;
                next = xadj[u + 1];
            }
            ru = rec[u];
        }
        const NodeRec rv = rec[adjncy[e]];
        const float du = (float)ru.deg, dv = (float)rv.deg;
        const float w = (float)adjwgt[e];
        const float cu = (float)ru.c, cv = (float)rv.c;
        const float cn = (float)__popcll(ru.sig & rv.sig);
        const float agree = (float)((ru.l0 == rv.l0) + (ru.l1 == rv.l1) + (ru.l2 == rv.l2));
        const float es2 = (w * w) / fmaxf(cu * cv, 1.0f);
        const float f[kFeatures] = {lg(fminf(du, dv)), lg(fmaxf(du, dv)), lg(w),
                                    lg(fminf(cu, cv)), lg(fmaxf(cu, cv)), cn * 0.125f, agree,
                                    lg(es2)};
        if constexpr (sizeof(OUT) == sizeof(float)) {
#pragma unroll
            for (int j = 0; j < kFeatures; ++j) out[e * kFeatures + j] = f[j];
        } else {
            bf16x8 b;
#pragma unroll
            for (int j = 0; j < kFeatures; ++j) b[j] = (__bf16)f[j];
            out[e] = b;
        }
    }
}

__global__ void expansion_star2_kernel(int n, long long m2, const int* __restrict__ xadj,
                                       const int* __restrict__ adjncy,
                                       const int* __restrict__ adjwgt,
                                       const int* __restrict__ vwgt, int shift,
                                       int* __restrict__ score) {
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
    const long long e1 = e0 + 8 < m2 ? e0 + 8 : m2;
    for (long long e = e0; e < e1; ++e) {
        while (e >= next) {
            ++u;
            next = xadj[u + 1];
        }
        const long long w = adjwgt[e];
        const long long c = (long long)vwgt[u] * (long long)vwgt[adjncy[e]];
        // round(1024 w^2 / c) in integers
        const long long num = 1024 * w * w;
        long long q = c > 0 ? (num + c / 2) / c : (long long)kScoreMax;
        if (q < 1) q = 1;
        if (q > kScoreMax) q = kScoreMax;
        score[e] = max((int)q >> shift, 1);
    }
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

__global__ void max_degree_kernel(int n, const int* __restrict__ xadj, int* __restrict__ out) {
    __shared__ int best;
    if (threadIdx.x == 0) best = 0;
    __syncthreads();
    int d = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        d = max(d, xadj[v + 1] - xadj[v]);
    atomicMax(&best, d);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, best);
}

// ------------------------------------------------------------------ MLP forward on MFMA
// Packed weights (bf16 bits), one 8-element vector per lane:
//   A1[t][lane]        t = 0..3 (hidden rows 16t..16t+15): W1[16t + lane%16][8g + j], g = lane/16,
//                      zero for 8g + j >= 8
//   A2[t2][kb][lane]   W2[16 t2 + lane%16][hid(kb, g, j)],
//                      hid(kb, g, j) = 16 (2 kb + (j >= 4)) + 4 g + (j & 3)
// followed by b1[64], b2[64], w3[64], b3 as fp32 (stored in the float tail of the same buffer).
constexpr int kPackA1 = 4 * 64 * 8;
constexpr int kPackA2 = 4 * 2 * 64 * 8;
constexpr int kPackBf16 = kPackA1 + kPackA2;          // bf16 elements
constexpr int kPackFloats = 64 + 64 + 64 + 1;        // fp32 tail
constexpr int kPackBytes = kPackBf16 * 2 + kPackFloats * 4;

__host__ __device__ inline int hid_of(int kb, int g, int j) {
    return 16 * (2 * kb + (j >= 4 ? 1 : 0)) + 4 * g + (j & 3);
}

// logit != nullptr: writes the logits; otherwise writes score = quantize(logit) >> shift, >= 1.
template <class IN>
__global__ __launch_bounds__(256) void mlp_forward_kernel(const IN* __restrict__ feat,
                                                          long long m,
                                                          const uint16_t* __restrict__ pack,
                                                          float* __restrict__ logit,
                                                          int* __restrict__ score, int shift) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, col = lane & 15;
    const bf16x8* A = reinterpret_cast<const bf16x8*>(pack);
    bf16x8 a1[4], a2[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) a1[t] = A[t * 64 + lane];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) a2[t][kb] = A[kPackA1 / 8 + (t * 2 + kb) * 64 + lane];
    const float* tail = reinterpret_cast<const float*>(pack + kPackBf16);
    float b1[4][4], b2[4][4], w3[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            b1[t][r] = tail[16 * t + 4 * g + r];
            b2[t][r] = tail[64 + 16 * t + 4 * g + r];
            w3[t][r] = tail[128 + 16 * t + 4 * g + r];
        }
    const float b3 = tail[192];
    const long long waves = (long long)gridDim.x * (blockDim.x >> 6);
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (long long base = wave * 16; base < m; base += waves * 16) {
        const long long e = base + col;
        // B operand of layer 1: the 8 features of edge `col` in lanes of group 0
        bf16x8 xb = {};
        if (g == 0 && e < m) {
            if constexpr (sizeof(IN) == sizeof(float)) {
#pragma unroll
                for (int j = 0; j < 8; ++j) xb[j] = (__bf16)feat[e * kFeatures + j];
            } else {
                xb = feat[e];
            }
        }
        f32x4 h1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
            z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[t], xb, z, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) h1[t][r] = fmaxf(z[r] + b1[t][r], 0.0f);
        }
        f32x4 h2[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) h2[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            bf16x8 hb;
#pragma unroll
            for (int j = 0; j < 8; ++j) hb[j] = (__bf16)h1[2 * kb + (j >> 2)][j & 3];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                h2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[t][kb], hb, h2[t], 0, 0, 0);
        }
        float acc = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc += w3[t][r] * fmaxf(h2[t][r] + b2[t][r], 0.0f);
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        if (g == 0 && e < m) {
            if (logit) logit[e] = acc + b3;
            else score[e] = max(quantize_logit(acc + b3) >> shift, 1);
        }
    }
}

// ------------------------------------------------------------------ training on MFMA
// One workgroup of four waves per kTB = 64 examples; wave w owns examples 16w..16w+15 and runs
// the forward exactly as mlp_forward_kernel (A operands built from the fp32 parameters, BF16
// inputs, weights and layer-1 activations, fp32 accumulation). The backward runs its three
// matrix products on v_mfma_f32_16x16x32_bf16 as well, with BF16 operands:
//   dH1 = W2^T D2      per wave: its 16 edges, K = the 64 hidden units of layer 2; the B operand
//                      is the wave's own D2 registers (same K permutation as layer 2's forward)
//   dW2 = D2^T H1      wave w: hidden rows 16w..16w+15, K = the workgroup's 64 edges (LDS)
//   dW1 = D1^T X       wave w: rows 16w..16w+15, 8 of the 16 columns used, K = 64 edges (LDS)
// where D2 = dL/dZ2 and D1 = dL/dZ1 are rounded to BF16 as product operands; the bias and w3
// gradients and the loss are fp32 sums over the edges (lane shuffles, then the four waves in
// order). Per-workgroup gradients go to `partial` and are summed in a fixed order, so training
// is deterministic.
constexpr int kTB = 64;  // examples per workgroup

__device__ inline float sum16(float x) {  // sum over the 16 lanes of a lane group (fixed tree)
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__global__ __launch_bounds__(256) void train_kernel(const float* __restrict__ feat,
                                                    const float* __restrict__ label, int batch,
                                                    const float* __restrict__ params,
                                                    float* __restrict__ partial,
                                                    float* __restrict__ loss_out) {
    __shared__ __bf16 Xs[kTB][16];          // features (columns 8..15 zero)
    __shared__ __bf16 H1s[kTB][kHidden];    // layer-1 activations
    __shared__ __bf16 D2s[kTB][kHidden];    // dL/dZ2
    __shared__ __bf16 D1s[kTB][kHidden];    // dL/dZ1
    __shared__ float red[4][3 * kHidden + 2];  // per-wave sums: dW3, db2, db1, db3, loss
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, col = lane & 15;
    const int le = wave * 16 + col;  // the lane's example within the workgroup
    const long long e = (long long)blockIdx.x * kTB + le;
    const bool valid = e < batch;
    // A operands of the forward (as the packed weights of mlp_forward_kernel)
    bf16x8 a1[4], a2[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            a1[t][j] = (__bf16)(k < kFeatures ? params[kOffW1 + (16 * t + col) * kFeatures + k] : 0.f);
        }
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                a2[t][kb][j] = (__bf16)params[kOffW2 + (16 * t + col) * kHidden + hid_of(kb, g, j)];
    }
    // forward, layer 1: B = the 8 features of edge `col` in lane group 0
    bf16x8 xb = {};
    if (g == 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) xb[j] = (__bf16)(valid ? feat[e * kFeatures + j] : 0.f);
#pragma unroll
        for (int j = 0; j < 16; ++j) Xs[le][j] = j < 8 ? xb[j] : (__bf16)0.f;
    }
    f32x4 z1[4];  // Z1[16t + 4g + r][col]
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[t], xb, z, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            z1[t][r] = z[r] + params[kOffB1 + 16 * t + 4 * g + r];
            H1s[le][16 * t + 4 * g + r] = (__bf16)fmaxf(z1[t][r], 0.f);
        }
    }
    // layer 2 (B operand: the lane's own layer-1 activations in the K order of hid_of)
    f32x4 z2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) z2[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        bf16x8 hb;
#pragma unroll
        for (int j = 0; j < 8; ++j) hb[j] = (__bf16)fmaxf(z1[2 * kb + (j >> 2)][j & 3], 0.f);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            z2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[t][kb], hb, z2[t], 0, 0, 0);
    }
    float w3[4][4];
    float acc = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int h = 16 * t + 4 * g + r;
            z2[t][r] += params[kOffB2 + h];
            w3[t][r] = params[kOffW3 + h];
            acc += w3[t][r] * fmaxf(z2[t][r], 0.f);
        }
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    const float z3 = acc + params[kOffB3];
    // loss and dL/dz3 of the lane's edge (every lane group holds the same values)
    float d = 0.f, l = 0.f;
    if (valid) {
        const float y = label[e];
        l = fmaxf(z3, 0.f) - z3 * y + log1pf(expf(-fabsf(z3)));  // BCE with logits, stable form
        d = (1.f / (1.f + expf(-z3)) - y) / (float)batch;
    }
    // D2 = d * w3 * (Z2 > 0); fp32 sums over the wave's edges for dW3 and db2
    f32x4 D2[4];
    float sw3[4][4], sb2[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            D2[t][r] = z2[t][r] > 0.f ? d * w3[t][r] : 0.f;
            D2s[le][16 * t + 4 * g + r] = (__bf16)D2[t][r];
            sw3[t][r] = sum16(d * fmaxf(z2[t][r], 0.f));
            sb2[t][r] = sum16(D2[t][r]);
        }
    // dH1 = W2^T D2 (M = 64 inputs of layer 2 in 4 tiles, N = the wave's 16 edges, K = 64)
    f32x4 D1[4];
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) D1[tk] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        bf16x8 db;
#pragma unroll
        for (int j = 0; j < 8; ++j) db[j] = (__bf16)D2[2 * kb + (j >> 2)][j & 3];
#pragma unroll
        for (int tk = 0; tk < 4; ++tk) {
            bf16x8 aw;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                aw[j] = (__bf16)params[kOffW2 + hid_of(kb, g, j) * kHidden + 16 * tk + col];
            D1[tk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw, db, D1[tk], 0, 0, 0);
        }
    }
    float sb1[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            D1[t][r] = z1[t][r] > 0.f ? D1[t][r] : 0.f;
            D1s[le][16 * t + 4 * g + r] = (__bf16)D1[t][r];
            sb1[t][r] = sum16(D1[t][r]);
        }
    const float sd = sum16(d), sl = sum16(l);
    if (col == 0) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int h = 16 * t + 4 * g + r;
                red[wave][h] = sw3[t][r];
                red[wave][kHidden + h] = sb2[t][r];
                red[wave][2 * kHidden + h] = sb1[t][r];
            }
        if (g == 0) {
            red[wave][3 * kHidden] = sd;
            red[wave][3 * kHidden + 1] = sl;
        }
    }
    __syncthreads();
    float* out = partial + (long long)blockIdx.x * kMlpParams;
    // dW2 rows 16 wave .. + 15 (A = D2^T from LDS, B = H1 from LDS, K = 64 edges)
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            bf16x8 a, b;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int ee = 32 * s2 + 8 * g + j;
                a[j] = D2s[ee][16 * wave + col];
                b[j] = H1s[ee][16 * tk + col];
            }
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
            out[kOffW2 + (16 * wave + 4 * g + r) * kHidden + 16 * tk + col] = c[r];
    }
    // dW1 rows 16 wave .. + 15 (B = X from LDS, columns 8..15 zero)
    {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            bf16x8 a, b;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int ee = 32 * s2 + 8 * g + j;
                a[j] = D1s[ee][16 * wave + col];
                b[j] = Xs[ee][col];
            }
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
        }
        if (col < kFeatures)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[kOffW1 + (16 * wave + 4 * g + r) * kFeatures + col] = c[r];
    }
    // fp32 gradients of w3, b2, b1, b3 and the loss: the four waves' sums in order
    if (threadIdx.x < kHidden) {
        const int h = threadIdx.x;
        out[kOffW3 + h] = ((red[0][h] + red[1][h]) + red[2][h]) + red[3][h];
        out[kOffB2 + h] = ((red[0][kHidden + h] + red[1][kHidden + h]) + red[2][kHidden + h]) +
                          red[3][kHidden + h];
        out[kOffB1 + h] = ((red[0][2 * kHidden + h] + red[1][2 * kHidden + h]) +
                           red[2][2 * kHidden + h]) + red[3][2 * kHidden + h];
    } else if (threadIdx.x == kHidden) {
        out[kOffB3] = ((red[0][3 * kHidden] + red[1][3 * kHidden]) + red[2][3 * kHidden]) +
                      red[3][3 * kHidden];
        loss_out[blockIdx.x] = ((red[0][3 * kHidden + 1] + red[1][3 * kHidden + 1]) +
                                red[2][3 * kHidden + 1]) + red[3][3 * kHidden + 1];
    }
}

__global__ void reduce_grads(int blocks, int batch, const float* __restrict__ partial,
                             const float* __restrict__ block_loss, float* __restrict__ grads,
                             float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float s = 0.f;
        for (int b = 0; b < blocks; ++b) s += partial[(long long)b * kMlpParams + p];
        grads[p] = s;
    } else if (p == kMlpParams && loss) {
        float s = 0.f;
        for (int b = 0; b < blocks; ++b) s += block_loss[b];
        *loss = s / (float)batch;
    }
}

__global__ void adam_kernel(float* __restrict__ params, float* __restrict__ m,
                            float* __restrict__ v, const float* __restrict__ g, float lr,
                            float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m[p] = b1 * m[p] + (1.f - b1) * g[p];
    v[p] = b2 * v[p] + (1.f - b2) * g[p] * g[p];
    params[p] -= lr * (m[p] / c1) / (sqrtf(v[p] / c2) + eps);
}

void pack_weights(const MlpWeights& w, std::vector<uint8_t>& out) {
    out.assign(kPackBytes, 0);
    uint16_t* bf = reinterpret_cast<uint16_t*>(out.data());
    for (int t = 0; t < 4; ++t)
        for (int lane = 0; lane < 64; ++lane) {
            const int g = lane >> 4, row = lane & 15;
            for (int j = 0; j < 8; ++j) {
                const int k = 8 * g + j;
                const float x = k < kFeatures ? w.p[kOffW1 + (16 * t + row) * kFeatures + k] : 0.f;
                bf[(t * 64 + lane) * 8 + j] = f2bf(x);
            }
        }
    for (int t = 0; t < 4; ++t)
        for (int kb = 0; kb < 2; ++kb)
            for (int lane = 0; lane < 64; ++lane) {
                const int g = lane >> 4, row = lane & 15;
                for (int j = 0; j < 8; ++j) {
                    const float x = w.p[kOffW2 + (16 * t + row) * kHidden + hid_of(kb, g, j)];
                    bf[kPackA1 + ((t * 2 + kb) * 64 + lane) * 8 + j] = f2bf(x);
                }
            }
    float* tail = reinterpret_cast<float*>(bf + kPackBf16);
    for (int h = 0; h < kHidden; ++h) {
        tail[h] = w.p[kOffB1 + h];
        tail[64 + h] = w.p[kOffB2 + h];
        tail[128 + h] = w.p[kOffW3 + h];
    }
    tail[192] = w.p[kOffB3];
}

}  // namespace

void RatingWorkspace::sweeps_and_sketch(const DGraph& g, int bound, uint32_t seed, hipStream_t s,
                                        LpWorkspace& lp, Profiler* prof) {
    sig_.reserve(g.n);
    lab_.reserve(3LL * g.n);
    size_.reserve(g.n);
    rec_.reserve((long long)g.n * (sizeof(NodeRec) / sizeof(int)));
    if (g.n == 0) return;
    {
        ProfScope ps(prof, "rating_sketch", 0.0, g.m2, s);
        heavy_.reserve(2 + g.m2 / (kSketchLight + 1));
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, heavy_.p);
        hipLaunchKernelGGL(sketch_light_kernel, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n,
                           g.xadj, g.adjncy, sig_.p, heavy_.p);
        hipLaunchKernelGGL(sketch_heavy_kernel, dim3(kSketchHeavyGrid), dim3(256), 0, s, g.xadj,
                           g.adjncy, heavy_.p, sig_.p);
    }
    {
        // three one-round sweeps with a single sub-round: the node lists do not depend on the
        // seed then, so they are built once; the seeds only change the tie-breaking
        ProfScope ps(prof, "rating_sweeps", 0.0, 3 * g.m2, s);
        BinConfig bins;
        lp.prepare(LpMode::Cluster, g, g.n, 1, 0u, bins, s);
        for (int i = 0; i < 3; ++i) {
            int* lab = lab_.p + (long long)i * g.n;
            const uint32_t sd = hash3(seed, 0xA6u, (uint32_t)i);
            LpWorkspace::init_cluster_labels(g, lab, size_.p, s);
            lp.run(LpMode::Cluster, g, lab, size_.p, bound, 1, sd, s, nullptr, "rating_sweep");
        }
        lp.check(s);
    }
    hipLaunchKernelGGL(node_rec_kernel, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n, g.xadj,
                       g.vwgt, sig_.p, lab_.p, lab_.p + g.n, lab_.p + 2LL * g.n,
                       reinterpret_cast<NodeRec*>(rec_.p));
    KH_CHECK(hipGetLastError());
}

void RatingWorkspace::features(const DGraph& g, int bound, uint32_t seed, float* out,
                               hipStream_t s, LpWorkspace& lp, Profiler* prof) {
    sweeps_and_sketch(g, bound, seed, s, lp, prof);
    ProfScope ps(prof, "rating_features", 0.0, g.m2, s);
    if (g.m2 > 0)
        hipLaunchKernelGGL(features_kernel<float>, dim3(grid_for((g.m2 + 7) / 8, 256)), dim3(256),
                           0, s, g.n, g.m2, g.xadj, g.adjncy, g.adjwgt,
                           reinterpret_cast<const NodeRec*>(rec_.p), out);
    KH_CHECK(hipGetLastError());
}

void RatingWorkspace::upload_weights(const MlpWeights& w, hipStream_t s) {
    std::vector<uint8_t> host;
    pack_weights(w, host);
    wpack_.reserve(kPackBytes / 2);
    KH_CHECK(hipMemcpyAsync(wpack_.p, host.data(), kPackBytes, hipMemcpyHostToDevice, s));
    KH_CHECK(hipStreamSynchronize(s));  // `host` must outlive the copy
}

void RatingWorkspace::mlp_forward(const float* feat, long long m, const MlpWeights& w,
                                  float* logits, hipStream_t s, Profiler* prof) {
    upload_weights(w, s);
    if (m > 0) {
        // 9344 useful FLOP per edge (2 x (8*64 + 64*64 + 64))
        ProfScope ps(prof, "mlp_forward", 9344.0 * (double)m, m, s);
        const long long groups = (m + 15) / 16;
        const int blocks = (int)std::min<long long>((groups + 3) / 4, 8192);
        hipLaunchKernelGGL(mlp_forward_kernel<float>, dim3(blocks), dim3(256), 0, s, feat, m,
                           wpack_.p, logits, nullptr, 0);
    }
    KH_CHECK(hipGetLastError());
}

int RatingWorkspace::score_shift(const DGraph& g, hipStream_t s) {
    map_.reserve(1);
    int* d = map_.p;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, d);
    hipLaunchKernelGGL(max_degree_kernel, dim3(std::min(grid_for(g.n, 256), 1024)), dim3(256), 0, s,
                       g.n, g.xadj, d);
    int maxdeg = 0;
    KH_CHECK(hipMemcpyAsync(&maxdeg, d, sizeof(int), hipMemcpyDeviceToHost, s));
    KH_CHECK(hipStreamSynchronize(s));
    int shift = 0;
    while (shift < 12 && (long long)maxdeg * (kScoreMax >> shift) >= (1LL << 31)) ++shift;
    KH_REQUIRE((long long)maxdeg * (kScoreMax >> shift) < (1LL << 31), "degree too large");
    return shift;
}

void RatingWorkspace::scores(RatingMode mode, const DGraph& g, int bound, uint32_t seed,
                             const MlpWeights* w, int* score, hipStream_t s, LpWorkspace& lp,
                             Profiler* prof) {
    if (g.m2 == 0) return;
    const int shift = score_shift(g, s);
    if (mode == RatingMode::ExpansionStar2) {
        ProfScope ps(prof, "rating_expansion_star2", 0.0, g.m2, s);
        hipLaunchKernelGGL(expansion_star2_kernel, dim3(grid_for((g.m2 + 7) / 8, 256)), dim3(256),
                           0, s, g.n, g.m2, g.xadj, g.adjncy, g.adjwgt, g.vwgt, shift, score);
        KH_CHECK(hipGetLastError());
        return;
    }
    KH_REQUIRE(mode == RatingMode::Learned && w, "learned rating needs MLP weights");
    // features rounded to BF16 (16 B per edge), then the MLP writes the quantised scores
    featb_.reserve(g.m2 * 8);
    upload_weights(*w, s);
    sweeps_and_sketch(g, bound, seed, s, lp, prof);
    {
        ProfScope ps(prof, "rating_features", 0.0, g.m2, s);
        hipLaunchKernelGGL(features_kernel<bf16x8>, dim3(grid_for((g.m2 + 7) / 8, 256)),
                           dim3(256), 0, s, g.n, g.m2, g.xadj, g.adjncy, g.adjwgt,
                           reinterpret_cast<const NodeRec*>(rec_.p),
                           reinterpret_cast<bf16x8*>(featb_.p));
    }
    {
        ProfScope ps(prof, "mlp_forward", 9344.0 * (double)g.m2, g.m2, s);
        const long long groups = (g.m2 + 15) / 16;
        const int blocks = (int)std::min<long long>((groups + 3) / 4, 8192);
        hipLaunchKernelGGL(mlp_forward_kernel<bf16x8>, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<const bf16x8*>(featb_.p), g.m2, wpack_.p, nullptr,
                           score, shift);
    }
    KH_CHECK(hipGetLastError());
}

void RatingWorkspace::train_step(const float* feat, const float* labels, int batch,
                                 float* params, float* m, float* v, float* grads, int step,
                                 float lr, float* loss, hipStream_t s) {
    KH_REQUIRE(batch > 0, "empty batch");
    const int blocks = (batch + kTB - 1) / kTB;
    partial_.reserve((long long)blocks * kMlpParams);
    loss_.reserve(blocks);
    hipLaunchKernelGGL(train_kernel, dim3(blocks), dim3(256), 0, s, feat, labels, batch, params,
                       partial_.p, loss_.p);
    hipLaunchKernelGGL(reduce_grads, dim3(grid_for(kMlpParams + 1, 256)), dim3(256), 0, s, blocks,
                       batch, partial_.p, loss_.p, grads, loss);
    if (lr > 0) {
        const float c1 = 1.f - std::pow(0.9f, (float)step);
        const float c2 = 1.f - std::pow(0.999f, (float)step);
        hipLaunchKernelGGL(adam_kernel, dim3(grid_for(kMlpParams, 256)), dim3(256), 0, s, params,
                           m, v, grads, lr, c1, c2);
    }
    KH_CHECK(hipGetLastError());
}

}  // namespace kahip
