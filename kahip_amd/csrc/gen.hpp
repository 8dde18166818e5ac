// Synthetic graph generators that build the CSR directly in GPU memory.
#pragma once

#include "contract.hpp"

namespace kahip {

void generate_rmat(int scale, long long samples, unsigned long long seed, double a, double b,
                   double c, OwnedGraph& out);
// Graph i of a generated batch: kind 0 = RGG with `size` nodes, 1 = R-MAT of scale `size` (16
// samples per node); x = log2_min + u (log2_max - log2_min) with u = (d >> 11) 2^-53,
// d = SplitMix64(seed*K1 + 3*K2 + i); RGG size = floor(2^x + 0.5), R-MAT scale = floor(x + 0.5);
// kind = i mod 2; graph seed = SplitMix64(seed*K1 + 4*K2 + i).
void batch_graph_spec(unsigned long long seed, int i, double log2_min, double log2_max, int* kind,
                      int* size, unsigned long long* graph_seed);
// radius <= 0: the default r = 0.55 sqrt(ln n / n)
void generate_rgg(int n, unsigned long long seed, double radius, OwnedGraph& out);

}  // namespace kahip
