// Synthetic graph generators that build the CSR directly in GPU memory.
#pragma once

#include "contract.hpp"

namespace kahip {

void generate_rmat(int scale, long long samples, unsigned long long seed, double a, double b,
                   double c, OwnedGraph& out);

}  // namespace kahip
