// planner.h -- radix chains and launch shapes (host side).
#pragma once

#include <vector>

#include "plan_desc.h"

namespace fsa {

// Factors n = 2^a 3^b 5^c; returns false for anything else (or 5^2 and up,
// which the reference cannot plan either: radix 5 is final-only,
// src/fftset_vec.c:944-971).
bool factor235(unsigned n, int* a, int* b, int* c);

// Default chain for the LDS engine: few passes of register-resident radices
// from {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2}, largest first.
std::vector<int> default_chain(unsigned n);

// Candidate chains for the autotuner: every multiset of register radices
// with product n and at most `max_passes` passes (at most one radix-2
// pass), each in largest-first and smallest-first order; the default chain
// comes first. At most `limit` chains.
std::vector<std::vector<int>> candidate_chains(unsigned n, int max_passes, int limit);

// Fills a ChainD for `radices` over length n (twiddle pointers left null).
bool build_chain(const std::vector<int>& radices, unsigned n, ChainD* out);

// Lanes per transform / transforms per workgroup for element budget E.
// `batch_hint` caps G for small batches. Returns false when n does not fit.
bool choose_shape(const ChainD& c, int E, long long batch_hint, ShapeD* out);

// Largest n the single-kernel LDS engine accepts.
bool fits_lds(unsigned n);

}  // namespace fsa
