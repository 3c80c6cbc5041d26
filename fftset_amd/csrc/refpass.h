// refpass.h -- the per-size-launch baseline with the reference's pass
// structure (host side; kernels in kernels_refpass.hip).
#pragma once

#include <vector>

#include "context.h"

namespace fsa {

// The reference's min-cost radix chain for an inner transform of `length`
// at CPU vector width vw (1, 4 or 8): build_float_graph over
// FFTSET_FLOAT_PASSES (src/fftset_vec.c:944-971, 1004-1062). Empty when the
// length has no chain.
std::vector<int> reference_chain(unsigned length, unsigned vw);

// The reference's pass structure for (modulation, L): its vector width vw,
// the radices of the first pass's split into vw sub-transforms ([4] at
// width 4, [4, 2] at width 8, none at width 1) and the inner chain of
// length L / vw (src/fftset_mod_freqoffsetreal.c:1012-1131,
// src/fftset_mod_cplx.c:320-347).
struct RefStructure {
    unsigned vw = 1;
    std::vector<int> split;
    std::vector<int> inner;
};
bool reference_structure(int mod, unsigned L, RefStructure* out);

// One batch class run with the reference's pass structure.
struct RefPassClass {
    int mod = 0, op = 0;
    int L = 0;
    long long ntrans = 0;
    int nsplit = 0;   // passes of the first-pass split
    unsigned V = 1;   // sub-transforms per transform (the reference's vector width)
    ChainD c{};       // split + inner chain over L, in-place DIF view (twiddles attached)
    const float2* tmod = nullptr;
    const float* const* in = nullptr;  // device pointer arrays of the class
    float* const* out = nullptr;
    float2* work = nullptr;   // ntrans * L
    float2* work2 = nullptr;  // ntrans * L (Stockham ping-pong)
    float2* kref = nullptr;   // ntrans * L (convolution: this structure's kernel spectra)
    bool own_kref = true;     // false: kref is another class's get_kernel output
    int slot[8] = {0}, resid[8] = {0};
};

// Plans and allocates one class (transforms `rc`'s pointer arrays, op) and,
// for convolutions, gathers the class's kernel spectra into the baseline's
// own order -- or, with kref_from, takes them from another class's
// get_kernel output (already in that order, one per transform). Returns an
// FFTSET_* code.
int refpass_setup(Ctx* ctx, const fftset_fft* p, int op, const RClass& rc, long long ntrans, RefPassClass* out,
                  hipStream_t st, float2* kref_from = nullptr);
// Queues every sweep of the class on `st`.
int refpass_run(const RefPassClass& k, hipStream_t st);
void refpass_free(RefPassClass* k);

}  // namespace fsa
