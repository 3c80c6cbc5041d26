// rcol.h -- four-step transforms n = A * 256 whose column step runs in
// registers (rcol_impl.h, instantiated in kernels_rcol.hip): declarations
// shared with the host runtime.
#pragma once

#include "plan_desc.h"

namespace fsa {

// Operations of the column-register kernel.
enum RcolOp : int { RC_FWD = 0, RC_INV = 1, RC_KERN = 2, RC_CONV = 3 };

struct RcolArgs {
    const float2* tw_row;  // rows, first pass of [16, 16]: W_256^{j k} at (k-1)*16 + j
    const float2* tfs;     // W_n^{n2 k1} at k1*256 + n2
    const float2* tfs_in;  // column-step output twiddles: tfs, or (FREQOFFSETREAL) W_{4n}^{n2 (4 k1 + 1)}
    const float2* tmod;    // FREQOFFSETREAL modulation W_{4n}^f
    long long ntrans;
    int conj_kern;         // convolution with conj(K) (signal adjoint)
    const float* in;
    float* out;
    const float* kern;     // convolution: kernel spectra (stride 0: one shared spectrum)
    float* spec;           // convolution: the signal's spectrum out (optional)
    long long in_st, out_st, kern_st, spec_st;  // floats
    const float* const* in_p;  // per-transform pointers (ragged batches), or null
    float* const* out_p;
    const float* const* kern_p;
};

// True for a split A x 256 with row chain [16, 16] and a column length A
// the kernel is compiled for (the column chain does not matter: the column
// step's output is in natural order whatever its chain).
bool rcol_matches(const ChainD& colc, const ChainD& rowc);
// Column lengths the kernel is compiled for.
bool rcol_length(unsigned A);
hipError_t launch_rcol(int A, int mod, int opk, const RcolArgs& a, hipStream_t st);

}  // namespace fsa
