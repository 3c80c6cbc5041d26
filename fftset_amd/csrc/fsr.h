// fsr.h -- 61440-bin transforms (240 x 256) with the whole transform
// resident on one CU (fsr_impl.h, instantiated in kernels_fsr_cx.hip and
// kernels_fsr_fo.hip): declarations shared with the host runtime.
#pragma once

#include "plan_desc.h"

namespace fsa {

// Operations of the resident kernel. FREQOFFSETREAL forward and inverse
// transforms keep the two-launch four-step path (their bin interleave makes
// strided output runs); every other operation of a 240 x 256 plan runs here.
enum FsrOp : int { FSR_FWD = 0, FSR_INV = 1, FSR_KERN = 2, FSR_CONV = 3 };

// Complex values of the per-transform spill region: batch rows 48..79 of
// the row phase pass through it while the column tile holds their LDS.
constexpr int kFsrSpill = 32 * 256;

struct FsrArgs {
    const float2* tw_col;    // column DIF, first pass of [16, 15]: W_240^{u k} at (k-1)*15 + u
    const float2* tw_close;  // closing columns, first pass of [15, 16]: W_240^{u k} at (k-1)*16 + u
    const float2* tw_row;    // rows, first pass of [16, 16]: W_256^{j k} at (k-1)*16 + j
    const float2* tfs;       // W_n^{n2 k1} at k1*256 + n2
    const float2* tfs_in;    // column-step output twiddles: tfs, or (FREQOFFSETREAL) W_{4n}^{n2 (4 k1 + 1)}
    const float2* trow;      // FREQOFFSETREAL column input factor W_{4A}^{n1}
    const float2* tmod;      // FREQOFFSETREAL modulation W_{4n}^f
    long long ntrans;
    int conj_kern;           // convolution with conj(K) (signal adjoint)
    const float* in;
    float* out;
    const float* kern;       // convolution: kernel spectra (stride 0: one shared spectrum)
    float* spec;             // convolution: the signal's spectrum out (optional)
    float2* spill;           // kFsrSpill complex values per transform
    long long in_st, out_st, kern_st, spec_st, spill_st;  // floats (spill: complex values)
    const float* const* in_p;  // per-transform pointers (ragged batches), or null
    float* const* out_p;
    const float* const* kern_p;
    float* const* spill_p;
};

// True for the 240 x 256 split with column chain [16, 15] and row chain
// [16, 16], the shape the resident kernel is compiled for.
bool fsr_matches(const ChainD& colc, const ChainD& rowc);
// True when (mod, op) runs the resident kernel on such a plan.
bool fsr_op(int mod, int op);
hipError_t launch_fsr(int mod, int opk, const FsrArgs& a, hipStream_t st);

}  // namespace fsa
