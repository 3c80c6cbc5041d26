// fs1.h -- four-step transforms with one workgroup per transform
// (kernels_fs1.hip): declarations shared with the host runtime.
#pragma once

#include "plan_desc.h"

namespace fsa {

// Four-step transforms of the 240 x 256 split (61440 bins) with one
// workgroup per transform and its intermediate in a per-workgroup scratch
// slot (kernels_fs1.hip).
enum Fs1Op : int { FS1_FWD = 0, FS1_INV = 1, FS1_CONV = 2 };
struct Fs1Args {
    ChainD colc, rowc;    // column (length A) and row (length B) chains
    const float2* tmod;   // FREQOFFSETREAL modulation W_{4n}^e
    const float2* tfs;    // tfs[k1 B + n2] = W_n^{n2 k1}
    const float2* tfs_fo; // FREQOFFSETREAL forward columns: W_{4n}^{n2 (4 k1 + 1)}
    const float2* trow;   // W_{4A}^{n1}
    long long ntrans;
    int conj_kern;
    const float* in;
    float* out;
    const float* kern;
    long long in_st, out_st, kern_st;
    const float* const* in_p;  // per-transform pointers (ragged batches), or null
    float* const* out_p;
    const float* const* kern_p;
    float2* scratch;  // one slot of n complex values per workgroup of the grid
    // chunking, set by launch_fs1: columns per column chunk, column chunks,
    // rows per row chunk, row chunks, padded LDS row length
    int cols, col_chunks, rows, row_chunks, row_ld;
};
bool fs1_matches(const ChainD& colc, const ChainD& rowc, int colT, int rowT);
long long fs1_grid(long long ntrans);
hipError_t launch_fs1(int mod, int opk, const Fs1Args& a, long long grid, hipStream_t st);

}  // namespace fsa
