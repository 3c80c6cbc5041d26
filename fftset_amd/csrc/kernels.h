// kernels.h -- kernel argument blocks and host-side launchers.
#pragma once

#include "engine.h"
#include "plan_desc.h"

namespace fsa {

// One launch of the single-kernel LDS engine over a uniform batch.
struct LdsArgs {
    ChainD c;
    const float2* tmod;  // FREQOFFSETREAL modulation: tmod[e] = W_{4L}^e, e < L
    int L;               // complex bins (== c.n)
    long long ntrans;  // transforms in the batch
    int G, T;          // transforms per workgroup, lanes per transform
    const float* in;
    float* out;
    const float* kern;
    long long in_st, out_st, kern_st;  // floats between consecutive transforms
};

hipError_t launch_lds(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st);

// Four-step (n = A*B) steps for transforms larger than LDS.
enum ColLoad : int { CL_FOMOD = 0, CL_FOBINS = 1, CL_CX = 2, CL_CXCONJ = 3 };
enum ColStore : int { CS_WORKTW = 0, CS_FODEMOD = 1, CS_CXCONJ = 2 };
enum RowKind : int { RK_STOCK = 0, RK_CONV = 1, RK_KERN = 2 };
enum RowStore : int { RS_FOBINS = 0, RS_CX = 1, RS_FODEMOD = 2, RS_CXCONJ = 3, RS_NONE = 4 };

struct FsArgs {
    ChainD c;            // column chain (length A) or row chain (length B)
    const float2* tmod;  // FREQOFFSETREAL modulation: tmod[e] = W_{4L}^e
    const float2* tfs;   // tfs[k1*B + n2] = W_L^{n2 k1}
    int L, A, B;         // n = L = A*B
    long long ntrans;
    int C;       // columns (or rows) per workgroup
    int T;       // lanes per column (or row) transform
    int groups;  // workgroups per transform
    int conj_kern;
    const float* in;
    float* out;
    const float* kern;
    long long in_st, out_st, kern_st;
};

hipError_t launch_cols(int load, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st);
hipError_t launch_rows(int kind, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st);
// out[r*cols + c] = W_N^{((a0 + r) * c) mod N}, computed in FP64.
hipError_t launch_twiddles(float2* out, long long N, int a0, int rows, int cols, hipStream_t st);

}  // namespace fsa
