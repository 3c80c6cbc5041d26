// kernels.h -- kernel argument blocks and host-side launchers.
#pragma once

#include "engine.h"
#include "plan_desc.h"

namespace fsa {

// One launch of the single-kernel LDS engine over a uniform batch.
struct LdsArgs {
    ChainD c;
    const float2* tw;  // shared table of M-th roots of unity
    int ms;            // FREQOFFSETREAL modulation stride: W_{4L}^e = tw[e * ms]
    int L;             // complex bins (== c.n)
    long long ntrans;  // transforms in the batch
    int G, T;          // transforms per workgroup, lanes per transform
    const float* in;
    float* out;
    const float* kern;
    long long in_st, out_st, kern_st;  // floats between consecutive transforms
};

hipError_t launch_lds(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st);
hipError_t launch_twiddles(float2* tw, int M, hipStream_t st);

}  // namespace fsa
