// plan_desc.h -- plain descriptors shared by the host planner and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsa {

constexpr int kMaxPasses = 16;
constexpr int kMaxLdsBytes = 160 * 1024;  // gfx950 LDS per workgroup

enum Mod : int { MOD_FO = 0, MOD_CX = 1 };

// Operations of one kernel launch. GRAD_KERN is built from KERN + GRAD_IN.
enum Op : int { OP_FWD = 0, OP_INV = 1, OP_KERN = 2, OP_CONV = 3, OP_GRAD_IN = 4, OP_COUNT = 5 };

struct PassD {
    int R;  // radix of this pass
    int m;  // sub-length after the pass (n / (s R))
    int s;  // number of sub-transforms entering the pass
    int pad_;
    const float2* tw;  // tw[(k-1)*m + j] = W_{R m}^{jk}; null when m == 1
};

// One radix chain for a transform of length n.
struct ChainD {
    int n;
    int np;
    PassD p[kMaxPasses];
};

// Launch shape of the LDS engine for one chain.
struct ShapeD {
    int E;  // elements per thread budget (kernel template parameter)
    int T;  // threads per transform
    int G;  // transforms per workgroup
    int threads;  // workgroup size (>= G*T, multiple of 64)
    int lds_bytes;
};

}  // namespace fsa
