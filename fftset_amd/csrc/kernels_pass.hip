// kernels_pass.hip -- per-pass baseline: the reference's pass structure on
// the GPU, one launch per radix pass with an HBM round trip between passes.
//
// The reference runs a first (modulating) pass, its Stockham or in-place
// DIF/DIT passes and a last pass, each a full sweep over the transform in
// memory (src/fftset_vec.c:1161-1249). This baseline launches the same
// passes of this library's chains one at a time: workgroup j owns transform
// j of a class, and the data between passes lives in a global work tile
// instead of LDS. It exists to measure what fusing the passes into one
// LDS-resident kernel buys; results equal the fused kernels'.
#include "drivers.h"
#include "kernels.h"

namespace fsa {

// Transform data in HBM, natural addressing.
struct TileGlobal {
    float2* base;
    __device__ __forceinline__ float2& at(int e) const { return base[e]; }
};

constexpr int kPassIdle = 1 << 28;

template <int MOD, int OP>
struct PassIo {
    static __device__ __forceinline__ auto ld(const RClass& k, int j)
    {
        const float* in = k.in[j];
        if constexpr (OP == OP_INV) {
            if constexpr (MOD == MOD_FO)
                return LdFoBins{(const float2*)in, k.L};
            else
                return LdCx<true>{(const float2*)in};
        } else {
            if constexpr (MOD == MOD_FO)
                return LdFoMod{in, k.L, k.tmod};
            else
                return LdCx<false>{(const float2*)in};
        }
    }
    static __device__ __forceinline__ auto st(const RClass& k, int j)
    {
        float* out = k.out[j];
        if constexpr (OP == OP_FWD) {
            if constexpr (MOD == MOD_FO)
                return StFoBins{(float2*)out, k.L};
            else
                return StCx<false>{(float2*)out};
        } else {
            if constexpr (MOD == MOD_FO)
                return StFoDemod{out, k.L, k.tmod};
            else
                return StCx<true>{(float2*)out};
        }
    }
};

// Step `step` of transform blockIdx.x of class k:
//   forward / inverse: 0 first pass, 1..np-2 Stockham passes, np-1 last pass
//   get_kernel:        0 first pass, 1..np-2 DIF passes, np-1 bottom (writes K)
//   conv / grad_input: 0 first pass, 1..np-2 DIF passes, np-1 bottom (product),
//                      np..2np-3 DIT passes np-2..1, 2np-2 last pass
template <int MOD, int OP>
__device__ __forceinline__ void pass_step(const RClass& k, int step, float2* work)
{
    constexpr int E = 16;
    const int j = blockIdx.x;
    const int T = k.T;
    const int lt = (int)threadIdx.x < T ? (int)threadIdx.x : kPassIdle;
    const TileGlobal tl{work + (long long)j * k.L};
    const ChainD& c = k.c;
    const int np = c.np;
    using Io = PassIo<MOD, OP>;
    if (np == 1) {
        if constexpr (OP == OP_KERN) {
            FSA_RADIX_SWITCH(c.p[0].R, E, (single_kern<RR>(lt, Io::ld(k, j), (float2*)k.out[j])));
        } else if constexpr (OP == OP_CONV || OP == OP_GRAD_IN) {
            const BotConv bot{(const float2*)k.kern[j], OP == OP_GRAD_IN};
            FSA_RADIX_SWITCH(c.p[0].R, E, (single_conv<RR>(lt, Io::ld(k, j), Io::st(k, j), bot)));
        } else {
            FSA_RADIX_SWITCH(c.p[0].R, E, (single_pass<RR>(lt, Io::ld(k, j), Io::st(k, j))));
        }
        return;
    }
    if (step == 0) {
        FSA_RADIX_SWITCH(c.p[0].R, E, (pass_first<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, Io::ld(k, j))));
        return;
    }
    if constexpr (OP == OP_FWD || OP == OP_INV) {
        const PassD& p = c.p[step];
        if (step < np - 1) {
            FSA_RADIX_SWITCH(p.R, E, (stock_mid<RR, NB>(tl, lt, T, p.m, p.s, p.tw)));
        } else {
            FSA_RADIX_SWITCH(p.R, E, (stock_last<RR, NB>(tl, lt, T, p.s, Io::st(k, j))));
        }
    } else {
        if (step < np - 1) {
            const PassD& p = c.p[step];
            FSA_RADIX_SWITCH(p.R, E, (dif_mid<RR, NB>(tl, lt, T, p.m, p.s, p.tw)));
        } else if (step == np - 1) {
            const PassD& p = c.p[np - 1];
            if constexpr (OP == OP_KERN) {
                const BotKernOut bot{(float2*)k.out[j]};
                FSA_RADIX_SWITCH(p.R, E, (dif_bottom<RR, NB>(tl, lt, T, p.s, bot)));
            } else {
                const BotConv bot{(const float2*)k.kern[j], OP == OP_GRAD_IN};
                FSA_RADIX_SWITCH(p.R, E, (dif_bottom<RR, NB>(tl, lt, T, p.s, bot)));
            }
        } else if constexpr (OP == OP_CONV || OP == OP_GRAD_IN) {
            const int i = 2 * np - 2 - step;  // DIT pass index: np-2 .. 0
            const PassD& p = c.p[i];
            if (i > 0) {
                FSA_RADIX_SWITCH(p.R, E, (dit_mid<RR, NB>(tl, lt, T, p.m, p.s, p.tw)));
            } else {
                FSA_RADIX_SWITCH(p.R, E, (dit_last<RR, NB>(tl, lt, T, p.m, p.tw, Io::st(k, j))));
            }
        }
    }
}

__global__ void __launch_bounds__(1024) k_pass(const RClass* __restrict__ cls, int step, float2* work)
{
    const RClass& k = *cls;
    switch (k.mod * OP_COUNT + k.op) {
    case MOD_FO * OP_COUNT + OP_FWD: pass_step<MOD_FO, OP_FWD>(k, step, work); break;
    case MOD_FO * OP_COUNT + OP_INV: pass_step<MOD_FO, OP_INV>(k, step, work); break;
    case MOD_FO * OP_COUNT + OP_KERN: pass_step<MOD_FO, OP_KERN>(k, step, work); break;
    case MOD_FO * OP_COUNT + OP_CONV: pass_step<MOD_FO, OP_CONV>(k, step, work); break;
    case MOD_FO * OP_COUNT + OP_GRAD_IN: pass_step<MOD_FO, OP_GRAD_IN>(k, step, work); break;
    case MOD_CX * OP_COUNT + OP_FWD: pass_step<MOD_CX, OP_FWD>(k, step, work); break;
    case MOD_CX * OP_COUNT + OP_INV: pass_step<MOD_CX, OP_INV>(k, step, work); break;
    case MOD_CX * OP_COUNT + OP_KERN: pass_step<MOD_CX, OP_KERN>(k, step, work); break;
    case MOD_CX * OP_COUNT + OP_CONV: pass_step<MOD_CX, OP_CONV>(k, step, work); break;
    case MOD_CX * OP_COUNT + OP_GRAD_IN: pass_step<MOD_CX, OP_GRAD_IN>(k, step, work); break;
    default: break;
    }
}

int pass_steps(int op, int np)
{
    if (np <= 1)
        return 1;
    return (op == OP_CONV || op == OP_GRAD_IN) ? 2 * np - 1 : np;
}

hipError_t launch_pass(const RClass* d_cls, int step, int ntrans, int T, float2* work, hipStream_t st)
{
    if (ntrans <= 0)
        return hipSuccess;
    const int threads = ((T + 63) / 64) * 64;
    if (threads > 1024)
        return hipErrorInvalidValue;
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(k_pass, dim3(ntrans), dim3(threads), 0, st, d_cls, step, work);
    return hipGetLastError();
}

}  // namespace fsa
