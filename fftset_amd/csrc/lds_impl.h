// lds_impl.h -- the single-kernel LDS-resident transform (n fits in LDS).
//
// One workgroup holds G transforms of one plan in LDS, T lanes each. The
// radix chain is a runtime descriptor; every pass dispatches on its radix to
// a fully unrolled register butterfly. Modulation, bin interleave,
// conjugation and the convolution product are fused into the passes that
// read or write HBM, so each transform moves exactly its input, output and
// (for convolution) kernel spectrum through HBM once. Each modulation is
// instantiated in its own kernels_lds_*.hip.
#pragma once

#include "drivers.h"
#include "kernels.h"

namespace fsa {


// ---------------------------------------------------------------- kernel
// E = 16 keeps a butterfly set within 128 VGPRs (1024-lane workgroups);
// E = 32 needs up to 256 and is launched with at most 512 lanes.
template <int E>
constexpr int lds_max_threads() { return E <= 16 ? 1024 : 512; }

// HBM-facing functors of one transform for each (modulation, operation).
template <int MOD, int OP>
struct LdsIo {
    static __device__ __forceinline__ auto ld(const LdsArgs& a, long long t)
    {
        const float* in = a.in + t * a.in_st;
        if constexpr (OP == OP_INV) {
            if constexpr (MOD == MOD_FO)
                return LdFoBins{(const float2*)in, a.L};
            else
                return LdCx<true>{(const float2*)in};
        } else {
            if constexpr (MOD == MOD_FO)
                return LdFoFold{in, a.L, a.tmod, a.tfo};
            else
                return LdCx<false>{(const float2*)in};
        }
    }
    static __device__ __forceinline__ auto st(const LdsArgs& a, long long t)
    {
        float* out = a.out + t * a.out_st;
        if constexpr (OP == OP_FWD) {
            if constexpr (MOD == MOD_FO)
                return StFoBins{(float2*)out, a.L};
            else
                return StCx<false>{(float2*)out};
        } else if constexpr (OP == OP_INV) {
            if constexpr (MOD == MOD_FO)
                return StFoDemod{out, a.L, a.tmod};
            else
                return StCx<true>{(float2*)out};
        } else {
            if constexpr (MOD == MOD_FO)
                return StFoFold{out, a.L, a.tmod, a.tfo};
            else
                return StCx<true>{(float2*)out};
        }
    }
};

// Workgroup b handles tiles b, b + gridDim.x, ... (G transforms per tile).
// PERSIST: the grid is the resident set and the next tile's first-pass
// inputs are loaded into registers while the current tile runs its LDS
// passes. Otherwise the grid covers every tile once and nothing is held
// across the tail (lower register pressure).
template <int MOD, int OP, int E, bool PERSIST>
__global__ void __launch_bounds__(lds_max_threads<E>()) k_lds(const LdsArgs a)
{
    extern __shared__ float2 smem[];
    using Io = LdsIo<MOD, OP>;
    const int T = a.T, G = a.G;
    const int g = threadIdx.x / T;
    const int lt0 = threadIdx.x - g * T;
    const bool in_tile = g < G;
    const TileContig tl{smem + (in_tile ? g : 0) * lds_len(a.L)};
    constexpr int kIdle = 1 << 28;

    if constexpr (!PERSIST) {
        // one tile per workgroup, first pass fused with its loads
        const long long tr = (long long)blockIdx.x * G + g;
        const bool ok = in_tile && tr < a.ntrans;
        const int lt = ok ? lt0 : kIdle;
        const long long t0 = ok ? tr : 0;
        if constexpr (OP == OP_FWD || OP == OP_INV) {
            run_stockham<E>(a.c, tl, lt, T, Io::ld(a, t0), Io::st(a, t0));
        } else if constexpr (OP == OP_KERN) {
            run_kern<E>(a.c, tl, lt, T, Io::ld(a, t0), (float2*)(a.out + t0 * a.out_st));
        } else {
            const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN, a.kern_st == 0,
                          a.spec ? (float2*)(a.spec + t0 * a.spec_st) : nullptr};
            run_conv<E>(a.c, tl, lt, T, Io::ld(a, t0), Io::st(a, t0), bot);
        }
        return;
    }

    const long long ntiles = (a.ntrans + G - 1) / G;
    const bool single = a.c.np == 1;
    float2 pf[E];

    long long tile = blockIdx.x;
    if (tile >= ntiles)
        return;
    {
        const long long tr = tile * G + g;
        const bool ok = in_tile && tr < a.ntrans;
        load_first<E>(a.c, pf, ok ? lt0 : kIdle, T, Io::ld(a, ok ? tr : 0));
    }
    while (tile < ntiles) {
        const long long tr = tile * G + g;
        const bool ok = in_tile && tr < a.ntrans;
        const int lt = ok ? lt0 : kIdle;
        const long long t0 = ok ? tr : 0;
        const long long next = tile + gridDim.x;
        const long long trn = next * G + g;
        const bool okn = next < ntiles && in_tile && trn < a.ntrans;
        const int ltn = okn ? lt0 : kIdle;
        const long long tn = okn ? trn : 0;

        if (single) {
            if constexpr (OP == OP_KERN) {
                pf_single_kern<E>(a.c, pf, lt, (float2*)(a.out + t0 * a.out_st));
            } else if constexpr (OP == OP_FWD || OP == OP_INV) {
                pf_single_stockham<E>(a.c, pf, lt, Io::st(a, t0));
            } else {
                const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN, a.kern_st == 0,
                          a.spec ? (float2*)(a.spec + t0 * a.spec_st) : nullptr};
                pf_single_conv<E>(a.c, pf, lt, bot, Io::st(a, t0));
            }
            if (PERSIST && next < ntiles)
                load_first<E>(a.c, pf, ltn, T, Io::ld(a, tn));
        } else {
            first_to_tile<E>(a.c, pf, tl, lt, T);
            if (PERSIST && next < ntiles)
                load_first<E>(a.c, pf, ltn, T, Io::ld(a, tn));
            if constexpr (OP == OP_FWD || OP == OP_INV) {
                stockham_tail<E>(a.c, tl, lt, T, Io::st(a, t0));
            } else if constexpr (OP == OP_KERN) {
                kern_tail<E>(a.c, tl, lt, T, (float2*)(a.out + t0 * a.out_st));
            } else {
                const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN, a.kern_st == 0,
                          a.spec ? (float2*)(a.spec + t0 * a.spec_st) : nullptr};
                conv_tail<E>(a.c, tl, lt, T, Io::st(a, t0), bot);
            }
        }
        __syncthreads();
        tile = next;
    }
}

template <int MOD, int OP, int E>
hipError_t launch_one(const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    const long long tiles = (a.ntrans + sh.G - 1) / sh.G;
    if (tiles <= 0)
        return hipSuccess;
    const bool persist = persistent_mode();
    auto fn = persist ? k_lds<MOD, OP, E, true> : k_lds<MOD, OP, E, false>;
    if (sh.lds_bytes > 64 * 1024) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, sh.lds_bytes);
    }
    const long long grid = persist ? persistent_grid((const void*)fn, sh.threads, sh.lds_bytes, tiles) : tiles;
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(sh.threads), sh.lds_bytes, st, a);
    return hipGetLastError();
}

template <int MOD, int E>
hipError_t launch_op(int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    switch (op) {
    case OP_FWD: return launch_one<MOD, OP_FWD, E>(a, sh, st);
    case OP_INV: return launch_one<MOD, OP_INV, E>(a, sh, st);
    case OP_KERN: return launch_one<MOD, OP_KERN, E>(a, sh, st);
    case OP_CONV: return launch_one<MOD, OP_CONV, E>(a, sh, st);
    case OP_GRAD_IN: return launch_one<MOD, OP_GRAD_IN, E>(a, sh, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace fsa
