// drivers.h -- pass-chain drivers over a runtime radix chain.
//
// Every pass dispatches on its radix (uniform across the workgroup) to a
// fully unrolled register butterfly; E is the per-thread element budget that
// bounds the butterflies a lane may own in one pass.
#pragma once

#include "engine.h"
#include "plan_desc.h"

namespace fsa {

// Radix dispatch: BODY sees constexpr RR (radix) and NB (butterflies per
// thread for the element budget E).
#define FSA_RADIX_CASE(RV, E, BODY)                                   \
    case RV: {                                                        \
        constexpr int RR = RV;                                        \
        constexpr int NB = ((E) / RV) > 0 ? ((E) / RV) : 1;           \
        (void)NB;                                                     \
        BODY;                                                         \
    } break;
#define FSA_RADIX_SWITCH(R, E, BODY)                                  \
    switch (R) {                                                      \
        FSA_RADIX_CASE(2, E, BODY)                                    \
        FSA_RADIX_CASE(3, E, BODY)                                    \
        FSA_RADIX_CASE(4, E, BODY)                                    \
        FSA_RADIX_CASE(5, E, BODY)                                    \
        FSA_RADIX_CASE(6, E, BODY)                                    \
        FSA_RADIX_CASE(8, E, BODY)                                    \
        FSA_RADIX_CASE(9, E, BODY)                                    \
        FSA_RADIX_CASE(10, E, BODY)                                   \
        FSA_RADIX_CASE(12, E, BODY)                                   \
        FSA_RADIX_CASE(15, E, BODY)                                   \
        FSA_RADIX_CASE(16, E, BODY)                                   \
        default: break;                                               \
    }

// ---------------------------------------------------------------- drivers
template <int R, class Load, class Store>
__device__ __forceinline__ void single_conv(int lt, const Load& ld, const Store& st, const BotConv& bot)
{
    if (lt == 0) {
        float2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = ld(t);
        DFT<R>::run(v);
        if (bot.spec) {
#pragma unroll
            for (int k = 0; k < R; ++k) bot.spec[k] = v[k];
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float2 K = bot.kern[k];
            v[k] = cconj(bot.conj_kern ? cmulc(v[k], K) : cmul(v[k], K));
        }
        DFT<R>::run(v);
#pragma unroll
        for (int k = 0; k < R; ++k) st(k, v[k]);
    }
}

template <int R, class Load>
__device__ __forceinline__ void single_kern(int lt, const Load& ld, float2* out)
{
    if (lt == 0) {
        float2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = ld(t);
        DFT<R>::run(v);
#pragma unroll
        for (int k = 0; k < R; ++k) out[k] = v[k];
    }
}

// Stockham chain: natural order in (through ld), natural order out (st).
template <int E, class Tile, class Load, class Store>
__device__ __forceinline__ void run_stockham(const ChainD& c, const Tile& tl, int lt, int T,
                                             const Load& ld, const Store& st)
{
    const int np = c.np;
    if (np == 1) {
        FSA_RADIX_SWITCH(c.p[0].R, E, (single_pass<RR>(lt, ld, st)));
        return;
    }
    FSA_RADIX_SWITCH(c.p[0].R, E, (pass_first<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, ld)));
    __syncthreads();
    for (int i = 1; i < np - 1; ++i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (stock_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    FSA_RADIX_SWITCH(c.p[np - 1].R, E, (stock_last<RR, NB>(tl, lt, T, c.p[np - 1].s, st)));
}

// In-place DIF chain, product at the bottom, in-place DIT chain back.
template <int E, class Tile, class Load, class Store>
__device__ __forceinline__ void run_conv(const ChainD& c, const Tile& tl, int lt, int T,
                                         const Load& ld, const Store& st, const BotConv& bot)
{
    const int np = c.np;
    if (np == 1) {
        FSA_RADIX_SWITCH(c.p[0].R, E, (single_conv<RR>(lt, ld, st, bot)));
        return;
    }
    FSA_RADIX_SWITCH(c.p[0].R, E, (pass_first<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, ld)));
    __syncthreads();
    for (int i = 1; i < np - 1; ++i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (dif_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    FSA_RADIX_SWITCH(c.p[np - 1].R, E, (dif_bottom<RR, NB>(tl, lt, T, c.p[np - 1].s, bot)));
    __syncthreads();
    for (int i = np - 2; i >= 1; --i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (dit_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    FSA_RADIX_SWITCH(c.p[0].R, E, (dit_last<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, st)));
}

// In-place DIF chain whose bottom writes the kernel spectrum.
template <int E, class Tile, class Load>
__device__ __forceinline__ void run_kern(const ChainD& c, const Tile& tl, int lt, int T,
                                         const Load& ld, float2* out)
{
    const int np = c.np;
    if (np == 1) {
        FSA_RADIX_SWITCH(c.p[0].R, E, (single_kern<RR>(lt, ld, out)));
        return;
    }
    FSA_RADIX_SWITCH(c.p[0].R, E, (pass_first<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, ld)));
    __syncthreads();
    for (int i = 1; i < np - 1; ++i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (dif_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    const BotKernOut bot{out};
    FSA_RADIX_SWITCH(c.p[np - 1].R, E, (dif_bottom<RR, NB>(tl, lt, T, c.p[np - 1].s, bot)));
}

// ------------------------------------------------ persistent-loop stages
// A workgroup that loops over transforms calls load_first() for the next
// transform right after first_to_tile() of the current one has consumed the
// prefetch registers, so HBM latency overlaps the remaining LDS passes.
// Chains with a single pass (n == R) use the single_* stages instead.

template <int E, class Load>
__device__ __forceinline__ void load_first(const ChainD& c, float2* pf, int lt, int T, const Load& ld)
{
    FSA_RADIX_SWITCH(c.p[0].R, E, (first_load<RR, NB>(pf, lt, T, c.p[0].m, ld)));
}

template <int E, class Tile>
__device__ __forceinline__ void first_to_tile(const ChainD& c, float2* pf, const Tile& tl, int lt, int T)
{
    FSA_RADIX_SWITCH(c.p[0].R, E, (first_compute<RR, NB>(pf, tl, lt, T, c.p[0].m, c.p[0].tw)));
}

template <int R, class Store>
__device__ __forceinline__ void single_fft_pf(float2* pf, int lt, const Store& st)
{
    if (lt == 0) {
        DFT<R>::run(pf);
#pragma unroll
        for (int k = 0; k < R; ++k) st(k, pf[k]);
    }
}

template <int R, class Store>
__device__ __forceinline__ void single_conv_pf(float2* pf, int lt, const BotConv& bot, const Store& st)
{
    if (lt == 0) {
        DFT<R>::run(pf);
        if (bot.spec) {
#pragma unroll
            for (int k = 0; k < R; ++k) bot.spec[k] = pf[k];
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float2 K = bot.kern[k];
            pf[k] = cconj(bot.conj_kern ? cmulc(pf[k], K) : cmul(pf[k], K));
        }
        DFT<R>::run(pf);
#pragma unroll
        for (int k = 0; k < R; ++k) st(k, pf[k]);
    }
}

// n == R: the whole transform from the prefetch registers.
template <int E, class Store>
__device__ __forceinline__ void pf_single_stockham(const ChainD& c, float2* pf, int lt, const Store& st)
{
    FSA_RADIX_SWITCH(c.p[0].R, E, (single_fft_pf<RR>(pf, lt, st)));
}

template <int E, class Store>
__device__ __forceinline__ void pf_single_conv(const ChainD& c, float2* pf, int lt, const BotConv& bot,
                                               const Store& st)
{
    FSA_RADIX_SWITCH(c.p[0].R, E, (single_conv_pf<RR>(pf, lt, bot, st)));
}

template <int E>
__device__ __forceinline__ void pf_single_kern(const ChainD& c, float2* pf, int lt, float2* out)
{
    FSA_RADIX_SWITCH(c.p[0].R, E, (single_fft_pf<RR>(pf, lt, StCx<false>{out})));
}

template <int E, class Tile>
__device__ __forceinline__ void pf_single_tile(const ChainD& c, float2* pf, int lt, const Tile& tl)
{
    FSA_RADIX_SWITCH(c.p[0].R, E, (single_fft_pf<RR>(pf, lt, TileStore<Tile>{tl})));
}

// Remaining Stockham passes after first_to_tile().
template <int E, class Tile, class Store>
__device__ __forceinline__ void stockham_tail(const ChainD& c, const Tile& tl, int lt, int T, const Store& st)
{
    const int np = c.np;
    __syncthreads();
    for (int i = 1; i < np - 1; ++i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (stock_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    FSA_RADIX_SWITCH(c.p[np - 1].R, E, (stock_last<RR, NB>(tl, lt, T, c.p[np - 1].s, st)));
}

// Remaining Stockham passes, result left in the tile in natural order.
template <int E, class Tile>
__device__ __forceinline__ void stockham_tail_to_tile(const ChainD& c, const Tile& tl, int lt, int T)
{
    const int np = c.np;
    for (int i = 1; i < np; ++i) {
        __syncthreads();
        FSA_RADIX_SWITCH(c.p[i].R, E, (stock_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
    }
}

// Remaining in-place DIF passes, product at the bottom, DIT back.
template <int E, class Tile, class Store>
__device__ __forceinline__ void conv_tail(const ChainD& c, const Tile& tl, int lt, int T, const Store& st,
                                          const BotConv& bot)
{
    const int np = c.np;
    __syncthreads();
    for (int i = 1; i < np - 1; ++i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (dif_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    FSA_RADIX_SWITCH(c.p[np - 1].R, E, (dif_bottom<RR, NB>(tl, lt, T, c.p[np - 1].s, bot)));
    __syncthreads();
    for (int i = np - 2; i >= 1; --i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (dit_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    FSA_RADIX_SWITCH(c.p[0].R, E, (dit_last<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, st)));
}

// Remaining in-place DIF passes; the bottom writes the kernel spectrum.
template <int E, class Tile>
__device__ __forceinline__ void kern_tail(const ChainD& c, const Tile& tl, int lt, int T, float2* out)
{
    const int np = c.np;
    __syncthreads();
    for (int i = 1; i < np - 1; ++i) {
        FSA_RADIX_SWITCH(c.p[i].R, E, (dif_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
        __syncthreads();
    }
    const BotKernOut bot{out};
    FSA_RADIX_SWITCH(c.p[np - 1].R, E, (dif_bottom<RR, NB>(tl, lt, T, c.p[np - 1].s, bot)));
}

// Stockham chain whose last pass leaves the result in the tile, natural
// order (for outputs that must be re-mapped before they go to HBM).
template <int E, class Tile, class Load>
__device__ __forceinline__ void run_stockham_to_tile(const ChainD& c, const Tile& tl, int lt, int T,
                                                     const Load& ld)
{
    const int np = c.np;
    if (np == 1) {
        FSA_RADIX_SWITCH(c.p[0].R, E, (single_pass<RR>(lt, ld, TileStore<Tile>{tl})));
        return;
    }
    FSA_RADIX_SWITCH(c.p[0].R, E, (pass_first<RR, NB>(tl, lt, T, c.p[0].m, c.p[0].tw, ld)));
    __syncthreads();
    for (int i = 1; i < np; ++i) {
        if (i > 1)
            __syncthreads();
        FSA_RADIX_SWITCH(c.p[i].R, E, (stock_mid<RR, NB>(tl, lt, T, c.p[i].m, c.p[i].s, c.p[i].tw)));
    }
}

// ------------------------------------------- two-pass chains, compile time
// A chain [R0, R1] fixed at compile time (16 elements per lane): the same
// passes, twiddle tables and data order as the runtime-chain drivers above,
// with constant radices, strides and lane counts.
template <int R0, int R1>
struct Fix2 {
    static_assert(R0 > 0 && R1 > 0, "a two-pass chain");
    static constexpr int n = R0 * R1;
    static constexpr int m0 = n / R0, s1 = R0;
    static constexpr int NB0 = (16 / R0) > 0 ? 16 / R0 : 1;
    static constexpr int NB1 = (16 / R1) > 0 ? 16 / R1 : 1;
    static constexpr int T0 = (m0 + NB0 - 1) / NB0, T1 = (s1 + NB1 - 1) / NB1;
    static constexpr int T = T0 > T1 ? T0 : T1;  // lanes per transform

    // true when a runtime chain and lane count are exactly this one
    static bool matches(const ChainD& c, int lanes)
    {
        return c.np == 2 && c.p[0].R == R0 && c.p[1].R == R1 && c.p[0].m == m0 && c.p[1].m == 1 &&
               c.p[1].s == s1 && lanes == T;
    }
};

template <int R0, int R1, class Tile, class Load, class Store>
__device__ __forceinline__ void run_stockham_fix2(const ChainD& c, const Tile& tl, int lt, const Load& ld,
                                                  const Store& st)
{
    using F = Fix2<R0, R1>;
    pass_first<R0, F::NB0>(tl, lt, F::T, F::m0, c.p[0].tw, ld);
    __syncthreads();
    stock_last<R1, F::NB1>(tl, lt, F::T, F::s1, st);
}

template <int R0, int R1, class Tile, class Load>
__device__ __forceinline__ void run_stockham_to_tile_fix2(const ChainD& c, const Tile& tl, int lt, const Load& ld)
{
    using F = Fix2<R0, R1>;
    pass_first<R0, F::NB0>(tl, lt, F::T, F::m0, c.p[0].tw, ld);
    __syncthreads();
    stock_mid<R1, F::NB1>(tl, lt, F::T, 1, F::s1, c.p[1].tw);
}

template <int R0, int R1, class Tile, class Load, class Store, class Bottom>
__device__ __forceinline__ void run_conv_fix2(const ChainD& c, const Tile& tl, int lt, const Load& ld,
                                              const Store& st, const Bottom& bot)
{
    using F = Fix2<R0, R1>;
    pass_first<R0, F::NB0>(tl, lt, F::T, F::m0, c.p[0].tw, ld);
    __syncthreads();
    dif_bottom<R1, F::NB1>(tl, lt, F::T, F::s1, bot);
    __syncthreads();
    dit_last<R0, F::NB0>(tl, lt, F::T, F::m0, c.p[0].tw, st);
}

template <int R0, int R1, class Tile, class Load>
__device__ __forceinline__ void run_kern_fix2(const ChainD& c, const Tile& tl, int lt, const Load& ld, float2* out)
{
    using F = Fix2<R0, R1>;
    pass_first<R0, F::NB0>(tl, lt, F::T, F::m0, c.p[0].tw, ld);
    __syncthreads();
    dif_bottom<R1, F::NB1>(tl, lt, F::T, F::s1, BotKernOut{out});
}

}  // namespace fsa
