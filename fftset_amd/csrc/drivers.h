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

}  // namespace fsa
