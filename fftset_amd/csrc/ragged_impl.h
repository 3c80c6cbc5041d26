// ragged_impl.h -- one launch over a ragged batch of transforms (the kernel;
// each workgroup size is instantiated in its own kernels_ragged_w*.hip).
//
// Every workgroup reads one tile descriptor and runs it with the same LDS
// engine the uniform kernels use: a tile is either G LDS-resident transforms
// of one class, or one column / row group of a four-step transform. Classes
// mix sizes, modulations and operations freely. Tiles are sorted largest
// first; workgroups are dispatched in blockIdx order, which gives a
// longest-processing-time-first schedule without a software work queue.
#pragma once

#include "drivers.h"
#include "kernels.h"

namespace fsa {

constexpr int kRaggedE = 16;
constexpr int kIdleLane = 1 << 28;

// FOLD_ST: the FREQOFFSETREAL demodulation folded into the convolution's
// closing pass too. Only the 128- and 256-lane kernels do (measured no
// faster at 512 lanes).
template <int MOD, int OP, bool FOLD_ST>
__device__ __forceinline__ void ragged_lds(const RClass& k, const RTile& t, float2* smem)
{
    const int T = k.T;
    const int g = threadIdx.x / T;
    const int lt0 = threadIdx.x - g * T;
    const bool ok = g < t.count;
    const int lt = ok ? lt0 : kIdleLane;
    const int j = t.job + (ok ? g : 0);
    const TileContig tl{smem + (ok ? g : 0) * lds_len(k.L)};
    const float* in = k.in[j];
    float* out = k.out[j];
    const int L = k.L;
    if constexpr (OP == OP_FWD) {
        if constexpr (MOD == MOD_FO)
            run_stockham<kRaggedE>(k.c, tl, lt, T, LdFoFold{in, L, k.tmod, k.tfo}, StFoBins{(float2*)out, L});
        else
            run_stockham<kRaggedE>(k.c, tl, lt, T, LdCx<false>{(const float2*)in}, StCx<false>{(float2*)out});
    } else if constexpr (OP == OP_INV) {
        if constexpr (MOD == MOD_FO)
            run_stockham<kRaggedE>(k.c, tl, lt, T, LdFoBins{(const float2*)in, L}, StFoDemod{out, L, k.tmod});
        else
            run_stockham<kRaggedE>(k.c, tl, lt, T, LdCx<true>{(const float2*)in}, StCx<true>{(float2*)out});
    } else if constexpr (OP == OP_KERN) {
        if constexpr (MOD == MOD_FO)
            run_kern<kRaggedE>(k.c, tl, lt, T, LdFoFold{in, L, k.tmod, k.tfo}, (float2*)out);
        else
            run_kern<kRaggedE>(k.c, tl, lt, T, LdCx<false>{(const float2*)in}, (float2*)out);
    } else {
        const BotConv bot{(const float2*)k.kern[j], OP == OP_GRAD_IN};
        if constexpr (MOD == MOD_FO && FOLD_ST)
            run_conv<kRaggedE>(k.c, tl, lt, T, LdFoFold{in, L, k.tmod, k.tfo}, StFoFold{out, L, k.tmod, k.tfo},
                               bot);
        else if constexpr (MOD == MOD_FO)
            run_conv<kRaggedE>(k.c, tl, lt, T, LdFoFold{in, L, k.tmod, k.tfo}, StFoDemod{out, L, k.tmod}, bot);
        else
            run_conv<kRaggedE>(k.c, tl, lt, T, LdCx<false>{(const float2*)in}, StCx<true>{(float2*)out}, bot);
    }
}

// Column step of a four-step transform: first (COL) or, for convolutions,
// the closing step (COL2) that reads the work buffer.
template <int MOD, int OP, bool SECOND>
__device__ __forceinline__ void ragged_col(const RClass& k, const RTile& t, float2* smem)
{
    const int C = k.colC, T = k.colT;
    const int g = threadIdx.x % C;
    const int lt0 = threadIdx.x / C;
    const int n2 = t.count * C + g;
    const bool ok = lt0 < T && n2 < k.B;
    const int lt = ok ? lt0 : kIdleLane;
    const int n2c = ok ? n2 : 0;
    const TileInter tl{smem + g, C};
    const int L = k.L, B = k.B;
    const int j = t.job;
    if constexpr (SECOND) {
        // work -> out, closing the conjugate-data inverse of the convolution
        const LdCol<LdCx<false>> ld{{(const float2*)k.work[j]}, B, n2c};
        if constexpr (MOD == MOD_FO)
            run_stockham<kRaggedE>(k.c, tl, lt, T, ld, StCol<StFoDemod>{{k.out[j], L, k.tmod}, B, n2c});
        else
            run_stockham<kRaggedE>(k.c, tl, lt, T, ld, StCol<StCx<true>>{{(float2*)k.out[j]}, B, n2c});
    } else {
        // in -> work (get_kernel: in -> out, which it then transforms in place)
        float2* dst = (float2*)(OP == OP_KERN ? k.out[j] : k.work[j]);
        const StWorkTw st{dst, B, n2c, k.tfs};
        const float* in = k.in[j];
        if constexpr (OP == OP_INV) {
            if constexpr (MOD == MOD_FO)
                run_stockham<kRaggedE>(k.c, tl, lt, T, LdCol<LdFoBins>{{(const float2*)in, L}, B, n2c}, st);
            else
                run_stockham<kRaggedE>(k.c, tl, lt, T, LdCol<LdCx<true>>{{(const float2*)in}, B, n2c}, st);
        } else {
            if constexpr (MOD == MOD_FO)
                run_stockham<kRaggedE>(k.c, tl, lt, T, LdCol<LdFoMod>{{in, L, k.tmod}, B, n2c}, st);
            else
                run_stockham<kRaggedE>(k.c, tl, lt, T, LdCol<LdCx<false>>{{(const float2*)in}, B, n2c}, st);
        }
    }
}

// Row step of a four-step transform.
template <int MOD, int OP>
__device__ __forceinline__ void ragged_row(const RClass& k, const RTile& t, float2* smem)
{
    const int T = k.rowT, C = k.rowR;
    const int g = threadIdx.x / T;
    const int lt0 = threadIdx.x - g * T;
    const int k1_0 = t.count * C;
    const int k1 = k1_0 + g;
    const bool ok = g < C && k1 < k.A;
    const int lt = ok ? lt0 : kIdleLane;
    const int k1c = ok ? k1 : 0;
    const int LS = lds_len(k.B);
    const TileContig tl{smem + (ok ? g : 0) * LS};
    const int j = t.job;
    const float2* src = (const float2*)(OP == OP_KERN ? k.out[j] : k.work[j]);
    const LdCx<false> ld{src + (long long)k1c * k.B};
    if constexpr (OP == OP_FWD || OP == OP_INV) {
        run_stockham_to_tile<kRaggedE>(k.c2, tl, lt, T, ld);
        __syncthreads();
        const int nrows = min(C, k.A - k1_0);
        float* out = k.out[j];
        const int L = k.L, A = k.A, B = k.B;
        for (int idx = threadIdx.x; idx < nrows * B; idx += blockDim.x) {
            const int gg = idx % nrows, k2 = idx / nrows;
            const float2 v = smem[gg * LS + lds_pad(k2)];
            const int f = k1_0 + gg + A * k2;
            if constexpr (OP == OP_FWD) {
                if constexpr (MOD == MOD_FO)
                    StFoBins{(float2*)out, L}(f, v);
                else
                    StCx<false>{(float2*)out}(f, v);
            } else {
                if constexpr (MOD == MOD_FO)
                    StFoDemod{out, L, k.tmod}(f, v);
                else
                    StCx<true>{(float2*)out}(f, v);
            }
        }
    } else if constexpr (OP == OP_KERN) {
        run_kern<kRaggedE>(k.c2, tl, lt, T, ld, k.out[j] ? (float2*)k.out[j] + (long long)k1c * k.B : nullptr);
    } else {
        float2* row = (float2*)k.work[j] + (long long)k1c * k.B;
        const BotConv bot{(const float2*)k.kern[j] + (long long)k1c * k.B, OP == OP_GRAD_IN};
        run_conv<kRaggedE>(k.c2, tl, lt, T, ld, StRowTw{row, k.tfs + (long long)k1c * k.B}, bot);
    }
}

// LDS tiles and four-step tiles are compiled into separate kernels (each
// launch range holds one kind), which keeps each kernel's code small.
template <int MOD, int OP, bool FS, bool FOLD_ST>
__device__ __forceinline__ void ragged_dispatch_kind(const RClass& k, const RTile& t, float2* smem)
{
    if constexpr (!FS) {
        ragged_lds<MOD, OP, FOLD_ST>(k, t, smem);
    } else {
        switch (t.kind) {
        case TK_COL: ragged_col<MOD, OP, false>(k, t, smem); break;
        case TK_ROW: ragged_row<MOD, OP>(k, t, smem); break;
        case TK_COL2:
            if constexpr (OP == OP_CONV || OP == OP_GRAD_IN)
                ragged_col<MOD, OP, true>(k, t, smem);
            break;
        default: break;
        }
    }
}

template <bool FS, bool FOLD_ST>
__device__ __forceinline__ void ragged_tile(const RaggedArgs& a, int ti, float2* smem)
{
    const RTile t = a.tiles[ti];
    const RClass& k = a.classes[t.cls];
    const int sel = k.mod * OP_COUNT + k.op;
    switch (sel) {
    case MOD_FO * OP_COUNT + OP_FWD: ragged_dispatch_kind<MOD_FO, OP_FWD, FS, FOLD_ST>(k, t, smem); break;
    case MOD_FO * OP_COUNT + OP_INV: ragged_dispatch_kind<MOD_FO, OP_INV, FS, FOLD_ST>(k, t, smem); break;
    case MOD_FO * OP_COUNT + OP_KERN: ragged_dispatch_kind<MOD_FO, OP_KERN, FS, FOLD_ST>(k, t, smem); break;
    case MOD_FO * OP_COUNT + OP_CONV: ragged_dispatch_kind<MOD_FO, OP_CONV, FS, FOLD_ST>(k, t, smem); break;
    case MOD_FO * OP_COUNT + OP_GRAD_IN: ragged_dispatch_kind<MOD_FO, OP_GRAD_IN, FS, FOLD_ST>(k, t, smem); break;
    case MOD_CX * OP_COUNT + OP_FWD: ragged_dispatch_kind<MOD_CX, OP_FWD, FS, FOLD_ST>(k, t, smem); break;
    case MOD_CX * OP_COUNT + OP_INV: ragged_dispatch_kind<MOD_CX, OP_INV, FS, FOLD_ST>(k, t, smem); break;
    case MOD_CX * OP_COUNT + OP_KERN: ragged_dispatch_kind<MOD_CX, OP_KERN, FS, FOLD_ST>(k, t, smem); break;
    case MOD_CX * OP_COUNT + OP_CONV: ragged_dispatch_kind<MOD_CX, OP_CONV, FS, FOLD_ST>(k, t, smem); break;
    case MOD_CX * OP_COUNT + OP_GRAD_IN: ragged_dispatch_kind<MOD_CX, OP_GRAD_IN, FS, FOLD_ST>(k, t, smem); break;
    default: break;
    }
}

// One tile per workgroup (grid = tiles), or persistent: a resident grid
// whose workgroups take tiles in table order from a counter, so the
// largest-first order holds across the whole launch. Tiles of one launch
// never depend on each other, so no workgroup ever waits on another.
//
// At most 128 VGPRs (four waves per SIMD): left to itself the compiler gives
// the kernel 132 for its FOLD_ST convolution, and with them every tile of
// the launch, whatever its operation, runs at three waves per SIMD. The cap
// costs no spills; a batch of 20000 1440-bin COMPLEX forward transforms took
// 140.5 us with it against 147.0 without.
template <int THREADS, bool FS>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_ragged(const RaggedArgs a)
{
    extern __shared__ float2 smem[];
    __shared__ int s_tile;
    int ti = blockIdx.x;
    for (;;) {  // one call site of the (large, inlined) tile body
        if (a.next) {
            if (threadIdx.x == 0)
                s_tile = atomicAdd(a.next, 1);
            __syncthreads();
            ti = s_tile;
            if (ti >= a.ntiles)
                return;  // the same value in every lane: the whole workgroup leaves
        }
        ragged_tile<FS, (THREADS < 512)>(a, ti, smem);
        if (!a.next)
            return;
        __syncthreads();  // LDS and s_tile are reused by the next tile
    }
}

template <int THREADS>
__global__ void k_ragged_reset(int* next)
{
    *next = 0;
}

template <int THREADS>
hipError_t launch_w(const RaggedArgs& a, int lds_bytes, hipStream_t st)
{
    auto fn = a.four_step ? k_ragged<THREADS, true> : k_ragged<THREADS, false>;
    if (lds_bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    long long grid = a.ntiles;
    if (a.next) {
        // a kernel node, not a memset: the counter is reset inside captured graphs too
        (void)hipGetLastError();
        hipLaunchKernelGGL(k_ragged_reset<THREADS>, dim3(1), dim3(1), 0, st, a.next);
        if (hipGetLastError() != hipSuccess)
            return hipErrorLaunchFailure;
        grid = persistent_grid((const void*)fn, THREADS, lds_bytes, a.ntiles);
    }
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(THREADS), lds_bytes, st, a);
    return hipGetLastError();
}

}  // namespace fsa
