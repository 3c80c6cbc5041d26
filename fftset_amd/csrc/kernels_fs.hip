// kernels_fs.hip -- four-step steps for transforms larger than LDS.
//
// n = A*B. The column step runs B transforms of length A over elements
// B*n1 + n2 (C adjacent columns per workgroup, lanes across columns so every
// global access is a contiguous run of C values) and writes
// Y[k1][n2] * W_n^{n2 k1} to a row-major work buffer. The row step runs A
// transforms of length B over contiguous rows. Forward / inverse:
// X[k1 + A*k2] = DFT_B(row k1)[k2], written through LDS so the transposed
// store is again contiguous across rows. Convolution: the row step does
// DIF -> product -> DIT in place (the kernel spectrum is stored per row in
// the same slot-major order) and a second column step finishes the
// conjugate-data inverse.
#include "drivers.h"
#include "kernels.h"

namespace fsa {

// Transform t's buffers: base + t * stride, or the per-transform pointer
// arrays of a ragged batch.
__device__ __forceinline__ const float* fs_in(const FsArgs& a, long long t)
{
    return a.in_p ? a.in_p[t] : a.in + t * a.in_st;
}
__device__ __forceinline__ float* fs_out(const FsArgs& a, long long t)
{
    return a.out_p ? a.out_p[t] : a.out + t * a.out_st;
}
__device__ __forceinline__ const float* fs_kern(const FsArgs& a, long long t)
{
    return a.kern_p ? a.kern_p[t] : a.kern + t * a.kern_st;
}
// one kernel spectrum shared by every transform (kernel stride 0)
__device__ __forceinline__ bool fs_kern_shared(const FsArgs& a) { return !a.kern_p && a.kern_st == 0; }

// Column-step functors of one (transform, column) pair.
template <int LK, int SK>
struct ColIo {
    static __device__ __forceinline__ auto ld(const FsArgs& a, long long t, int n2)
    {
        const float* in = fs_in(a, t);
        if constexpr (LK == CL_FOMOD)
            return LdCol<LdFoMod>{{in, a.L, a.tmod}, a.B, n2};
        else if constexpr (LK == CL_FOMODC)
            return LdColFoMod{in, a.L, a.B, n2, a.trow};
        else if constexpr (LK == CL_FOBINS)
            return LdCol<LdFoBins>{{(const float2*)in, a.L}, a.B, n2};
        else if constexpr (LK == CL_CX)
            return LdCol<LdCx<false>>{{(const float2*)in}, a.B, n2};
        else
            return LdCol<LdCx<true>>{{(const float2*)in}, a.B, n2};
    }
    static __device__ __forceinline__ auto st(const FsArgs& a, long long t, int n2)
    {
        float* out = fs_out(a, t);
        if constexpr (SK == CS_WORKTW)
            return StWorkTw{(float2*)out, a.B, n2, a.tfs};
        else if constexpr (SK == CS_FODEMOD)
            return StCol<StFoDemod>{{out, a.L, a.tmod}, a.B, n2};
        else
            return StCol<StCx<true>>{{(float2*)out}, a.B, n2};
    }
};

// Persistent column step: tile = (transform, group of C adjacent columns).
// R0 > 0: the column chain is [R0, R1], fixed at compile time.
template <int LK, int SK, int E, bool PERSIST, int R0 = 0, int R1 = 0>
__global__ void __launch_bounds__(1024) k_cols(const FsArgs a)
{
    extern __shared__ float2 smem[];
    using Io = ColIo<LK, SK>;
    constexpr int kIdle = 1 << 28;
    const int C = a.C, T = a.T;
    const int g = threadIdx.x % C;
    const int lt0 = threadIdx.x / C;
    const bool in_col = lt0 < T;
    const TileInter tl{smem + g, C};

    if constexpr (!PERSIST) {
        const long long t = blockIdx.x / a.groups;
        const int n2 = (int)(blockIdx.x - t * a.groups) * C + g;
        const bool ok = in_col && n2 < a.B && t < a.ntrans;
        if constexpr (R0 > 0)
            run_stockham_fix2<R0, R1>(a.c, tl, ok ? lt0 : kIdle, Io::ld(a, ok ? t : 0, ok ? n2 : 0),
                                      Io::st(a, ok ? t : 0, ok ? n2 : 0));
        else
            run_stockham<E>(a.c, tl, ok ? lt0 : kIdle, T, Io::ld(a, ok ? t : 0, ok ? n2 : 0),
                            Io::st(a, ok ? t : 0, ok ? n2 : 0));
        return;
    }

    const long long ntiles = a.ntrans * a.groups;
    const bool single = a.c.np == 1;
    float2 pf[E];

    auto lane = [&](long long tile, long long* t, int* n2) {
        *t = tile / a.groups;
        *n2 = (int)(tile - *t * a.groups) * C + g;
        const bool ok = in_col && *n2 < a.B && *t < a.ntrans;
        if (!ok) {
            *t = 0;
            *n2 = 0;
        }
        return ok ? lt0 : kIdle;
    };
    long long tile = blockIdx.x;
    if (tile >= ntiles)
        return;
    {
        long long t;
        int n2;
        const int lt = lane(tile, &t, &n2);
        load_first<E>(a.c, pf, lt, T, Io::ld(a, t, n2));
    }
    while (tile < ntiles) {
        long long t, tn;
        int n2, n2n;
        const int lt = lane(tile, &t, &n2);
        const long long next = tile + gridDim.x;
        const int ltn = next < ntiles ? lane(next, &tn, &n2n) : kIdle;
        if (single) {
            pf_single_stockham<E>(a.c, pf, lt, Io::st(a, t, n2));
            if (PERSIST && next < ntiles)
                load_first<E>(a.c, pf, ltn, T, Io::ld(a, tn, n2n));
        } else {
            first_to_tile<E>(a.c, pf, tl, lt, T);
            if (PERSIST && next < ntiles)
                load_first<E>(a.c, pf, ltn, T, Io::ld(a, tn, n2n));
            stockham_tail<E>(a.c, tl, lt, T, Io::st(a, t, n2));
        }
        __syncthreads();
        tile = next;
    }
}

// Persistent row step: tile = (transform, group of C rows).
// R0 > 0: the row chain is [R0, R1], fixed at compile time. BOT >= 0 (with
// R0 > 0, convolution): the adjoint (bit 0) and shared-spectrum (bit 1)
// choices of the bottom fixed at compile time too.
template <int KIND, int SK, int E, bool PERSIST, int R0 = 0, int R1 = 0, int BOT = -1>
__global__ void __launch_bounds__(1024) k_rows(const FsArgs a)
{
    extern __shared__ float2 smem[];
    constexpr int kIdle = 1 << 28;
    int T = a.T;
    if constexpr (R0 > 0)
        T = Fix2<R0, R1>::T;
    const int C = a.C;
    const int g = threadIdx.x / T;
    const int lt0 = threadIdx.x - g * T;
    const int LS = lds_len(a.B);
    const TileContig tl{smem + (g < C ? g : 0) * LS};

    if constexpr (!PERSIST) {
        const long long t = blockIdx.x / a.groups;
        const int k1_0 = (int)(blockIdx.x - t * a.groups) * C;
        const int k1 = k1_0 + g;
        const bool ok = g < C && k1 < a.A && t < a.ntrans;
        const int lt = ok ? lt0 : kIdle;
        const int k1c = ok ? k1 : 0;
        const long long t0 = t < a.ntrans ? t : 0;
        const LdCx<false> ld{(const float2*)fs_in(a, t0) + (long long)k1c * a.B};
        float2* row_out = (float2*)fs_out(a, t0) + (long long)k1c * a.B;
        if constexpr (KIND == RK_STOCK) {
            if constexpr (R0 > 0)
                run_stockham_to_tile_fix2<R0, R1>(a.c, tl, lt, ld);
            else
                run_stockham_to_tile<E>(a.c, tl, lt, T, ld);
            __syncthreads();
            const int nrows = min(C, a.A - k1_0);
            if (t >= a.ntrans || nrows <= 0)
                return;
            float* out = fs_out(a, t);
            const int L = a.L, A = a.A;
            for (int idx = threadIdx.x; idx < nrows * a.B; idx += blockDim.x) {
                const int gg = idx % nrows, k2 = idx / nrows;
                const float2 v = smem[gg * LS + lds_pad(k2)];
                const int f = k1_0 + gg + A * k2;
                if constexpr (SK == RS_FOBINS)
                    StFoBins{(float2*)out, L}(f, v);
                else if constexpr (SK == RS_CX)
                    StCx<false>{(float2*)out}(f, v);
                else if constexpr (SK == RS_FODEMOD)
                    StFoDemod{out, L, a.tmod}(f, v);
                else
                    StCx<true>{(float2*)out}(f, v);
            }
        } else if constexpr (KIND == RK_CONV) {
            float2* spec = a.spec ? (float2*)(a.spec + t0 * a.spec_st) + (long long)k1c * a.B : nullptr;
            const BotConv bot{(const float2*)fs_kern(a, t0) + (long long)k1c * a.B, a.conj_kern != 0, fs_kern_shared(a),
                              spec};
            const StRowTw st{row_out, a.tfs + (long long)k1c * a.B};
            if constexpr (R0 > 0 && BOT >= 0)
                run_conv_fix2<R0, R1>(a.c, tl, lt, ld, st,
                                      BotConvC<(BOT & 1) != 0, (BOT & 2) != 0>{bot.kern, spec});
            else if constexpr (R0 > 0)
                run_conv_fix2<R0, R1>(a.c, tl, lt, ld, st, bot);
            else
                run_conv<E>(a.c, tl, lt, T, ld, st, bot);
        } else {
            if constexpr (R0 > 0)
                run_kern_fix2<R0, R1>(a.c, tl, lt, ld, row_out);
            else
                run_kern<E>(a.c, tl, lt, T, ld, row_out);
        }
        return;
    }

    const long long ntiles = a.ntrans * a.groups;
    const bool single = a.c.np == 1;
    float2 pf[E];

    auto lane = [&](long long tile, long long* t, int* k1) {
        *t = tile / a.groups;
        *k1 = (int)(tile - *t * a.groups) * C + g;
        const bool ok = g < C && *k1 < a.A && *t < a.ntrans;
        if (!ok) {
            *t = 0;
            *k1 = 0;
        }
        return ok ? lt0 : kIdle;
    };
    auto row_ld = [&](long long t, int k1) {
        return LdCx<false>{(const float2*)fs_in(a, t) + (long long)k1 * a.B};
    };
    long long tile = blockIdx.x;
    if (tile >= ntiles)
        return;
    {
        long long t;
        int k1;
        const int lt = lane(tile, &t, &k1);
        load_first<E>(a.c, pf, lt, T, row_ld(t, k1));
    }
    while (tile < ntiles) {
        long long t, tn;
        int k1, k1n;
        const int lt = lane(tile, &t, &k1);
        const long long next = tile + gridDim.x;
        const int ltn = next < ntiles ? lane(next, &tn, &k1n) : kIdle;
        float2* row_out = (float2*)fs_out(a, t) + (long long)k1 * a.B;

        if constexpr (KIND == RK_STOCK) {
            if (single)
                pf_single_tile<E>(a.c, pf, lt, tl);
            else
                first_to_tile<E>(a.c, pf, tl, lt, T);
            if (PERSIST && next < ntiles)
                load_first<E>(a.c, pf, ltn, T, row_ld(tn, k1n));
            if (!single)
                stockham_tail_to_tile<E>(a.c, tl, lt, T);
            __syncthreads();
            // transposed store: lanes run across the rows of this tile so
            // consecutive lanes write consecutive outputs f = k1 + A*k2
            const long long tt = tile / a.groups;
            const int k1_0 = (int)(tile - tt * a.groups) * C;
            const int nrows = min(C, a.A - k1_0);
            float* out = fs_out(a, tt);
            const int L = a.L, A = a.A;
            for (int idx = threadIdx.x; idx < nrows * a.B; idx += blockDim.x) {
                const int gg = idx % nrows, k2 = idx / nrows;
                const float2 v = smem[gg * LS + lds_pad(k2)];
                const int f = k1_0 + gg + A * k2;
                if constexpr (SK == RS_FOBINS)
                    StFoBins{(float2*)out, L}(f, v);
                else if constexpr (SK == RS_CX)
                    StCx<false>{(float2*)out}(f, v);
                else if constexpr (SK == RS_FODEMOD)
                    StFoDemod{out, L, a.tmod}(f, v);
                else
                    StCx<true>{(float2*)out}(f, v);
            }
        } else if constexpr (KIND == RK_CONV) {
            const BotConv bot{(const float2*)fs_kern(a, t) + (long long)k1 * a.B, a.conj_kern != 0, fs_kern_shared(a),
                              a.spec ? (float2*)(a.spec + t * a.spec_st) + (long long)k1 * a.B : nullptr};
            const StRowTw st{row_out, a.tfs + (long long)k1 * a.B};
            if (single) {
                pf_single_conv<E>(a.c, pf, lt, bot, st);
                if (PERSIST && next < ntiles)
                    load_first<E>(a.c, pf, ltn, T, row_ld(tn, k1n));
            } else {
                first_to_tile<E>(a.c, pf, tl, lt, T);
                if (PERSIST && next < ntiles)
                    load_first<E>(a.c, pf, ltn, T, row_ld(tn, k1n));
                conv_tail<E>(a.c, tl, lt, T, st, bot);
            }
        } else {
            if (single) {
                pf_single_kern<E>(a.c, pf, lt, row_out);
                if (PERSIST && next < ntiles)
                    load_first<E>(a.c, pf, ltn, T, row_ld(tn, k1n));
            } else {
                first_to_tile<E>(a.c, pf, tl, lt, T);
                if (PERSIST && next < ntiles)
                    load_first<E>(a.c, pf, ltn, T, row_ld(tn, k1n));
                kern_tail<E>(a.c, tl, lt, T, row_out);
            }
        }
        __syncthreads();
        tile = next;
    }
}

template <class F>
static hipError_t do_launch(F fn, F fn_persist, const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    const long long tiles = a.ntrans * a.groups;
    if (tiles <= 0)
        return hipSuccess;
    const bool persist = persistent_mode();
    if (persist)
        fn = fn_persist;
    if (sh.lds_bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, sh.lds_bytes);
    const long long grid = persist ? persistent_grid((const void*)fn, sh.threads, sh.lds_bytes, tiles) : tiles;
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(sh.threads), sh.lds_bytes, st, a);
    return hipGetLastError();
}

// Chains with a compile-time form (the 240 x 256 split of 61440: columns
// [16,15], rows [16,16]) use it unless the kernel selection asks for the
// runtime-chain kernels.

template <int LK, int SK>
static hipError_t cols(const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    const bool fix = !persistent_mode() && !generic_forced() && Fix2<16, 15>::matches(a.c, a.T);
    count_launch(fix);
    if (fix)
        return do_launch(k_cols<LK, SK, 16, false, 16, 15>, k_cols<LK, SK, 16, true>, a, sh, st);
    return do_launch(k_cols<LK, SK, 16, false>, k_cols<LK, SK, 16, true>, a, sh, st);
}

template <int KIND, int SK>
static hipError_t rows(const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if constexpr (KIND != RK_CONV) {  // convolutions: see launch_rows
        const bool fix = !persistent_mode() && !generic_forced() && Fix2<16, 16>::matches(a.c, a.T);
        count_launch(fix);
        if (fix)
            return do_launch(k_rows<KIND, SK, 16, false, 16, 16>, k_rows<KIND, SK, 16, true>, a, sh, st);
    } else {
        count_launch(false);
    }
    return do_launch(k_rows<KIND, SK, 16, false>, k_rows<KIND, SK, 16, true>, a, sh, st);
}

hipError_t launch_cols(int load, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E != 16)
        return hipErrorInvalidValue;
    if (store == CS_WORKTW) {
        switch (load) {
        case CL_FOMOD: return cols<CL_FOMOD, CS_WORKTW>(a, sh, st);
        case CL_FOMODC: return cols<CL_FOMODC, CS_WORKTW>(a, sh, st);
        case CL_FOBINS: return cols<CL_FOBINS, CS_WORKTW>(a, sh, st);
        case CL_CX: return cols<CL_CX, CS_WORKTW>(a, sh, st);
        case CL_CXCONJ: return cols<CL_CXCONJ, CS_WORKTW>(a, sh, st);
        }
    } else if (load == CL_CX) {
        if (store == CS_FODEMOD)
            return cols<CL_CX, CS_FODEMOD>(a, sh, st);
        if (store == CS_CXCONJ)
            return cols<CL_CX, CS_CXCONJ>(a, sh, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rows(int kind, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E != 16)
        return hipErrorInvalidValue;
    if (kind == RK_STOCK) {
        switch (store) {
        case RS_FOBINS: return rows<RK_STOCK, RS_FOBINS>(a, sh, st);
        case RS_CX: return rows<RK_STOCK, RS_CX>(a, sh, st);
        case RS_FODEMOD: return rows<RK_STOCK, RS_FODEMOD>(a, sh, st);
        case RS_CXCONJ: return rows<RK_STOCK, RS_CXCONJ>(a, sh, st);
        }
        return hipErrorInvalidValue;
    }
    if (kind == RK_CONV) {
        if (!persistent_mode() && !generic_forced() && Fix2<16, 16>::matches(a.c, a.T)) {
            count_launch(true);
            const int bot = (a.conj_kern ? 1 : 0) | (!a.kern_p && a.kern_st == 0 ? 2 : 0);
            auto fp = k_rows<RK_CONV, RS_NONE, 16, true>;
            switch (bot) {
            case 0: return do_launch(k_rows<RK_CONV, RS_NONE, 16, false, 16, 16, 0>, fp, a, sh, st);
            case 1: return do_launch(k_rows<RK_CONV, RS_NONE, 16, false, 16, 16, 1>, fp, a, sh, st);
            case 2: return do_launch(k_rows<RK_CONV, RS_NONE, 16, false, 16, 16, 2>, fp, a, sh, st);
            default: return do_launch(k_rows<RK_CONV, RS_NONE, 16, false, 16, 16, 3>, fp, a, sh, st);
            }
        }
        return rows<RK_CONV, RS_NONE>(a, sh, st);
    }
    if (kind == RK_KERN)
        return rows<RK_KERN, RS_NONE>(a, sh, st);
    return hipErrorInvalidValue;
}

}  // namespace fsa
