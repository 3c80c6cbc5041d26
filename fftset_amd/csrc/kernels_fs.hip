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

template <int LK, int SK, int E>
__global__ void __launch_bounds__(1024) k_cols(const FsArgs a)
{
    extern __shared__ float2 smem[];
    const int C = a.C;
    const int g = threadIdx.x % C;
    const int lt0 = threadIdx.x / C;
    const long long t = blockIdx.x / a.groups;
    const int grp = (int)(blockIdx.x - t * a.groups);
    const int n2 = grp * C + g;
    const bool valid = lt0 < a.T && n2 < a.B && t < a.ntrans;
    const int lt = valid ? lt0 : (1 << 28);
    const long long t0 = valid ? t : 0;
    const int n2c = valid ? n2 : 0;
    const TileInter tl{smem + g, C};
    const float* in = a.in + t0 * a.in_st;
    float* out = a.out + t0 * a.out_st;
    const int L = a.L, B = a.B;

    auto run = [&](const auto& ld) {
        if constexpr (SK == CS_WORKTW)
            run_stockham<E>(a.c, tl, lt, a.T, ld, StWorkTw{(float2*)out, B, n2c, a.tfs});
        else if constexpr (SK == CS_FODEMOD)
            run_stockham<E>(a.c, tl, lt, a.T, ld, StCol<StFoDemod>{{out, L, a.tmod}, B, n2c});
        else
            run_stockham<E>(a.c, tl, lt, a.T, ld, StCol<StCx<true>>{{(float2*)out}, B, n2c});
    };
    if constexpr (LK == CL_FOMOD)
        run(LdCol<LdFoMod>{{in, L, a.tmod}, B, n2c});
    else if constexpr (LK == CL_FOBINS)
        run(LdCol<LdFoBins>{{(const float2*)in, L}, B, n2c});
    else if constexpr (LK == CL_CX)
        run(LdCol<LdCx<false>>{{(const float2*)in}, B, n2c});
    else
        run(LdCol<LdCx<true>>{{(const float2*)in}, B, n2c});
}

template <int KIND, int SK, int E>
__global__ void __launch_bounds__(1024) k_rows(const FsArgs a)
{
    extern __shared__ float2 smem[];
    const int T = a.T, C = a.C;
    const int g = threadIdx.x / T;
    const long long t = blockIdx.x / a.groups;
    const int grp = (int)(blockIdx.x - t * a.groups);
    const int k1 = grp * C + g;
    const bool valid = g < C && k1 < a.A && t < a.ntrans;
    const int lt = valid ? (int)(threadIdx.x - g * T) : (1 << 28);
    const long long t0 = t < a.ntrans ? t : 0;
    const int k1c = valid ? k1 : 0;
    const int LS = lds_len(a.B);
    const TileContig tl{smem + (valid ? g : 0) * LS};
    const float2* row_in = (const float2*)(a.in + t0 * a.in_st) + (long long)k1c * a.B;
    float2* row_out = (float2*)(a.out + t0 * a.out_st) + (long long)k1c * a.B;

    if constexpr (KIND == RK_STOCK) {
        run_stockham_to_tile<E>(a.c, tl, lt, T, LdCx<false>{row_in});
        __syncthreads();
        // transposed store: lanes run across the rows of this workgroup so
        // consecutive lanes write consecutive outputs f = k1 + A*k2
        const int k1_0 = grp * C;
        const int nrows = min(C, a.A - k1_0);
        if (t >= a.ntrans || nrows <= 0)
            return;
        float* out = a.out + t * a.out_st;
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
        const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st) + (long long)k1c * a.B, a.conj_kern != 0};
        run_conv<E>(a.c, tl, lt, T, LdCx<false>{row_in}, StRowTw{row_out, a.tfs + (long long)k1c * a.B}, bot);
    } else {
        run_kern<E>(a.c, tl, lt, T, LdCx<false>{row_in}, row_out);
    }
}

template <class F>
static hipError_t do_launch(F fn, const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    const long long grid = a.ntrans * a.groups;
    if (grid <= 0)
        return hipSuccess;
    if (sh.lds_bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, sh.lds_bytes);
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(sh.threads), sh.lds_bytes, st, a);
    return hipGetLastError();
}

hipError_t launch_cols(int load, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E != 16)
        return hipErrorInvalidValue;
    if (store == CS_WORKTW) {
        switch (load) {
        case CL_FOMOD: return do_launch(k_cols<CL_FOMOD, CS_WORKTW, 16>, a, sh, st);
        case CL_FOBINS: return do_launch(k_cols<CL_FOBINS, CS_WORKTW, 16>, a, sh, st);
        case CL_CX: return do_launch(k_cols<CL_CX, CS_WORKTW, 16>, a, sh, st);
        case CL_CXCONJ: return do_launch(k_cols<CL_CXCONJ, CS_WORKTW, 16>, a, sh, st);
        }
    } else if (load == CL_CX) {
        if (store == CS_FODEMOD)
            return do_launch(k_cols<CL_CX, CS_FODEMOD, 16>, a, sh, st);
        if (store == CS_CXCONJ)
            return do_launch(k_cols<CL_CX, CS_CXCONJ, 16>, a, sh, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rows(int kind, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E != 16)
        return hipErrorInvalidValue;
    if (kind == RK_STOCK) {
        switch (store) {
        case RS_FOBINS: return do_launch(k_rows<RK_STOCK, RS_FOBINS, 16>, a, sh, st);
        case RS_CX: return do_launch(k_rows<RK_STOCK, RS_CX, 16>, a, sh, st);
        case RS_FODEMOD: return do_launch(k_rows<RK_STOCK, RS_FODEMOD, 16>, a, sh, st);
        case RS_CXCONJ: return do_launch(k_rows<RK_STOCK, RS_CXCONJ, 16>, a, sh, st);
        }
        return hipErrorInvalidValue;
    }
    if (kind == RK_CONV)
        return do_launch(k_rows<RK_CONV, RS_NONE, 16>, a, sh, st);
    if (kind == RK_KERN)
        return do_launch(k_rows<RK_KERN, RS_NONE, 16>, a, sh, st);
    return hipErrorInvalidValue;
}

}  // namespace fsa
