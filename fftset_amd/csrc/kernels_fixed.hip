// kernels_fixed.hip -- LDS-resident transforms specialised at compile time
// for the radix chains of the hot sizes.
//
// The generic kernel (lds_impl.h) reads its chain at run time: every pass
// dispatches on its radix and computes its strides and LDS addresses from
// runtime values. Here a three-pass chain [R0, R1, R2] and the lanes per
// transform are template constants, so pass strides, butterfly bounds and
// most index arithmetic fold at compile time. Same passes, same
// twiddle tables, same data order (kernel spectra are interchangeable with
// the generic kernel's), fewer instructions.
#include "lds_impl.h"

namespace fsa {

// ONE: one transform per workgroup of exactly T lanes (G = 1). The
// transform index is then blockIdx.x, uniform across the workgroup, so its
// signal, output and kernel bases live in SGPRs and every access is that
// base plus a 32-bit lane offset; without it each lane carries 64-bit
// addresses.
template <int MOD, int OP, int R0, int R1, int R2, int T, bool ONE>
__global__ void __launch_bounds__(1024) k_fix(const LdsArgs a)
{
    constexpr int E = 16;
    constexpr int n = R0 * R1 * R2;
    constexpr int m0 = n / R0, s1 = R0, m1 = n / (R0 * R1), s2 = R0 * R1;
    constexpr int NB0 = (E / R0) > 0 ? E / R0 : 1;
    constexpr int NB1 = (E / R1) > 0 ? E / R1 : 1;
    constexpr int NB2 = (E / R2) > 0 ? E / R2 : 1;
    constexpr int kIdle = 1 << 28;
    extern __shared__ float2 smem[];
    using Io = LdsIo<MOD, OP>;
    const int G = ONE ? 1 : a.G;
    const int g = ONE ? 0 : threadIdx.x / T;
    const int lt0 = threadIdx.x - g * T;
    const bool in_tile = ONE || g < G;
    const TileContig tl{smem + (in_tile ? g : 0) * lds_len(n)};
    const long long tr = (long long)blockIdx.x * G + g;
    const bool ok = ONE || (in_tile && tr < a.ntrans);  // ONE: the grid is ntrans workgroups
    const int lt = ok ? lt0 : kIdle;
    const long long t0 = ok ? tr : 0;
    const float2* tw0 = a.c.p[0].tw;
    const float2* tw1 = a.c.p[1].tw;

    pass_first<R0, NB0>(tl, lt, T, m0, tw0, Io::ld(a, t0));
    __syncthreads();
    if constexpr (OP == OP_FWD || OP == OP_INV) {
        stock_mid<R1, NB1>(tl, lt, T, m1, s1, tw1);
        __syncthreads();
        stock_last<R2, NB2>(tl, lt, T, s2, Io::st(a, t0));
    } else if constexpr (OP == OP_KERN) {
        dif_mid<R1, NB1>(tl, lt, T, m1, s1, tw1);
        __syncthreads();
        dif_bottom<R2, NB2>(tl, lt, T, s2, BotKernOut{(float2*)(a.out + t0 * a.out_st)});
    } else {
        dif_mid<R1, NB1>(tl, lt, T, m1, s1, tw1);
        __syncthreads();
        const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN, a.kern_st == 0,
                          a.spec ? (float2*)(a.spec + t0 * a.spec_st) : nullptr};
        dif_bottom<R2, NB2>(tl, lt, T, s2, bot);
        __syncthreads();
        dit_mid<R1, NB1>(tl, lt, T, m1, s1, tw1);
        __syncthreads();
        dit_last<R0, NB0>(tl, lt, T, m0, tw0, Io::st(a, t0));
    }
}

template <int MOD, int OP, int R0, int R1, int R2, int T>
static hipError_t launch_fix(const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    const long long tiles = (a.ntrans + sh.G - 1) / sh.G;
    if (tiles <= 0)
        return hipSuccess;
    if (tiles > 0x7fffffffLL)
        return hipErrorInvalidValue;
    auto fn = (sh.G == 1 && sh.threads == T) ? k_fix<MOD, OP, R0, R1, R2, T, true> : k_fix<MOD, OP, R0, R1, R2, T, false>;
    if (sh.lds_bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, sh.lds_bytes);
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(fn, dim3((unsigned)tiles), dim3(sh.threads), sh.lds_bytes, st, a);
    return hipGetLastError();
}

static bool chain_is(const ChainD& c, int r0, int r1, int r2)
{
    return c.np == 3 && c.p[0].R == r0 && c.p[1].R == r1 && c.p[2].R == r2;
}

// The specialisations: COMPLEX 3840 = [15,16,16] forward / inverse,
// FREQOFFSETREAL 4096 = [16,16,16] and 2048 = [8,16,16] get_kernel /
// convolution / signal adjoint. hipErrorNotSupported: use the generic kernel.
hipError_t launch_lds_fixed(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E != 16)
        return hipErrorNotSupported;
    if (mod == MOD_CX && chain_is(a.c, 15, 16, 16) && sh.T == 256) {
        if (op == OP_FWD)
            return launch_fix<MOD_CX, OP_FWD, 15, 16, 16, 256>(a, sh, st);
        if (op == OP_INV)
            return launch_fix<MOD_CX, OP_INV, 15, 16, 16, 256>(a, sh, st);
    }
    if (mod == MOD_FO && chain_is(a.c, 16, 16, 16) && sh.T == 256) {
        if (op == OP_KERN)
            return launch_fix<MOD_FO, OP_KERN, 16, 16, 16, 256>(a, sh, st);
        if (op == OP_CONV)
            return launch_fix<MOD_FO, OP_CONV, 16, 16, 16, 256>(a, sh, st);
        if (op == OP_GRAD_IN)
            return launch_fix<MOD_FO, OP_GRAD_IN, 16, 16, 16, 256>(a, sh, st);
    }
    if (mod == MOD_FO && chain_is(a.c, 8, 16, 16) && sh.T == 128) {
        if (op == OP_KERN)
            return launch_fix<MOD_FO, OP_KERN, 8, 16, 16, 128>(a, sh, st);
        if (op == OP_CONV)
            return launch_fix<MOD_FO, OP_CONV, 8, 16, 16, 128>(a, sh, st);
        if (op == OP_GRAD_IN)
            return launch_fix<MOD_FO, OP_GRAD_IN, 8, 16, 16, 128>(a, sh, st);
    }
    return hipErrorNotSupported;
}

}  // namespace fsa
