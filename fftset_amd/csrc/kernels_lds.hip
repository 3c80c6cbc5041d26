// kernels_lds.hip -- single-kernel LDS-resident transforms (n fits in LDS)
// and the FP64 twiddle generator.
//
// One workgroup holds G transforms of one plan in LDS, T lanes each. The
// radix chain is a runtime descriptor; every pass dispatches on its radix to
// a fully unrolled register butterfly. Modulation, bin interleave,
// conjugation and the convolution product are fused into the passes that
// read or write HBM, so each transform moves exactly its input, output and
// (for convolution) kernel spectrum through HBM once.
#include "drivers.h"
#include "kernels.h"

namespace fsa {

// ---------------------------------------------------------------- kernel
// E = 16 keeps a butterfly set within 128 VGPRs (1024-lane workgroups);
// E = 32 needs up to 256 and is launched with at most 512 lanes.
template <int E>
constexpr int lds_max_threads() { return E <= 16 ? 1024 : 512; }

template <int MOD, int OP, int E>
__global__ void __launch_bounds__(lds_max_threads<E>()) k_lds(const LdsArgs a)
{
    extern __shared__ float2 smem[];
    const int T = a.T;
    const int g = threadIdx.x / T;
    const long long tr = (long long)blockIdx.x * a.G + g;
    const bool valid = (g < a.G) && (tr < a.ntrans);
    const int lt = valid ? (int)(threadIdx.x - g * T) : (1 << 28);
    const long long t0 = valid ? tr : 0;
    const TileContig tl{smem + (valid ? g : 0) * lds_len(a.L)};
    const float* in = a.in + t0 * a.in_st;
    float* out = a.out + t0 * a.out_st;
    const int L = a.L;

    if constexpr (OP == OP_FWD) {
        if constexpr (MOD == MOD_FO)
            run_stockham<E>(a.c, tl, lt, T, LdFoMod{in, L, a.tmod}, StFoBins{(float2*)out, L});
        else
            run_stockham<E>(a.c, tl, lt, T, LdCx<false>{(const float2*)in}, StCx<false>{(float2*)out});
    } else if constexpr (OP == OP_INV) {
        if constexpr (MOD == MOD_FO)
            run_stockham<E>(a.c, tl, lt, T, LdFoBins{(const float2*)in, L}, StFoDemod{out, L, a.tmod});
        else
            run_stockham<E>(a.c, tl, lt, T, LdCx<true>{(const float2*)in}, StCx<true>{(float2*)out});
    } else if constexpr (OP == OP_KERN) {
        if constexpr (MOD == MOD_FO)
            run_kern<E>(a.c, tl, lt, T, LdFoMod{in, L, a.tmod}, (float2*)out);
        else
            run_kern<E>(a.c, tl, lt, T, LdCx<false>{(const float2*)in}, (float2*)out);
    } else {
        const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN};
        if constexpr (MOD == MOD_FO)
            run_conv<E>(a.c, tl, lt, T, LdFoMod{in, L, a.tmod}, StFoDemod{out, L, a.tmod}, bot);
        else
            run_conv<E>(a.c, tl, lt, T, LdCx<false>{(const float2*)in}, StCx<true>{(float2*)out}, bot);
    }
}

// ---------------------------------------------------------------- twiddles
// out[r*cols + c] = W_N^{x}, x = ((a0 + r) * c) mod N, W_N = exp(-2 pi i / N):
// evaluated in FP64 (exact integer reduction of the angle) and rounded once
// to FP32. This is the only FP64 arithmetic in the library.
__global__ void k_twiddle_fp64(float2* __restrict__ out, long long N, int a0, int rows, int cols)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)rows * cols)
        return;
    const long long r = idx / cols, c = idx - r * cols;
    long long x = ((a0 + r) % N) * (c % N) % N;
    if ((4 * x) % N == 0) {
        const int q = (int)((4 * x) / N);  // quarter turns are exact
        const float cr[4] = {1.f, 0.f, -1.f, 0.f};
        const float ci[4] = {0.f, -1.f, 0.f, 1.f};
        out[idx] = make_float2(cr[q], ci[q]);
        return;
    }
    if (2 * x > N)
        x -= N;  // angle in (-pi, pi]
    double sn, cs;
    sincospi(-2.0 * (double)x / (double)N, &sn, &cs);
    out[idx] = make_float2((float)cs, (float)sn);
}

hipError_t launch_twiddles(float2* out, long long N, int a0, int rows, int cols, hipStream_t st)
{
    const long long n = (long long)rows * cols;
    const int threads = 256;
    hipLaunchKernelGGL(k_twiddle_fp64, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0, st, out, N,
                       a0, rows, cols);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launcher
template <int MOD, int OP, int E>
static hipError_t launch_one(const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    const long long grid = (a.ntrans + sh.G - 1) / sh.G;
    if (grid <= 0)
        return hipSuccess;
    auto fn = k_lds<MOD, OP, E>;
    if (sh.lds_bytes > 64 * 1024) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, sh.lds_bytes);
    }
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(sh.threads), sh.lds_bytes, st, a);
    return hipGetLastError();
}

template <int MOD, int E>
static hipError_t launch_op(int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
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

hipError_t launch_lds(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E == 16)
        return mod == MOD_FO ? launch_op<MOD_FO, 16>(op, a, sh, st) : launch_op<MOD_CX, 16>(op, a, sh, st);
    if (sh.E == 32)
        return mod == MOD_FO ? launch_op<MOD_FO, 32>(op, a, sh, st) : launch_op<MOD_CX, 32>(op, a, sh, st);
    return hipErrorInvalidValue;
}

}  // namespace fsa
