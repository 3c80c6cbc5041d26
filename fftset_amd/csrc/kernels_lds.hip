// kernels_lds.hip -- single-kernel LDS-resident transforms (n fits in LDS)
// and the FP64 twiddle generator.
//
// One workgroup holds G transforms of one plan in LDS, T lanes each. The
// radix chain is a runtime descriptor; every pass dispatches on its radix to
// a fully unrolled register butterfly. Modulation, bin interleave,
// conjugation and the convolution product are fused into the passes that
// read or write HBM, so each transform moves exactly its input, output and
// (for convolution) kernel spectrum through HBM once.
#include <cstdlib>

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
                return LdFoMod{in, a.L, a.tmod};
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
        } else {
            if constexpr (MOD == MOD_FO)
                return StFoDemod{out, a.L, a.tmod};
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
            const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN};
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
                const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN};
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
                const BotConv bot{(const float2*)(a.kern + t0 * a.kern_st), OP == OP_GRAD_IN};
                conv_tail<E>(a.c, tl, lt, T, Io::st(a, t0), bot);
            }
        }
        __syncthreads();
        tile = next;
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
// Workgroups resident at once (occupancy x CUs), capped by the work.
long long persistent_grid(const void* fn, int threads, int lds_bytes, long long tiles)
{
    static thread_local int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds_bytes) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const long long resident = (long long)per_cu * cus * persistent_waves();
    return tiles < resident ? tiles : resident;
}

// Persistent launches (grid = resident set, register prefetch of the next
// tile) are opt-in through FFTSET_AMD_PERSIST=1.
bool persistent_mode()
{
    static int m = -1;
    if (m < 0) {
        const char* e = getenv("FFTSET_AMD_PERSIST");
        m = (e && atoi(e) > 0) ? 1 : 0;
    }
    return m == 1;
}

// Resident-grid multiplier (waves of workgroups per launch); 1 = exactly the
// resident set. Read once from FFTSET_AMD_GRID_WAVES.
int persistent_waves()
{
    static int w = 0;
    if (!w) {
        const char* e = getenv("FFTSET_AMD_GRID_WAVES");
        w = e ? atoi(e) : 1;
        if (w < 1)
            w = 1;
    }
    return w;
}
template <int MOD, int OP, int E>
static hipError_t launch_one(const LdsArgs& a, const ShapeD& sh, hipStream_t st)
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
