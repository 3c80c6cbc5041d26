// kernels_refpass.hip -- the reference's pass structure on the GPU, one
// launch per sweep: the per-size-launch baseline that the fused batch is
// measured against.
//
// The reference runs every transform as a sequence of full sweeps over
// memory (src/fftset_vec.c:1161-1249; src/fftset_mod_freqoffsetreal.c:
// 36-470, 884-1010; src/fftset_mod_cplx.c:33-318):
//   * a first pass that modulates (FREQOFFSETREAL) and splits the transform
//     into vw interleaved sub-transforms by radix-4 (and radix-2 at vector
//     width 8) DIF butterflies, or a modulation-only sweep at width 1;
//   * the inner passes of its planner's chain for length L/vw, each a sweep:
//     Stockham (ping-pong) for forward / inverse, in-place DIF, a separate
//     product sweep and in-place DIT for the convolution;
//   * a last pass: the lane / bin reorder (forward), or the radix-4 (and
//     radix-2) DIT combine with demodulation or conjugation (inverse,
//     convolution).
// Here each of those sweeps is one kernel launch over every transform of a
// batch class, one thread per butterfly (or element), with the data in HBM
// between sweeps. The vw sub-transforms live as contiguous lanes of a
// transform's work row: lane slot q at [q M, (q+1) M), M = L / vw. The CPU
// ping-pong copies (memcpy when a Stockham chain ends in the other buffer)
// have no counterpart: a GPU pass reads and writes where it needs to.
#include "kernels.h"

namespace fsa {

// Per-thread element index over a class: transform t and element e of it.
__device__ __forceinline__ long long rp_gid() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }

// Modulated FREQOFFSETREAL input element e (src/fftset_mod_freqoffsetreal.c:
// 917-924): (x[e] - i x[L+e]) W_{4L}^e.
__device__ __forceinline__ float2 rp_ld_fo(const float* x, int L, const float2* tmod, int e)
{
    return LdFoMod{x, L, tmod}(e);
}

// In-place DIF pass over every transform: positions b R m + j + q m of the
// work row (s = L / (R m) blocks), twiddle W_{Rm}^{jk} after the DFT. SRC:
// RP_WORK reads the work row; RP_IN_FO / RP_IN_CX (first pass, s = 1) read
// the transform's input at the natural index, modulated for FREQOFFSETREAL.
template <int R, int SRC>
__global__ void __launch_bounds__(256) k_rp_dif(const RpArgs a)
{
    const long long per = a.L / R;
    const long long g = rp_gid();
    if (g >= a.ntrans * per)
        return;
    const long long t = g / per;
    const int u = (int)(g - t * per);
    const int b = u / a.m, j = u - b * a.m;
    const int base = b * R * a.m + j;
    float2* w = a.work + t * a.L;
    float2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int e = base + q * a.m;
        if constexpr (SRC == RP_IN_FO)
            v[q] = rp_ld_fo(a.in[t], a.L, a.tmod, e);
        else if constexpr (SRC == RP_IN_CX)
            v[q] = ldg((const float2*)a.in[t] + e);
        else
            v[q] = w[e];
    }
    DFT<R>::run(v);
    if (a.m > 1) {
#pragma unroll
        for (int k = 1; k < R; ++k) v[k] = cmul(v[k], a.tw[(k - 1) * a.m + j]);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) w[base + k * a.m] = v[k];
}

// In-place DIT pass (forward twiddles before the DFT: the conjugate-data
// inverse, src/fftset_vec.c:78-89). DST: RP_WORK writes back in place;
// RP_OUT_FO / RP_OUT_CX (last pass, s = 1) write the transform's output at
// the natural index, demodulated (FREQOFFSETREAL) or conjugated (COMPLEX).
template <int R, int DST>
__global__ void __launch_bounds__(256) k_rp_dit(const RpArgs a)
{
    const long long per = a.L / R;
    const long long g = rp_gid();
    if (g >= a.ntrans * per)
        return;
    const long long t = g / per;
    const int u = (int)(g - t * per);
    const int b = u / a.m, j = u - b * a.m;
    const int base = b * R * a.m + j;
    float2* w = a.work + t * a.L;
    float2 v[R];
    v[0] = w[base];
#pragma unroll
    for (int q = 1; q < R; ++q) v[q] = a.m > 1 ? cmul(w[base + q * a.m], a.tw[(q - 1) * a.m + j]) : w[base + q * a.m];
    DFT<R>::run(v);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int e = base + k * a.m;
        if constexpr (DST == RP_OUT_FO)
            StFoDemod{a.out[t], a.L, a.tmod}(e, v[k]);
        else if constexpr (DST == RP_OUT_CX)
            StCx<true>{(float2*)a.out[t]}(e, v[k]);
        else
            w[e] = v[k];
    }
}

// Stockham pass inside every lane (src/fftset_vec.c:90-106): lane slot q of
// transform t, butterfly (b, j): in[b R m + j + q' m] -> out[b m + j + k m s]
// with W_{Rm}^{jk}. SRC RP_IN_CX: the first pass of a width-1 COMPLEX
// forward reads the input directly; DST RP_OUT_CX: its last pass writes the
// output directly (no reorder sweep at width 1).
template <int R, int SRC, int DST>
__global__ void __launch_bounds__(256) k_rp_stock(const RpArgs a)
{
    const int M = a.L / a.V;
    const long long per = (long long)a.V * (M / R);
    const long long g = rp_gid();
    if (g >= a.ntrans * per)
        return;
    const long long t = g / per;
    const int r = (int)(g - t * per);
    const int lane = r / (M / R), u = r - lane * (M / R);
    const int b = u / a.m, j = u - b * a.m;
    const int ib = b * R * a.m + j;
    const float2* src = SRC == RP_IN_CX ? (const float2*)a.in[t] : a.work + t * a.L + lane * M;
    float2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = SRC == RP_IN_CX ? ldg(src + ib + q * a.m) : src[ib + q * a.m];
    DFT<R>::run(v);
    if (a.m > 1) {
#pragma unroll
        for (int k = 1; k < R; ++k) v[k] = cmul(v[k], a.tw[(k - 1) * a.m + j]);
    }
    const int ob = b * a.m + j;
    const int step = a.m * a.s;
    if constexpr (DST == RP_OUT_CX) {
        float2* o = (float2*)a.out[t];
#pragma unroll
        for (int k = 0; k < R; ++k) stg(o + ob + k * step, v[k]);
    } else {
        float2* o = a.work2 + t * a.L + lane * M;
#pragma unroll
        for (int k = 0; k < R; ++k) o[ob + k * step] = v[k];
    }
}

// The convolution's product sweep (src/fftset_vec.c:847-867):
// w <- conj(w K), or conj(w conj(K)) for the signal adjoint.
__global__ void __launch_bounds__(256) k_rp_mulconj(const RpArgs a)
{
    const long long g = rp_gid();
    if (g >= a.ntrans * a.L)
        return;
    const float2 K = a.kref[g];
    a.work[g] = cconj(a.conj_kern ? cmulc(a.work[g], K) : cmul(a.work[g], K));
}

// Forward's last sweep: lane slot slot[f % V] holds bins f = V k' + l in
// natural k' order; write them in natural bin order (COMPLEX) or with the
// FREQOFFSETREAL bin interleave (src/fftset_mod_freqoffsetreal.c:520-531,
// 958-971; src/fftset_mod_cplx.c:181-210). One thread per output bin.
template <int MOD>
__global__ void __launch_bounds__(256) k_rp_inter(const RpArgs a)
{
    const long long g = rp_gid();
    if (g >= a.ntrans * a.L)
        return;
    const long long t = g / a.L;
    const int f = (int)(g - t * a.L);
    const int M = a.L / a.V;
    const int l = f % a.V, kk = f / a.V;
    const float2 v = a.work[t * a.L + a.slot[l] * M + kk];
    if constexpr (MOD == MOD_FO)
        StFoBins{(float2*)a.out[t], a.L}(f, v);
    else
        stg((float2*)a.out[t] + f, v);
}

// Inverse's first sweep: bins into lane slots, conjugated (the inverse runs
// forward passes on conjugate data): FREQOFFSETREAL bin de-interleave
// (src/fftset_mod_freqoffsetreal.c:546-557, 858-870, 985-998), COMPLEX conj
// (src/fftset_mod_cplx.c:221-227, 300-304). One thread per lane element.
template <int MOD>
__global__ void __launch_bounds__(256) k_rp_deinter(const RpArgs a)
{
    const long long g = rp_gid();
    if (g >= a.ntrans * a.L)
        return;
    const long long t = g / a.L;
    const int e = (int)(g - t * a.L);
    const int M = a.L / a.V;
    const int lane = e / M, kk = e - lane * M;
    const int f = a.V * kk + a.resid[lane];
    float2 v;
    if constexpr (MOD == MOD_FO)
        v = LdFoBins{(const float2*)a.in[t], a.L}(f);
    else
        v = cconj(ldg((const float2*)a.in[t] + f));
    a.work[t * a.L + e] = v;
}

// Set-up: the baseline's kernel spectra, gathered from the library's
// spectra: kref[t][p] = K_t[perm[p]].
__global__ void __launch_bounds__(256) k_rp_gather(const RpArgs a, const int* __restrict__ perm)
{
    const long long g = rp_gid();
    if (g >= a.ntrans * a.L)
        return;
    const long long t = g / a.L;
    const int p = (int)(g - t * a.L);
    a.work[g] = ((const float2*)a.in[t])[perm[p]];
}

// ---------------------------------------------------------------- launchers
static hipError_t rp_launch(void (*fn)(const RpArgs), const RpArgs& a, long long threads, hipStream_t st)
{
    if (threads <= 0)
        return hipSuccess;
    const long long blocks = (threads + 255) / 256;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidValue;
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

#define FSA_RP_RADIX(R, FN, ...)                        \
    switch (R) {                                        \
    case 1: fn = FN<1, __VA_ARGS__>; break;             \
    case 2: fn = FN<2, __VA_ARGS__>; break;             \
    case 3: fn = FN<3, __VA_ARGS__>; break;             \
    case 4: fn = FN<4, __VA_ARGS__>; break;             \
    case 5: fn = FN<5, __VA_ARGS__>; break;             \
    case 6: fn = FN<6, __VA_ARGS__>; break;             \
    case 8: fn = FN<8, __VA_ARGS__>; break;             \
    case 16: fn = FN<16, __VA_ARGS__>; break;           \
    default: return hipErrorInvalidValue;               \
    }

hipError_t launch_refpass(int kind, int R, int io, const RpArgs& a, hipStream_t st)
{
    void (*fn)(const RpArgs) = nullptr;
    switch (kind) {
    case RPK_DIF:
        if (io == RP_IN_FO) {
            FSA_RP_RADIX(R, k_rp_dif, RP_IN_FO)
        } else if (io == RP_IN_CX) {
            FSA_RP_RADIX(R, k_rp_dif, RP_IN_CX)
        } else {
            FSA_RP_RADIX(R, k_rp_dif, RP_WORK)
        }
        return rp_launch(fn, a, a.ntrans * (a.L / R), st);
    case RPK_DIT:
        if (io == RP_OUT_FO) {
            FSA_RP_RADIX(R, k_rp_dit, RP_OUT_FO)
        } else if (io == RP_OUT_CX) {
            FSA_RP_RADIX(R, k_rp_dit, RP_OUT_CX)
        } else {
            FSA_RP_RADIX(R, k_rp_dit, RP_WORK)
        }
        return rp_launch(fn, a, a.ntrans * (a.L / R), st);
    case RPK_STOCK:
        if (io == RP_IN_CX) {
            FSA_RP_RADIX(R, k_rp_stock, RP_IN_CX, RP_WORK)
        } else if (io == RP_OUT_CX) {
            FSA_RP_RADIX(R, k_rp_stock, RP_WORK, RP_OUT_CX)
        } else if (io == RP_IN_OUT_CX) {
            FSA_RP_RADIX(R, k_rp_stock, RP_IN_CX, RP_OUT_CX)
        } else {
            FSA_RP_RADIX(R, k_rp_stock, RP_WORK, RP_WORK)
        }
        return rp_launch(fn, a, a.ntrans * (a.L / R), st);
    case RPK_MULCONJ: return rp_launch(k_rp_mulconj, a, a.ntrans * a.L, st);
    case RPK_INTER: return rp_launch(io == RP_OUT_FO ? k_rp_inter<MOD_FO> : k_rp_inter<MOD_CX>, a, a.ntrans * a.L, st);
    case RPK_DEINTER:
        return rp_launch(io == RP_IN_FO ? k_rp_deinter<MOD_FO> : k_rp_deinter<MOD_CX>, a, a.ntrans * a.L, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_refpass_gather(const RpArgs& a, const int* perm, hipStream_t st)
{
    const long long n = a.ntrans * a.L;
    if (n <= 0)
        return hipSuccess;
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidValue;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_rp_gather, dim3((unsigned)blocks), dim3(256), 0, st, a, perm);
    return hipGetLastError();
}

}  // namespace fsa
