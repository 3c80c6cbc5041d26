// engine.h -- LDS-resident mixed-radix passes for gfx950 (wave64).
//
// A "tile" is one transform of length n living in LDS; T threads (lanes)
// work on it. Two pass families are provided, both following the index maps
// of the reference's CPU passes so the data orders (and the opaque kernel
// spectrum layout) are well defined:
//
//  * Stockham DIF (self-sorting; natural order in, natural order out) for the
//    forward and inverse transforms -- src/fftset_vec.c:54-63, 90-106:
//      out[b m + j + k m s] = W_{rm}^{jk} DFT_r(in[b r m + j + q m])[k]
//    Each pass stages its butterflies in VGPRs, so one LDS array suffices.
//  * In-place DIF then in-place DIT (no reordering) for convolution, with the
//    pointwise product fused at the bottom -- src/fftset_vec.c:46-53, 66-89,
//    1161-1220: positions b r m + j + q m are read and written back.
//
// The first pass reads global memory through a Load functor and the last
// pass writes global memory through a Store functor, so modulation, bin
// interleave and conjugation are fused into the passes that touch HBM.
//
// Twiddles: each pass of sub-length N = R*m reads its own FP32 table laid
// out tw[(k-1)*m + j] = W_N^{jk}, so lanes with consecutive j read
// consecutive addresses. Tables are generated in FP64 on the GPU and shared
// between plans by key.
#pragma once

#include "dev_math.h"

namespace fsa {

// ------------------------------------------------------------ LDS addressing
// One padding slot every 16 complex values breaks the power-of-two strides of
// the late passes across the 64 LDS banks.
__device__ __forceinline__ int lds_pad(int e) { return e + (e >> 4); }
__host__ __device__ constexpr int lds_len(int n) { return n + (n >> 4) + 1; }

struct TileContig {
    float2* base;
    __device__ __forceinline__ float2& at(int e) const { return base[lds_pad(e)]; }
};

// Element e of tile g at lds[e * C + g] (base = lds + g): tiles interleaved
// so that lanes running across tiles touch consecutive banks.
struct TileInter {
    float2* base;
    int C;
    __device__ __forceinline__ float2& at(int e) const { return base[e * C]; }
};

// ------------------------------------------------------------ pass functions
// NB = butterflies a thread may own in one pass; u = lt + q*T for q < NB.
// The host sizes T so that T*NB >= butterflies of every pass.

// FREQOFFSETREAL modulation folded into the first pass (and demodulation
// into the last DIT pass). In the first pass (radix R, m = n/R, s = 1) the
// inputs of butterfly u are e = u + t m, and W_{4n}^e = W_{4n}^u W_{4R}^t:
// the second factor is a compile-time root per input, the first is common
// to the butterfly's DFT and merges with its output twiddles W_n^{uk} into
// one table tfo[k m + u] = W_{4n}^{u (1 + 4k)}, k < R. The closing DIT pass
// mirrors it: inputs times tfo[t m + u], outputs times W_{4R}^k, which is
// the demodulation W_{4n}^f of output f = u + k m. A functor opts in with
// kFoFold = true and provides raw(e) = x[e] - i x[n+e] (load) or
// put_demod(f, v) (store) and tfo; its operator() keeps the full modulation
// for every other pass that reaches it.
template <class F, class = void>
struct FoFold : std::false_type {};
template <class F>
struct FoFold<F, std::void_t<decltype(F::kFoFold)>> : std::bool_constant<F::kFoFold> {};

// First pass (s = 1, m = n/R): reads x[j + q m] through `ld`, writes the
// tile at j + k m. Shared by Stockham and in-place DIF.
template <int R, int NB, class Tile, class Load>
__device__ __forceinline__ void pass_first(const Tile& tl, int lt, int T, int m, const float2* __restrict__ tw,
                                           const Load& ld)
{
    float2 v[NB][R];
    if constexpr (FoFold<Load>::value) {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int u = lt + q * T;
            if (u < m)
                static_for<R>([&](auto Tt) {
                    constexpr int t = decltype(Tt)::value;
                    v[q][t] = rootmul<t, 4 * R>(ld.raw(u + t * m));
                });
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int u = lt + q * T;
            if (u < m) {
                DFT<R>::run(v[q]);
#pragma unroll
                for (int k = 0; k < R; ++k) tl.at(u + k * m) = cmul(v[q][k], at2(ld.tfo, k * m, u));
            }
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < m) {
#pragma unroll
            for (int t = 0; t < R; ++t) v[q][t] = ld(u + t * m);
        }
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < m) {
            DFT<R>::run(v[q]);
#pragma unroll
            for (int k = 1; k < R; ++k) v[q][k] = cmul(v[q][k], at2(tw, (k - 1) * m, u));
#pragma unroll
            for (int k = 0; k < R; ++k) tl.at(u + k * m) = v[q][k];
        }
    }
}

// The first pass split in two so a persistent workgroup can prefetch the
// next transform's inputs: first_load() fills pf[q*R + t] with x[u + t m]
// (u = lt + q T); first_compute() runs the butterflies from pf into the tile.
template <int R, int NB, class Load>
__device__ __forceinline__ void first_load(float2* pf, int lt, int T, int m, const Load& ld)
{
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < m) {
#pragma unroll
            for (int t = 0; t < R; ++t) pf[q * R + t] = ld(u + t * m);
        }
    }
}

template <int R, int NB, class Tile>
__device__ __forceinline__ void first_compute(float2* pf, const Tile& tl, int lt, int T, int m,
                                              const float2* __restrict__ tw)
{
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < m) {
            float2* v = pf + q * R;
            DFT<R>::run(v);
#pragma unroll
            for (int k = 1; k < R; ++k) v[k] = cmul(v[k], at2(tw, (k - 1) * m, u));
#pragma unroll
            for (int k = 0; k < R; ++k) tl.at(u + k * m) = v[k];
        }
    }
}

// Middle Stockham pass, LDS to LDS through registers.
template <int R, int NB, class Tile>
__device__ __forceinline__ void stock_mid(const Tile& tl, int lt, int T, int m, int s, const float2* __restrict__ tw)
{
    float2 v[NB][R];
    int bq[NB], jq[NB];
    const int nbf = m * s;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < nbf) {
            const int b = u / m, j = u - b * m;
            bq[q] = b;
            jq[q] = j;
            const int base = b * R * m + j;
#pragma unroll
            for (int t = 0; t < R; ++t) v[q][t] = tl.at(base + t * m);
            DFT<R>::run(v[q]);
            if (m > 1) {
#pragma unroll
                for (int k = 1; k < R; ++k) v[q][k] = cmul(v[q][k], at2(tw, (k - 1) * m, j));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < nbf) {
            const int o = bq[q] * m + jq[q];
#pragma unroll
            for (int k = 0; k < R; ++k) tl.at(o + k * m * s) = v[q][k];
        }
    }
}

// Last Stockham pass (m = 1, s = n/R): tile -> `st(f, value)` in natural
// order f = b + k s.
template <int R, int NB, class Tile, class Store>
__device__ __forceinline__ void stock_last(const Tile& tl, int lt, int T, int s, const Store& st)
{
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = lt + q * T;
        if (b < s) {
            float2 v[R];
#pragma unroll
            for (int t = 0; t < R; ++t) v[t] = tl.at(b * R + t);
            DFT<R>::run(v);
#pragma unroll
            for (int k = 0; k < R; ++k) st(b + k * s, v[k]);
        }
    }
}

// Whole transform in one butterfly (n == R): global -> registers -> global.
template <int R, class Load, class Store>
__device__ __forceinline__ void single_pass(int lt, const Load& ld, const Store& st)
{
    if (lt == 0) {
        float2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = ld(t);
        DFT<R>::run(v);
#pragma unroll
        for (int k = 0; k < R; ++k) st(k, v[k]);
    }
}

// Middle in-place DIF pass: positions b R m + j + q m, twiddle after the DFT.
template <int R, int NB, class Tile>
__device__ __forceinline__ void dif_mid(const Tile& tl, int lt, int T, int m, int s, const float2* __restrict__ tw)
{
    const int nbf = m * s;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < nbf) {
            const int b = u / m, j = u - b * m;
            const int base = b * R * m + j;
            float2 v[R];
#pragma unroll
            for (int t = 0; t < R; ++t) v[t] = tl.at(base + t * m);
            DFT<R>::run(v);
#pragma unroll
            for (int k = 1; k < R; ++k) v[k] = cmul(v[k], at2(tw, (k - 1) * m, j));
#pragma unroll
            for (int k = 0; k < R; ++k) tl.at(base + k * m) = v[k];
        }
    }
}

// Middle in-place DIT pass: forward twiddle before a forward DFT (the
// conjugate-data inverse of src/fftset_vec.c:78-89).
template <int R, int NB, class Tile>
__device__ __forceinline__ void dit_mid(const Tile& tl, int lt, int T, int m, int s, const float2* __restrict__ tw)
{
    const int nbf = m * s;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < nbf) {
            const int b = u / m, j = u - b * m;
            const int base = b * R * m + j;
            float2 v[R];
            v[0] = tl.at(base);
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = cmul(tl.at(base + t * m), at2(tw, (t - 1) * m, j));
            DFT<R>::run(v);
#pragma unroll
            for (int k = 0; k < R; ++k) tl.at(base + k * m) = v[k];
        }
    }
}

// Last DIT pass (s = 1, m = n/R): tile -> store in natural order j + k m.
template <int R, int NB, class Tile, class Store>
__device__ __forceinline__ void dit_last(const Tile& tl, int lt, int T, int m, const float2* __restrict__ tw,
                                         const Store& st)
{
    if constexpr (FoFold<Store>::value) {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int u = lt + q * T;
            if (u < m) {
                float2 v[R];
#pragma unroll
                for (int t = 0; t < R; ++t) v[t] = cmul(tl.at(u + t * m), at2(st.tfo, t * m, u));
                DFT<R>::run(v);
                static_for<R>([&](auto Kk) {
                    constexpr int k = decltype(Kk)::value;
                    st.put_demod(u + k * m, rootmul<k, 4 * R>(v[k]));
                });
            }
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int u = lt + q * T;
        if (u < m) {
            float2 v[R];
            v[0] = tl.at(u);
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = cmul(tl.at(u + t * m), at2(tw, (t - 1) * m, u));
            DFT<R>::run(v);
#pragma unroll
            for (int k = 0; k < R; ++k) st(u + k * m, v[k]);
        }
    }
}

// Bottom of the in-place DIF chain (m = 1, s = n/R): the R outputs of block
// b are handed to `bot(b, v)` while still in registers. The kernel spectrum
// layout used by the bottom operators is slot-major: value k of block b at
// kern[k * s + b] (coalesced across lanes).
template <int R, int NB, class Tile, class Bottom>
__device__ __forceinline__ void dif_bottom(const Tile& tl, int lt, int T, int s, const Bottom& bot)
{
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = lt + q * T;
        if (b < s) {
            float2 v[R];
#pragma unroll
            for (int t = 0; t < R; ++t) v[t] = tl.at(b * R + t);
            DFT<R>::run(v);
            bot.template apply<R>(tl, b, s, v);
        }
    }
}

// ------------------------------------------------------------ bottom operators
// Convolution: w <- conj(w * K) (src/fftset_vec.c:847-867), then the first
// DIT pass (m = 1: a plain forward DFT) in registers, back to the tile.
// spec (optional): the signal's spectrum, written before the product in the
// kernel-spectrum layout (what get_kernel of the signal gives), for the
// kernel gradient of a later backward pass.
struct BotConv {
    const float2* __restrict__ kern;
    bool conj_kern;      // true: w <- conj(w * conj(K)) (adjoint w.r.t. the signal)
    bool shared = false;  // one spectrum for every transform: cached loads
    float2* __restrict__ spec = nullptr;
    template <int R, class Tile>
    __device__ __forceinline__ void apply(const Tile& tl, int b, int s, float2* v) const
    {
        if (spec) {
#pragma unroll
            for (int k = 0; k < R; ++k) stg(spec + k * s + b, v[k]);
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float2 K = shared ? kern[k * s + b] : ldg(kern + k * s + b);
            v[k] = cconj(conj_kern ? cmulc(v[k], K) : cmul(v[k], K));
        }
        DFT<R>::run(v);
#pragma unroll
        for (int t = 0; t < R; ++t) tl.at(b * R + t) = v[t];
    }
};

// BotConv with the adjoint and shared-spectrum choices fixed at compile time.
template <bool CONJ, bool SHARED>
struct BotConvC {
    const float2* __restrict__ kern;
    float2* __restrict__ spec = nullptr;  // optional signal spectrum (see BotConv)
    template <int R, class Tile>
    __device__ __forceinline__ void apply(const Tile& tl, int b, int s, float2* v) const
    {
        if (spec) {
#pragma unroll
            for (int k = 0; k < R; ++k) stg(spec + k * s + b, v[k]);
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float2 K = SHARED ? kern[k * s + b] : ldg(kern + k * s + b);
            v[k] = cconj(CONJ ? cmulc(v[k], K) : cmul(v[k], K));
        }
        DFT<R>::run(v);
#pragma unroll
        for (int t = 0; t < R; ++t) tl.at(b * R + t) = v[t];
    }
};

// Kernel spectrum: written out slot-major.
struct BotKernOut {
    float2* __restrict__ out;
    template <int R, class Tile>
    __device__ __forceinline__ void apply(const Tile&, int b, int s, float2* v) const
    {
#pragma unroll
        for (int k = 0; k < R; ++k) stg(out + k * s + b, v[k]);
    }
};

// ------------------------------------------------------------ I/O functors
// FREQOFFSETREAL first pass (src/fftset_mod_freqoffsetreal.c:917-924):
// z[e] = (x[e] - i x[L+e]) e^{-i pi e / (2L)}; tmod[e] = W_{4L}^e.
struct LdFoMod {
    const float* __restrict__ x;
    int L;
    const float2* __restrict__ tmod;
    __device__ __forceinline__ float2 operator()(int e) const
    {
        const float re = ldg(x + e), im = ldg(x + L + e);
        const float2 w = at32(tmod, e);
        return make_float2(re * w.x + im * w.y, re * w.y - im * w.x);
    }
};

// LdFoMod with the modulation folded into the first pass (see FoFold):
// tfo[k m + u] = W_{4L}^{u (1 + 4k)} for the plan's first pass.
struct LdFoFold {
    static constexpr bool kFoFold = true;
    const float* __restrict__ x;
    int L;
    const float2* __restrict__ tmod;
    const float2* __restrict__ tfo;
    __device__ __forceinline__ float2 raw(int e) const { return make_float2(ldg(x + e), -ldg(x + L + e)); }
    __device__ __forceinline__ float2 operator()(int e) const { return LdFoMod{x, L, tmod}(e); }
};

// FREQOFFSETREAL inverse input (src/fftset_mod_freqoffsetreal.c:985-998):
// w[e] = conj X[2e] for e < ceil(L/2), else X[2(L-1-e)+1].
struct LdFoBins {
    const float2* __restrict__ in;
    int L;
    __device__ __forceinline__ float2 operator()(int e) const
    {
        return (e < ((L + 1) >> 1)) ? cconj(ldg(in + 2 * e)) : ldg(in + 2 * (L - 1 - e) + 1);
    }
};

template <bool CONJ>
struct LdCx {
    const float2* __restrict__ in;
    __device__ __forceinline__ float2 operator()(int e) const
    {
        const float2 v = ldg(in + e);
        return CONJ ? cconj(v) : v;
    }
};

// FREQOFFSETREAL forward output (src/fftset_mod_freqoffsetreal.c:958-971):
// bin 2p <- Z[p], bin 2p+1 <- conj Z[L-1-p].
struct StFoBins {
    float2* __restrict__ out;
    int L;
    __device__ __forceinline__ void operator()(int f, float2 v) const
    {
        if (f < ((L + 1) >> 1))
            stg(out + 2 * f, v);
        else
            stg(out + 2 * (L - 1 - f) + 1, cconj(v));
    }
};

// FREQOFFSETREAL last pass of the inverse and of the convolution
// (src/fftset_mod_freqoffsetreal.c:926-933, 1002-1009):
// y[f] = Re(v W), y[L+f] = Im(v W), W = e^{-i pi f / (2L)}.
struct StFoDemod {
    float* __restrict__ y;
    int L;
    const float2* __restrict__ tmod;
    __device__ __forceinline__ void operator()(int f, float2 v) const
    {
        const float2 w = at32(tmod, f);
        stg(y + f, v.x * w.x - v.y * w.y);
        stg(y + L + f, v.x * w.y + v.y * w.x);
    }
};

// StFoDemod with the demodulation folded into the closing DIT pass (see
// FoFold): put_demod(f, v) receives v already times W_{4L}^f.
struct StFoFold {
    static constexpr bool kFoFold = true;
    float* __restrict__ y;
    int L;
    const float2* __restrict__ tmod;
    const float2* __restrict__ tfo;
    __device__ __forceinline__ void put_demod(int f, float2 v) const
    {
        stg(y + f, v.x);
        stg(y + L + f, v.y);
    }
    __device__ __forceinline__ void operator()(int f, float2 v) const { StFoDemod{y, L, tmod}(f, v); }
};

template <bool CONJ>
struct StCx {
    float2* __restrict__ out;
    __device__ __forceinline__ void operator()(int f, float2 v) const { stg(out + f, CONJ ? cconj(v) : v); }
};

// Writes into the tile (natural order) instead of global memory.
template <class Tile>
struct TileStore {
    Tile tl;
    __device__ __forceinline__ void operator()(int f, float2 v) const { tl.at(f) = v; }
};

// ------------------------------------------------- four-step index views
// A transform of n = A*B is viewed as B columns of length A (element n1 of
// column n2 is element B*n1 + n2) and A rows of length B.
template <class Ld>
struct LdCol {
    Ld ld;
    int B, n2;
    __device__ __forceinline__ float2 operator()(int n1) const { return ld(B * n1 + n2); }
};

// FREQOFFSETREAL column load with the modulation split (four-step column
// step): element n1 of column n2 is z[e] = (x[e] - i x[L+e]) W_{4L}^e with
// e = n1 B + n2 and W_{4L}^e = W_{4A}^{n1} W_{4L}^{n2}. The column factor is
// constant over the column's DFT and goes into the output twiddles; only
// trow[n1] = W_{4A}^{n1} is read here.
struct LdColFoMod {
    const float* __restrict__ x;
    int L, B, n2;
    const float2* __restrict__ trow;
    __device__ __forceinline__ float2 operator()(int n1) const
    {
        const int e = B * n1 + n2;
        const float re = ldg(x + e), im = ldg(x + L + e);
        const float2 w = trow[n1];
        return make_float2(re * w.x + im * w.y, re * w.y - im * w.x);
    }
};

template <class St>
struct StCol {
    St st;
    int B, n2;
    __device__ __forceinline__ void operator()(int n1, float2 v) const { st(B * n1 + n2, v); }
};

// Column-step output: Y[k1][n2] W_n^{n2 k1} into the row-major work layout;
// tfs[k1*B + n2] = W_n^{n2 k1} has the same layout.
struct StWorkTw {
    float2* __restrict__ work;
    int B, n2;
    const float2* __restrict__ tfs;
    __device__ __forceinline__ void operator()(int k1, float2 v) const
    {
        stg(work + k1 * B + n2, cmul(v, tfs[k1 * B + n2]));
    }
};

// Row-step output of the convolution: element n2 of row k1, times
// W_n^{n2 k1} (trow = tfs + k1*B), back into the row.
struct StRowTw {
    float2* __restrict__ row;
    const float2* __restrict__ trow;
    __device__ __forceinline__ void operator()(int n2, float2 v) const { stg(row + n2, cmul(v, trow[n2])); }
};

}  // namespace fsa
