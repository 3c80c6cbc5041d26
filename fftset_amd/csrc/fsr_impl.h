// fsr_impl.h -- 61440-bin transforms with the whole transform resident on
// one CU (instantiated by kernels_fsr_cx.hip and kernels_fsr_fo.hip).
//
// n = A * B = 240 * 256. The two-launch four-step path (kernels_fs.hip)
// writes every transform's intermediate to HBM between its column and row
// launches: 2x (forward) or 2.33x (convolution) the algorithmic bytes. Here
// one 512-lane workgroup (one per CU: it takes all 160 KB of LDS and 256
// VGPRs per lane) keeps a whole transform (480 KB) on its CU. The 240 x 256
// intermediate Y[k1][n2] lives as
//   rows k1 = b + 16 k, k < 3 (48 rows):  LDS;
//   k = 3, 4 (32 rows):                   a 64 KB per-transform spill region,
//                                          written and read back by the same
//                                          workgroup (small enough for L2);
//   k >= 5 (160 rows):                    VGPRs, 80 complex values per lane
//                                          (lane (g, b) of a 32-column chunk).
// Columns: 8 chunks of 32 adjacent columns (chain [16, 15]) in a 60 KB LDS
//   tile over the LDS of rows 48..79; global loads in 256 B (COMPLEX) or
//   128 B (FREQOFFSETREAL) runs that start on a 128-byte line whatever the
//   input's alignment (the chunks start at the matching column); outputs
//   times W_n^{n2 k1} to their place above.
// Rows: 3 batches of 80 rows in LDS (chain [16, 16], in place; the spilled
//   rows return first):
//   forward / inverse: DIF, then the second pass's outputs go straight to
//     X[k1 + 240 k2], lanes across rows (512 B runs);
//   get_kernel: the DIF bottom writes the slot-major spectrum;
//   convolution: DIF -> product -> DIT -> W_n^{n2 k1}; the results swap
//     places with the next batch's VGPRs, and the last batch's rows 48..79
//     go to the spill region.
// Closing columns (convolution): chain [15, 16] on conjugate data, whose
//   first pass reads exactly one lane's rows b + 16 k from VGPRs, LDS and
//   the spill; outputs in runs of 32 columns.
// Index maps, twiddle tables and the kernel-spectrum layout (value k of
// block b of row k1 at k1*256 + 16 k + b) are the two-launch path's, so
// spectra are interchangeable between the two paths. Barriers order a
// workgroup's own steps; no workgroup waits on another.
#pragma once

#include "engine.h"
#include "fsr.h"

namespace fsa {
namespace fsr {

constexpr int kA = 240, kB = 256, kN = kA * kB;
constexpr int kThreads = 512;
constexpr int kTile = 48 * kB;           // column tile: complex offset in LDS (batch rows 48..79)
constexpr int kLds = 80 * kB * 8;        // one batch of 80 rows: 163840 bytes

// The thread index, opaque to the compiler: lane-dependent addresses
// derived from it are computed where they are used instead of being hoisted
// to the top of the kernel (dozens of address VGPRs live across every
// phase). The mask restores the known range.
__device__ __forceinline__ int fresh_tid()
{
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t & (kThreads - 1);
}

// Column lanes: lane lt (block b = lt) of a 32-column chunk. The chunks
// start at column s (0 <= s < 32) of the transform, chosen so that every
// chunk's run of 32 columns in a row begins on a 128-byte line of the
// input: a run that straddled two lines would have each boundary line
// fetched once for either chunk (the loads are nontemporal). In chunk c the
// lane holds column 32 cc(c) + g, with g = (lane + s) & 31 and d the carry.
struct ColLane {
    int g, lt, d, lo;  // lo = 256 lt + g
    __device__ __forceinline__ explicit ColLane(int s)
    {
        const int t = fresh_tid();
        const int x = (t & 31) + s;
        g = x & 31;
        d = x >> 5;
        lt = t >> 5;
        lo = kB * lt + g;
    }
    __device__ __forceinline__ int cc(int c) const { return (c + d) & 7; }
    // 256 lt + the lane's column in chunk c
    __device__ __forceinline__ int at(int c) const { return lo + 32 * cc(c); }
};

// LDS position of element p of batch row r, without padding (a batch of 80
// rows fills all 160 KB). ACROSS (forward / inverse, whose row lanes run
// across rows): p ^ (r & 31). Otherwise (row lanes run across the 16
// butterflies of a row, 4 rows per wave): the low digit of p XOR its high
// digit, and bit 4 XOR (r & 1). Either keeps every access pattern below at
// two lanes per bank pair, the minimum for 64-bit accesses of 64 lanes.
// Each pattern is a lane base and a constant XOR or offset per element, so
// an access costs one VALU instruction at most:
//   pa(r, j) ^ at_a(s): element 16 s + j   (radix-16 pass over stride 16)
//   pb(r, b) ^ s:       element 16 b + s   (radix-16 pass over stride 1)
//   pc(r, g) . at_c(c): element 32 c + g   (column chunk c, column g; c may
//                       be a lane value)
template <bool ACROSS>
struct Swz;

template <>
struct Swz<true> {
    static __device__ __forceinline__ int pos(int r, int p) { return r * kB + (p ^ (r & 31)); }
    static __device__ __forceinline__ int pa(int r, int j) { return r * kB + (j ^ (r & 31)); }
    static __device__ __forceinline__ int at_a(int base, int s) { return base ^ (16 * s); }
    static __device__ __forceinline__ int pb(int r, int b) { return r * kB + ((16 * b) ^ (r & 31)); }
    static __device__ __forceinline__ int pc(int r, int g) { return r * kB + (g ^ (r & 31)); }
    static __device__ __forceinline__ int at_c(int base, int c) { return base + 32 * c; }
};

template <>
struct Swz<false> {
    static __device__ __forceinline__ int pos(int r, int p) { return r * kB + (p ^ ((p >> 4) & 15) ^ ((r & 1) << 4)); }
    static __device__ __forceinline__ int pa(int r, int j) { return r * kB + (j | ((r & 1) << 4)); }
    static __device__ __forceinline__ int at_a(int base, int s) { return base ^ (17 * s); }
    static __device__ __forceinline__ int pb(int r, int b) { return r * kB + ((17 * b) ^ ((r & 1) << 4)); }
    static __device__ __forceinline__ int pc(int r, int g) { return r * kB + ((17 * (g >> 4)) ^ ((r & 1) << 4) ^ (g & 15)); }
    static __device__ __forceinline__ int at_c(int base, int c) { return base ^ (34 * c); }
};

// Powers W^{u m} (m = 0..15) of one per-lane root, from six stored ones:
// W^{u m} = W^{u (m & 3)} W^{4 u (m >> 2)}, at most one multiply each and
// two roundings. lo[0] is 1, or a common factor of every power (then every
// power carries it).
struct Pow16 {
    float2 lo[4];  // W^{u b} (times the common factor), b = 0..3
    float2 hi[3];  // W^{4 u a}, a = 1..3
    bool one;      // lo[0] is exactly 1
    // from a pass table laid out tw[(k-1) m + u] = W^{u k}
    __device__ __forceinline__ void load(const float2* tw, int m, int u)
    {
        one = true;
        lo[0] = make_float2(1.f, 0.f);
#pragma unroll
        for (int b = 1; b < 4; ++b) lo[b] = at2(tw, (b - 1) * m, u);
#pragma unroll
        for (int a = 1; a < 4; ++a) hi[a - 1] = at2(tw, (4 * a - 1) * m, u);
    }
    // v times W^{u k} (k a compile-time constant after unrolling)
    __device__ __forceinline__ float2 apply(float2 v, int k) const
    {
        const int b = k & 3, a = k >> 2;
        if (a == 0)
            return (b == 0 && one) ? v : cmul(v, lo[b]);
        return (b == 0 && one) ? cmul(v, hi[a - 1]) : cmul(v, cmul(lo[b], hi[a - 1]));
    }
};

// Column inputs for pass-1 lane u: elements n1 = u + 15 q of column n2, at
// 3840 q + lo (lo = 256 u + n2). COMPLEX: the value (conjugated for the
// inverse). FREQOFFSETREAL: the real pair (x[e], x[n + e]).
template <int MOD, int OPK>
__device__ __forceinline__ void col_load(float2* P, const float* in, int lo)
{
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        if constexpr (MOD == MOD_FO) {
            P[q] = make_float2(ldg(ub(in, 3840 * q, lo)), ldg(ub(in, kN + 3840 * q, lo)));
        } else {
            const float2 v = ldg(ub((const float2*)in, 3840 * q, lo));
            P[q] = OPK == FSR_INV ? cconj(v) : v;
        }
    }
}

// The four-step twiddle of row b + 16 k (k = 0..15) of one column (or, in
// the row phase, element j + 16 k of one row): W_n^{x (b + 16 k)} =
// W_n^{x b} W_n^{16 x (k & 3)} W_n^{64 x (k >> 2)} from seven table entries
// (e[0] may carry a modulation factor, as tfs_fo's rows do).
struct Tw7 {
    float2 e[7];  // W^{x b}; W^{16 x b'}, b' = 1..3; W^{64 x a}, a = 1..3
    // e[b'] <- e[0] e[b']: then every twiddle is at most one product away
    __device__ __forceinline__ void prep()
    {
#pragma unroll
        for (int b = 1; b < 4; ++b) e[b] = cmul(e[0], e[b]);
    }
    // v times the twiddle of row (element) k, after prep()
    __device__ __forceinline__ float2 apply(float2 v, int k) const
    {
        const int b = k & 3, a = k >> 2;
        const float2 w = a ? cmul(e[b], e[3 + a]) : e[b];
        return cmul(v, w);
    }
};

template <int MOD, int OPK, int BOT>
__global__ void __launch_bounds__(kThreads) k_fsr(const FsrArgs a)
{
    constexpr bool kConv = OPK == FSR_CONV, kKern = OPK == FSR_KERN;
    constexpr bool kAcross = OPK == FSR_FWD || OPK == FSR_INV;
    constexpr bool kConj = (BOT & 1) != 0, kShared = (BOT & 2) != 0;
    using S = Swz<kAcross>;
    extern __shared__ float2 lds[];
    float2* tile = lds + kTile;
    const int tid = threadIdx.x;
    const long long t = blockIdx.x;
    const float* in = a.in_p ? a.in_p[t] : a.in + t * a.in_st;
    float* out = a.out_p ? a.out_p[t] : a.out + t * a.out_st;
    float2* spill = a.spill_p ? (float2*)a.spill_p[t] : a.spill + t * a.spill_st;
    // first column of the chunks (ColLane): in + s elements is 128-byte aligned
    const unsigned long long neg = 0ull - reinterpret_cast<unsigned long long>(in);
    const int s = MOD == MOD_FO ? (int)((neg >> 2) & 31) : (int)((neg >> 3) & 15);

    // R[c][k - 5]: row b + 16 k of column 32 c + g (k = 5..14) after the
    // columns; for a convolution, after the row batches R[c][k] holds the
    // result rows k = 0..9.
    float2 R[8][10];

    // ---------------------------------------------------------------- columns
    // Pass 1 of lane u = lt (< 15): inputs n1 = u + 15 q, output twiddles
    // W_240^{u k}. FREQOFFSETREAL: the modulation's W_{4A}^{n1} =
    // W_960^u W_64^q, a compile-time root per input and a lane factor folded
    // into the output twiddles.
    Pow16 tw1;
    {
        const ColLane cl(s);
        const int u = cl.lt < 15 ? cl.lt : 14;
        tw1.load(fresh(a.tw_col), 15, u);
        if constexpr (MOD == MOD_FO) {
            const float2 au = fresh(a.trow)[u];
            tw1.one = false;
#pragma unroll
            for (int b = 0; b < 4; ++b) tw1.lo[b] = b ? cmul(tw1.lo[b], au) : au;
        }
    }
    float2 P[16];  // this chunk's pass-1 inputs (loaded one chunk ahead)
    {
        const ColLane cl(s);
        if (cl.lt < 15)
            col_load<MOD, OPK>(P, in, cl.at(0));
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const ColLane cl(s);
        const int g = cl.g, lt = cl.lt, cc = cl.cc(c), lo = cl.lo + 32 * cc, n2 = 32 * cc + g;
        // pass-2 twiddles of this chunk, issued before the next chunk's loads
        Tw7 e;
        e.e[0] = *ub(a.tfs_in, 0, lo);
#pragma unroll
        for (int b = 1; b < 4; ++b) e.e[b] = *ub(a.tfs, 4096 * b, n2);
#pragma unroll
        for (int m = 1; m < 4; ++m) e.e[3 + m] = *ub(a.tfs, 16384 * m, n2);
        if (lt < 15) {  // first pass, radix 16 over n1 = lt + 15 q
            float2 v[16];
            static_for<16>([&](auto Q) {
                constexpr int q = decltype(Q)::value;
                if constexpr (MOD == MOD_FO)
                    v[q] = rootmul<q, 64>(make_float2(P[q].x, -P[q].y));
                else
                    v[q] = P[q];
            });
            DFT<16>::run(v);
#pragma unroll
            for (int k = 0; k < 16; ++k) tile[(lt + 15 * k) * 32 + g] = tw1.apply(v[k], k);
            if (c < 7)
                col_load<MOD, OPK>(P, in, cl.at(c + 1));  // in flight over pass 2
        }
        __syncthreads();
        {  // second pass, radix 15: outputs k1 = lt + 16 k
            float2 v[15];
#pragma unroll
            for (int q = 0; q < 15; ++q) v[q] = tile[(15 * lt + q) * 32 + g];
            DFT<15>::run(v);
            e.prep();
#pragma unroll
            for (int k = 0; k < 15; ++k) {
                const float2 w = e.apply(v[k], k);
                if (k < 3)
                    lds[S::at_c(S::pc(lt + 16 * k, g), cc)] = w;
                else if (k < 5)
                    *ub(spill, 4096 * (k - 3), lo) = w;
                else {
                    R[c][k - 5] = w;
                    pin(R[c][k - 5]);
                }
            }
        }
        __syncthreads();
    }
    // the spilled rows 48..79 back into LDS, over the column tile
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
        const int idx = tid + kThreads * i;
        lds[S::pos(48 + (idx >> 8), idx & 255)] = spill[idx];
    }
    __syncthreads();

    // ------------------------------------------------------------------- rows
    // Row lanes: forward / inverse: butterfly j = tid >> 5 of rows
    // (tid & 31) + 32 i; otherwise butterfly j = tid & 15 of rows
    // (tid >> 4) + 32 i. Either way j is fixed per lane, so the radix-16
    // passes' twiddles W_256^{j k} come from one lane's six stored powers.
    const int rj = kAcross ? tid >> 5 : tid & 15;
    const int rr = kAcross ? tid & 31 : tid >> 4;
    constexpr int rrow = 32;                   // row step between a lane's butterflies
    const int rn = rr + 2 * rrow < 80 ? 3 : 2;  // butterflies per lane and pass
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q > 0 && !kConv) {  // this batch's rows from VGPRs
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const ColLane cl(s);
#pragma unroll
                for (int kk = 0; kk < 5; ++kk)
                    lds[S::at_c(S::pc(cl.lt + 16 * kk, cl.g), cl.cc(c))] = R[c][5 * (q - 1) + kk];
            }
            __syncthreads();
        }
        // DIF first pass, radix 16 over elements j + 16 s, in place
#pragma unroll 1
        for (int i = 0; i < rn; ++i) {
            const int r = rr + rrow * i, j = rj;
            Pow16 tw;
            tw.load(fresh(a.tw_row), 16, j);
            const int pa = S::pa(r, j);
            float2 v[16];
#pragma unroll
            for (int s = 0; s < 16; ++s) v[s] = lds[S::at_a(pa, s)];
            DFT<16>::run(v);
            const int pw = S::pa(fresh_int(r), fresh_int(j));
#pragma unroll
            for (int k = 0; k < 16; ++k) lds[S::at_a(pw, k)] = tw.apply(v[k], k);
        }
        __syncthreads();
        if constexpr (kAcross) {
            // second pass, radix 16 over elements 16 b + s: outputs
            // k2 = b + 16 k of row k1 = 80 q + r, stored transposed. The
            // lanes' rows are rotated by s2 so that each run of 32 rows
            // starts on a 128-byte line of the output.
            const int s2 = (int)((0ull - reinterpret_cast<unsigned long long>(out)) >> 3) & 15;
            float2* o = (float2*)out + 80 * q;
#pragma unroll 1
            for (int i = 0; i < rn; ++i) {
                const int r0 = rr + rrow * i + s2, r = r0 >= 80 ? r0 - 80 : r0, b = rj;
                const int pb = S::pb(r, b);
                float2 v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) v[s] = lds[pb ^ s];
                DFT<16>::run(v);
                const int lo2 = r + kA * b;
#pragma unroll
                for (int k = 0; k < 16; ++k) stg(ub(o, 3840 * k, lo2), OPK == FSR_INV ? cconj(v[k]) : v[k]);
            }
        } else if constexpr (kKern) {
            // DIF bottom: value k of block b of row k1 at k1*256 + 16 k + b
            float2* K = (float2*)out + 80 * q * kB;
#pragma unroll 1
            for (int i = 0; i < rn; ++i) {
                const int r = rr + rrow * i, b = rj;
                const int pb = S::pb(r, b);
                float2 v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) v[s] = lds[pb ^ s];
                DFT<16>::run(v);
                const int lo2 = kB * r + b;
#pragma unroll
                for (int k = 0; k < 16; ++k) stg(ub(K, 16 * k, lo2), v[k]);
            }
        } else {
            // DIF bottom, product w <- conj(w K) (conj(w conj K) for the
            // adjoint), then the DIT's first pass (a plain DFT), in place
            const float* kt = a.kern_p ? a.kern_p[t] : a.kern + t * a.kern_st;
            const float2* K = (const float2*)kt + 80 * q * kB;
            float2* Sp = a.spec ? (float2*)(a.spec + t * a.spec_st) + 80 * q * kB : nullptr;
#pragma unroll 1
            for (int i = 0; i < rn; ++i) {
                const int r = rr + rrow * i, b = rj;
                const int lo2 = kB * r + b;
                float2 kv[16];  // the spectrum, issued before the LDS reads
#pragma unroll
                for (int k = 0; k < 16; ++k) kv[k] = kShared ? *ub(K, 16 * k, lo2) : ldg(ub(K, 16 * k, lo2));
                const int pb = S::pb(r, b);
                float2 v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) v[s] = lds[pb ^ s];
                DFT<16>::run(v);
                if (Sp) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) stg(ub(Sp, 16 * k, lo2), v[k]);
                }
#pragma unroll
                for (int k = 0; k < 16; ++k) v[k] = cconj(kConj ? cmulc(v[k], kv[k]) : cmul(v[k], kv[k]));
                DFT<16>::run(v);
                const int pw = S::pb(fresh_int(r), fresh_int(b));
#pragma unroll
                for (int s = 0; s < 16; ++s) lds[pw ^ s] = v[s];
            }
            __syncthreads();
            // DIT last pass (forward twiddles before a forward DFT, on
            // conjugate data), then W_n^{n2 k1} for the closing columns
            const float2* tf = a.tfs + 80 * q * kB;
#pragma unroll 1
            for (int i = 0; i < rn; ++i) {
                const int r = rr + rrow * i, j = rj;
                const int lo2 = kB * r + j;
                Pow16 tw;  // W_256^{j k}, as in the first pass
                tw.load(fresh(a.tw_row), 16, j);
                Tw7 e;  // W_n^{(j + 16 k) k1}, k1 = 80 q + r
                e.e[0] = *ub(tf, 0, lo2);
#pragma unroll
                for (int b = 1; b < 4; ++b) e.e[b] = *ub(tf, 16 * b, lo2 - j);
#pragma unroll
                for (int m = 1; m < 4; ++m) e.e[3 + m] = *ub(tf, 64 * m, lo2 - j);
                const int pa = S::pa(r, j);
                float2 v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) v[s] = lds[S::at_a(pa, s)];
#pragma unroll
                for (int s = 1; s < 16; ++s) v[s] = tw.apply(v[s], s);
                DFT<16>::run(v);
                e.prep();
                const int pw = S::pa(fresh_int(r), fresh_int(j));
#pragma unroll
                for (int k = 0; k < 16; ++k) lds[S::at_a(pw, k)] = e.apply(v[k], k);
            }
            __syncthreads();
            if (q < 2) {
                // results <-> the next batch's rows: each lane swaps its own
                // positions (row b + 16 kk of its columns)
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const ColLane cl(s);
#pragma unroll
                    for (int kk = 0; kk < 5; ++kk) {
                        const int pc = S::at_c(S::pc(cl.lt + 16 * kk, cl.g), cl.cc(c));
                        const float2 o = lds[pc];
                        lds[pc] = R[c][5 * q + kk];
                        R[c][5 * q + kk] = o;
                        pin(R[c][5 * q + kk]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
                // last batch: rows 48..79 (k = 13, 14) to the spill region
#pragma unroll 4
                for (int i = 0; i < 16; ++i) {
                    const int idx = tid + kThreads * i;
                    spill[idx] = lds[S::pos(48 + (idx >> 8), idx & 255)];
                }
            }
        }
        __syncthreads();
    }

    // -------------------------------------------------------- closing columns
    if constexpr (kConv) {
        // first pass of [15, 16] for block u = lt: twiddles W_240^{u k}
        Pow16 tw3;
        float2 sp[2];  // this chunk's two spilled rows (loaded one chunk ahead)
        {
            const ColLane cl(s);
            tw3.load(fresh(a.tw_close), 16, cl.lt);
#pragma unroll
            for (int k = 0; k < 2; ++k) sp[k] = *ub(spill, 4096 * k, cl.at(0));
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const ColLane cl(s);
            const int g = cl.g, lt = cl.lt, cc = cl.cc(c), lo = cl.lo + 32 * cc;
            float2 dm = make_float2(1.f, 0.f);  // FREQOFFSETREAL: W_{4n}^{256 lt + n2}
            if constexpr (MOD == MOD_FO)
                dm = *ub(a.tmod, 0, lo);
            {  // first pass of [15, 16], radix 15 over k1 = lt + 16 k
                float2 v[15];
#pragma unroll
                for (int k = 0; k < 10; ++k) v[k] = R[c][k];
#pragma unroll
                for (int k = 10; k < 13; ++k) v[k] = lds[S::at_c(S::pc(lt + 16 * (k - 10), g), cc)];
                v[13] = sp[0];
                v[14] = sp[1];
                if (c < 7) {
#pragma unroll
                    for (int k = 0; k < 2; ++k) sp[k] = *ub(spill, 4096 * k, cl.at(c + 1));
                }
                DFT<15>::run(v);
#pragma unroll
                for (int k = 0; k < 15; ++k) tile[(lt + 16 * k) * 32 + g] = tw3.apply(v[k], k);
            }
            __syncthreads();
            if (lt < 15) {  // second pass, radix 16: outputs n1 = lt + 15 k
                float2 v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) v[s] = tile[(16 * lt + s) * 32 + g];
                DFT<16>::run(v);
                if constexpr (MOD == MOD_FO) {
                    // y[f] = Re(v W_{4n}^f), y[n + f] = Im(v W_{4n}^f), with
                    // W_{4n}^f = W_{4n}^{256 lt + n2} W_64^k
                    static_for<16>([&](auto Kk) {
                        constexpr int k = decltype(Kk)::value;
                        const float2 w = rootmul<k, 64>(cmul(v[k], dm));
                        stg(ub(out, 3840 * k, lo), w.x);
                        stg(ub(out, kN + 3840 * k, lo), w.y);
                    });
                } else {
#pragma unroll
                    for (int k = 0; k < 16; ++k) stg(ub((float2*)out, 3840 * k, lo), cconj(v[k]));
                }
            }
            __syncthreads();
        }
    }
}

template <int MOD>
hipError_t launch_fsr_mod(int opk, const FsrArgs& a, hipStream_t st);

}  // namespace fsr
}  // namespace fsa
