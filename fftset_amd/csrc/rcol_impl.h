// rcol_impl.h -- four-step transforms n = A * 256 with the column step in
// registers (instantiated by kernels_rcol_cx.hip and kernels_rcol_fo.hip).
//
// The LDS engine keeps a whole transform in LDS: at n = 15360 that is
// 123 KB, so one workgroup holds a CU and its load, pass and store phases
// never overlap another workgroup's. Here one 256-lane workgroup holds a
// transform in VGPRs instead: lane n2 owns column n2 of the A x 256 view
// (element n1 of column n2 is element n2 + 256 n1), 2A VGPRs of data.
//   Columns: a length-A DFT per lane in registers (no LDS, no barrier),
//     inputs read as 256-lane rows (2 KB runs), outputs times W_n^{n2 k1}.
//   Rows: batches of 15 rows (A = 60; 9 for A = 18; all A rows for
//     A <= 16) through one LDS buffer (33 KB at A = 60), so that three
//     (A = 60) to five workgroups share a CU: the radix-16 DIF pass over
//     elements j + 16 s with twiddles W_256^{j k}, then the radix-16 pass
//     over 16 b + s, which yields bin k2 = b + 16 k of row k1, X[k1 + A k2].
//     COMPLEX forward / inverse keep those bins in VGPRs and write the
//     transform at the end in contiguous chunks staged through the buffer;
//     FREQOFFSETREAL forward / inverse store each bin through the bin
//     interleave; get_kernel stores the slot-major spectrum (value k of
//     block b of row k1 at k1*256 + 16 k + b); convolution multiplies by
//     the spectrum, runs the DIT back (conjugate data, forward twiddles) and
//     returns the rows to their column lanes.
//   Closing columns (convolution): a second length-A DFT per lane in
//     registers; the output goes out as 256-lane rows in natural order.
// This is the decomposition, the twiddle tables and the spectrum layout of
// the two-launch four-step path (kernels_fs.hip) for the same split, so
// results agree to rounding and spectra are interchangeable between the
// two paths. Barriers order a workgroup's own steps; no workgroup waits on
// another.
#pragma once

#include "engine.h"
#include "rcol.h"

namespace fsa {
namespace rcol {

constexpr int kB = 256;
constexpr int kThreads = 256;
// Rows per LDS batch: 15 for A = 60 (four batches); 9 for A = 18 (two);
// all of them for A <= 16.
template <int A>
constexpr int rows_of()
{
    return A % 15 == 0 ? 15 : A % 16 == 0 ? 16 : A % 9 == 0 ? 9 : A <= 16 ? A : 1;
}
// LDS buffers for the row batches: with one, a batch waits for the previous
// batch's last LDS reads before it writes (one more barrier per batch), and
// three workgroups share a CU (LDS and <= 168 VGPRs); with two, two do.
constexpr int kBufs = 1;
// Workgroups per CU (= waves per SIMD) the register budget is set for: a
// 60-value column leaves room for three; an 18-value one for five, or four
// for the convolution (its spectrum and the held rows: 128 VGPRs).
template <int A, bool CONV>
constexpr int waves_of()
{
    return A <= 18 ? (CONV ? 4 : 5) : kBufs == 1 ? 3 : 2;
}

// LDS position of element p of batch row r: one padding slot per 16
// elements, rows RS apart. Every access of a pass is then one lane base plus
// a compile-time offset (17 per step of the pass). The LDS serves a 64-bit
// access 16 lanes at a time over 16 bank pairs, so each group of 16 lanes
// must touch 16 different pairs:
//   column lanes writing / reading row r: p = n2, 16 consecutive lanes;
//   lanes along rows, (r, j) = (tid >> 4, tid & 15), over j + 16 s or
//     16 j + s: a group is one row, whatever RS (272);
//   lanes across rows, (r, j) = (tid & 15, tid >> 4), over j + 16 s or
//     16 j + s: a group is 16 rows, which RS = 273 = 1 (mod 16) spreads over
//     all pairs (and 273 = 17 (mod 32) keeps two such groups apart).
template <bool ACROSS, int kRows>
struct Lay {
    static constexpr int RS = ACROSS ? 273 : 272;
    static constexpr int kBuf = kRows * RS;  // complex values per LDS buffer
    // element p of row r
    static __device__ __forceinline__ int pos(int r, int p) { return r * RS + p + (p >> 4); }
    // element j + 16 s of row r (j < 16) is at pa(r, j) + 17 s
    static __device__ __forceinline__ int pa(int r, int j) { return r * RS + j; }
    // element 16 b + s of row r (s < 16) is at pb(r, b) + s
    static __device__ __forceinline__ int pb(int r, int b) { return r * RS + 17 * b; }
    static constexpr int kLds = kBufs * kBuf * 8;
};

// Row-pass twiddles W_256^{j k} (k = 1..15) of lane j from six entries of
// the pass table tw[(k-1)*16 + j]: W^{j k} = W^{j (k & 3)} W^{4 j (k >> 2)}.
struct RowTw {
    float2 lo[3], hi[3];  // W^{j b}, b = 1..3; W^{4 j a}, a = 1..3
    __device__ __forceinline__ void load(const float2* tw, int j)
    {
#pragma unroll
        for (int b = 1; b < 4; ++b) lo[b - 1] = at2(tw, (b - 1) * 16, j);
#pragma unroll
        for (int a = 1; a < 4; ++a) hi[a - 1] = at2(tw, (4 * a - 1) * 16, j);
    }
    // v times W^{j k} (k a compile-time constant after unrolling, k >= 1)
    __device__ __forceinline__ float2 apply(float2 v, int k) const
    {
        const int b = k & 3, a = k >> 2;
        if (a == 0)
            return cmul(v, lo[b - 1]);
        return b == 0 ? cmul(v, hi[a - 1]) : cmul(v, cmul(lo[b - 1], hi[a - 1]));
    }
};

template <int A, int MOD, int OPK, int BOT>
__global__ void __launch_bounds__(kThreads, (waves_of<A, OPK == RC_CONV>())) k_rcol(const RcolArgs a)
{
    constexpr int kRows = rows_of<A>();
    constexpr int N = A * kB, NQ = A / kRows;
    static_assert(A % kRows == 0, "whole row batches");
    constexpr bool kConv = OPK == RC_CONV, kKern = OPK == RC_KERN;
    constexpr bool kAcross = OPK == RC_FWD || OPK == RC_INV;
    constexpr bool kConj = (BOT & 1) != 0, kShared = (BOT & 2) != 0;
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const long long t = blockIdx.x;
    const float* in = a.in_p ? a.in_p[t] : a.in + t * a.in_st;
    float* out = a.out_p ? a.out_p[t] : a.out + t * a.out_st;
    const int n2 = tid;

    // ---------------------------------------------------------------- columns
    float2 v[A];
    static_for<A>([&](auto N1) {
        constexpr int n1 = decltype(N1)::value;
        const int e = n2 + kB * n1;
        if constexpr (MOD == MOD_FO && OPK == RC_INV) {
            v[n1] = LdFoBins{(const float2*)in, N}(e);
        } else if constexpr (MOD == MOD_FO) {
            // z[e] = (x[e] - i x[n+e]) W_{4n}^e, W_{4n}^e = W_{4A}^{n1} W_{4n}^{n2}:
            // the first factor here, the second in tfs_in
            const float re = ldg(ub(in, kB * n1, n2)), im = ldg(ub(in, N + kB * n1, n2));
            v[n1] = rootmul<n1, 4 * A>(make_float2(re, -im));
        } else {
            const float2 z = ldg(ub((const float2*)in, kB * n1, n2));
            v[n1] = OPK == RC_INV ? cconj(z) : z;
        }
    });
    DFT<A>::run(v);
    // output twiddles W^{n2 k1} = W^{n2 a} W^{4 n2 b} (k1 = a + 4 b) from
    // A/4 + 3 table entries read in one go: A separate reads in dependent
    // groups cost a round trip each. FREQOFFSETREAL's tfs_in row a carries
    // the modulation factor W_{4n}^{n2}; row 4 b of plain tfs does not.
    {
        constexpr bool kFold = MOD == MOD_FO && OPK != RC_INV;
        const float2* tin = fresh(kFold ? a.tfs_in : a.tfs);
        const float2* tpl = fresh(a.tfs);
        float2 lo[4], hi[(A + 3) / 4];
        static_for<4>([&](auto Aa) {
            constexpr int aa = decltype(Aa)::value;
            // (rows a < A of the A x 256 tables only)
            lo[aa] = (aa > 0 || kFold) && aa < A ? *ub(tin, aa * kB, n2) : make_float2(1.f, 0.f);
        });
        static_for<(A + 3) / 4>([&](auto Bb) {
            constexpr int bb = decltype(Bb)::value;
            hi[bb] = bb > 0 ? *ub(tpl, 4 * bb * kB, n2) : make_float2(1.f, 0.f);
        });
        static_for<A>([&](auto K1) {
            constexpr int k1 = decltype(K1)::value, aa = k1 & 3, bb = k1 >> 2;
            if constexpr (bb > 0)
                v[k1] = cmul(v[k1], (aa > 0 || kFold) ? cmul(lo[aa], hi[bb]) : hi[bb]);
            else if constexpr (aa > 0 || kFold)
                v[k1] = cmul(v[k1], lo[aa]);
            pin(v[k1]);  // the product itself stays live, not its two factors
        });
    }

    // ------------------------------------------------------------------- rows
    // COMPLEX forward / inverse keep each batch's results in VGPRs (the
    // batch's inputs have left them) and write the transform at the end in
    // contiguous chunks staged through the LDS buffer: stored straight from
    // the second pass, a batch's bins are runs of only kRows values.
    constexpr bool kStage = kAcross && MOD == MOD_CX;
    float2 res[kStage ? NQ : 1][16];
    static_for<NQ>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        using S = Lay<kAcross, kRows>;
        float2* buf = lds + (q % kBufs) * S::kBuf;
        // this batch's rows into LDS: one buffer is free once every lane has
        // read the previous batch; of two, the other buffer's last readers
        // have passed the previous batch's first barrier
        if constexpr (kBufs == 1 && q > 0)
            __syncthreads();
        const int n2q = fresh_int(n2);  // lane values opaque per batch: no
                                        // addresses kept live across batches
        static_for<kRows>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            buf[S::pos(r, n2q)] = v[kRows * q + r];
        });
        __syncthreads();
        {  // DIF first pass, radix 16 over elements j + 16 s, in place
            const int tq = fresh_int(tid);
            const int r = kAcross ? tq & 15 : tq >> 4, j = kAcross ? tq >> 4 : tq & 15;
            if (r < kRows) {
                float2 w[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) w[s] = buf[S::pa(r, j) + 17 * s];
                DFT<16>::run(w);
                RowTw tw;
                tw.load(fresh(a.tw_row), j);
#pragma unroll
                for (int k = 1; k < 16; ++k) w[k] = tw.apply(w[k], k);
                const int rw = fresh_int(r), jw = fresh_int(j);
#pragma unroll
                for (int k = 0; k < 16; ++k) buf[S::pa(rw, jw) + 17 * k] = w[k];
            }
        }
        __syncthreads();
        if constexpr (kAcross) {
            // second pass over 16 b + s: bin k2 = b + 16 k of row k1 to
            // X[k1 + A k2], 15 rows (120 B) per run
            const int tq = fresh_int(tid), r = tq & 15, b = tq >> 4;
            if (r < kRows) {
                float2 w[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) w[s] = buf[S::pb(r, b) + s];
                DFT<16>::run(w);
                if constexpr (kStage) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        res[kStage ? q : 0][k] = OPK == RC_INV ? cconj(w[k]) : w[k];
                        pin(res[kStage ? q : 0][k]);
                    }
                }
                const int f0 = kRows * q + r + A * b;
#pragma unroll
                for (int k = 0; k < 16 && !kStage; ++k) {
                    const int f = f0 + 16 * A * k;
                    if constexpr (MOD == MOD_FO && OPK == RC_FWD)
                        StFoBins{(float2*)out, N}(f, w[k]);
                    else if constexpr (MOD == MOD_FO)
                        StFoDemod{out, N, a.tmod}(f, w[k]);
                    else
                        stg(ub((float2*)out, 16 * A * k, f0), OPK == RC_INV ? cconj(w[k]) : w[k]);
                }
            }
        } else if constexpr (kKern) {
            // DIF bottom: value k of block b of row k1 at k1*256 + 16 k + b
            const int tq = fresh_int(tid), r = tq >> 4, b = tq & 15;
            if (r < kRows) {
                float2 w[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) w[s] = buf[S::pb(r, b) + s];
                DFT<16>::run(w);
#pragma unroll
                for (int k = 0; k < 16; ++k) stg(ub((float2*)out, kRows * q * kB + 16 * k, r * kB + b), w[k]);
            }
        } else {
            const int tq = fresh_int(tid), r = tq >> 4;
            if (r < kRows) {
                // DIF bottom, product w <- conj(w K) (conj(w conj K) for the
                // adjoint), then the DIT's first pass (a plain DFT), in place
                const int b = tq & 15;
                const int lo = r * kB + b;  // plus the batch's uniform kRows*q*kB
                const float* kt = a.kern_p ? a.kern_p[t] : a.kern + t * a.kern_st;
                const float2* K = (const float2*)kt;
                float2 kv[16];  // the spectrum, issued before the LDS reads
#pragma unroll
                for (int k = 0; k < 16; ++k) kv[k] = kShared ? *ub(K, kRows * q * kB + 16 * k, lo) : ldg(ub(K, kRows * q * kB + 16 * k, lo));
                float2 w[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) w[s] = buf[S::pb(r, b) + s];
                DFT<16>::run(w);
                if (a.spec) {
                    float2* Sp = (float2*)(a.spec + t * a.spec_st);
#pragma unroll
                    for (int k = 0; k < 16; ++k) stg(ub(Sp, kRows * q * kB + 16 * k, lo), w[k]);
                }
#pragma unroll
                for (int k = 0; k < 16; ++k) w[k] = cconj(kConj ? cmulc(w[k], kv[k]) : cmul(w[k], kv[k]));
                DFT<16>::run(w);
                const int rw = fresh_int(r), bw = fresh_int(b);
#pragma unroll
                for (int s = 0; s < 16; ++s) buf[S::pb(rw, bw) + s] = w[s];
            }
            __syncthreads();
            // this batch's read-back twiddles W^{n2 k1} = W^{n2 a} W^{4 n2 b}
            // (k1 = a + 4 b), issued before the DIT pass so that their
            // latency hides behind it
            constexpr int b0 = (kRows * q) >> 2, nb = ((kRows * q + kRows - 1) >> 2) - b0 + 1;
            float2 rlo[3], rhi[nb];
            {
                const float2* tpl = fresh(a.tfs);
                static_for<3>([&](auto Aa) {
                    constexpr int aa = decltype(Aa)::value + 1;
                    rlo[Aa] = aa < A ? *ub(tpl, aa * kB, n2q) : make_float2(1.f, 0.f);
                });
                static_for<nb>([&](auto Bb) {
                    constexpr int bb = b0 + decltype(Bb)::value;
                    rhi[Bb] = bb ? *ub(tpl, 4 * bb * kB, n2q) : make_float2(1.f, 0.f);
                });
            }
            if (r < kRows) {
                // DIT last pass: forward twiddles before a forward DFT, on
                // conjugate data
                const int j = tq & 15;
                float2 w[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) w[s] = buf[S::pa(r, j) + 17 * s];
                RowTw tw;
                tw.load(fresh(a.tw_row), j);
#pragma unroll
                for (int s = 1; s < 16; ++s) w[s] = tw.apply(w[s], s);
                DFT<16>::run(w);
                const int rw = fresh_int(r), jw = fresh_int(j);
#pragma unroll
                for (int k = 0; k < 16; ++k) buf[S::pa(rw, jw) + 17 * k] = w[k];
            }
            __syncthreads();
            // back to the column lanes, times W_n^{n2 k1} for the closing columns
            static_for<kRows>([&](auto Rr) {
                constexpr int r2 = decltype(Rr)::value, k1 = kRows * q + r2, aa = k1 & 3, bb = k1 >> 2;
                const float2 y = buf[S::pos(r2, n2q)];
                if constexpr (aa == 0)
                    v[k1] = bb ? cmul(y, rhi[bb - b0]) : y;
                else
                    v[k1] = cmul(y, bb ? cmul(rlo[aa - 1], rhi[bb - b0]) : rlo[aa - 1]);
                pin(v[k1]);
            });
        }
    });

    if constexpr (kStage) {
        // bins k2 in [64 c, 64 c + 64) of every row are the A * 64 values
        // X[A * 64 c ...] in a row: lane (r, b) holds bins b + 16 k, so
        // chunk c takes its k = 4 c .. 4 c + 3 of every batch
        constexpr int kChunk = A * 64, kPer = (kChunk + kThreads - 1) / kThreads;
        static_assert(kChunk <= Lay<true, kRows>::kBuf, "a chunk fits the LDS buffer");
        const int tq = fresh_int(tid), r = tq & 15, b = tq >> 4;
        static_for<4>([&](auto C) {
            constexpr int c = decltype(C)::value;
            __syncthreads();  // every lane has read the buffer's previous contents
            if (r < kRows) {
                static_for<NQ>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    static_for<4>([&](auto Kk) {
                        constexpr int kk = decltype(Kk)::value;
                        lds[kRows * q + r + A * (b + 16 * kk)] = res[q][4 * c + kk];
                    });
                });
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int idx = tid + kThreads * i;
                if (kChunk % kThreads == 0 || idx < kChunk)
                    stg(ub((float2*)out, kChunk * c, idx), lds[idx]);
            }
        });
    }

    // -------------------------------------------------------- closing columns
    if constexpr (kConv) {
        DFT<A>::run(v);  // forward DFT over k1 on conjugate data: outputs n1
        if constexpr (MOD == MOD_FO) {
            // y[f] = Re(v W_{4n}^f), y[n + f] = Im(v W_{4n}^f), with
            // W_{4n}^f = W_{4n}^{n2} W_{4A}^{n1}
            const float2 dm = at2(fresh(a.tmod), 0, n2);
            static_for<A>([&](auto N1) {
                constexpr int n1 = decltype(N1)::value;
                const float2 w = rootmul<n1, 4 * A>(cmul(v[n1], dm));
                stg(ub(out, kB * n1, n2), w.x);
                stg(ub(out, N + kB * n1, n2), w.y);
            });
        } else {
            static_for<A>([&](auto N1) {
                constexpr int n1 = decltype(N1)::value;
                stg(ub((float2*)out, kB * n1, n2), cconj(v[n1]));
            });
        }
    }
}

template <int A, int MOD>
hipError_t launch_mod(int opk, const RcolArgs& a, hipStream_t st)
{
    const int bot = (a.conj_kern ? 1 : 0) | (!a.kern_p && a.kern_st == 0 ? 2 : 0);
    void* fn = nullptr;
    switch (opk) {
    case RC_FWD: fn = (void*)k_rcol<A, MOD, RC_FWD, 0>; break;
    case RC_INV: fn = (void*)k_rcol<A, MOD, RC_INV, 0>; break;
    case RC_KERN: fn = (void*)k_rcol<A, MOD, RC_KERN, 0>; break;
    case RC_CONV:
        fn = bot == 0 ? (void*)k_rcol<A, MOD, RC_CONV, 0>
           : bot == 1 ? (void*)k_rcol<A, MOD, RC_CONV, 1>
           : bot == 2 ? (void*)k_rcol<A, MOD, RC_CONV, 2>
                      : (void*)k_rcol<A, MOD, RC_CONV, 3>;
        break;
    default: return hipErrorInvalidValue;
    }
    constexpr int R = rows_of<A>();
    const int lds = opk == RC_FWD || opk == RC_INV ? Lay<true, R>::kLds : Lay<false, R>::kLds;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    RcolArgs args = a;
    void* kargs[] = {(void*)&args};
    const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)a.ntrans), dim3(kThreads), kargs, lds, st);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace rcol
}  // namespace fsa
