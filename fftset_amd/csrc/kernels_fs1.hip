// kernels_fs1.hip -- four-step transforms with one workgroup per transform.
//
// n = A * B = 240 * 256 (61440 bins). The two-launch four-step path
// (kernels_fs.hip) writes its intermediate for every transform of a batch
// to HBM between a column launch and a row launch: 2x (forward) or 2.33x
// (convolution) the algorithmic bytes through HBM. Here one 1024-lane
// workgroup runs all steps of one transform, in chunks that fit LDS:
//   columns: B / C chunks of C = 64 adjacent columns (length-A DFTs, chain
//            [16, 15]), twiddled, into the workgroup's scratch slot;
//   rows:    A / RPC chunks of RPC = 60 rows (length-B DFTs, chain [16, 16])
//            from the slot, to the output through an LDS transpose -- or, for
//            a convolution, DIF -> product -> DIT back into the slot;
//   closing columns (convolution): the conjugate-data inverse column DFTs
//            from the slot to the output.
// The slot (n complex values, 480 KB) is reused for every transform the
// workgroup takes, so the live intermediates of the whole grid (one
// workgroup per CU, 256 slots = 123 MB) fit the 256 MB Infinity Cache and
// the slot round trips never reach HBM. The slot is written and read only
// by its own workgroup, ordered by barriers: no workgroup waits on another.
// Grid: one workgroup per CU, transform t = blockIdx.x + i * gridDim.x.
// Index maps, twiddle tables and data order are the two-launch path's, so
// results and kernel spectra are interchangeable with it.
//
// Opt-in (FFTSET_SELECT_FOUR_STEP_FUSED): measured on MI355X it is slower
// than the two launches (forward 0.578 against 0.460 ms, convolution 1.074
// against 0.855 ms for 1250 transforms). With one 1024-lane workgroup per
// CU, every chunk's loads, LDS passes and stores run back to back on the
// CU, so memory latency is exposed; the two-launch kernels already move
// their bytes near the copy ceiling, and Infinity-Cache traffic is bound by
// the same fabric.
#include "drivers.h"
#include "fs1.h"
#include "kernels.h"

namespace fsa {

// Scratch-slot accesses: plain (cached) loads and stores, so the slot stays
// in L2 / the Infinity Cache; signals, outputs and kernel spectra stream
// with nontemporal accesses (ldg / stg).
struct LdSlot {
    const float2* p;
    __device__ __forceinline__ float2 operator()(int e) const { return p[e]; }
};

// Column-step output into the slot: Y[k1][n2] tw[k1 B + n2] at k1 B + n2.
struct StSlotTw {
    float2* slot;
    int B, n2;
    const float2* __restrict__ tw;
    __device__ __forceinline__ void operator()(int k1, float2 v) const
    {
        const int e = k1 * B + n2;
        slot[e] = cmul(v, tw[e]);
    }
};

// Row-step output of the convolution back into row k1 of the slot, times
// W_n^{n2 k1} (trow = tfs + k1 B).
struct StSlotRowTw {
    float2* row;
    const float2* __restrict__ trow;
    __device__ __forceinline__ void operator()(int n2, float2 v) const { row[n2] = cmul(v, trow[n2]); }
};

// The twiddle tables are the same for every chunk and transform; left to
// itself the compiler hoists their loads out of the loops and keeps both
// chains' twiddles live across all steps, which spills. The table pointers
// pass through fresh() (dev_math.h) each chunk, which keeps every load
// inside its chunk.

template <int MOD, int OPK, int BOT>
__global__ void __launch_bounds__(1024) k_fs1(const Fs1Args a)
{
    using FC = Fix2<16, 15>;  // column chain, A = 240
    using FR = Fix2<16, 16>;  // row chain, B = 256
    constexpr int A = FC::n, B = FR::n, L = A * B;
    constexpr int kIdle = 1 << 28;
    // Chunk geometry from the arguments (64 columns, 60 rows of 273 padded
    // slots): folded to constants, the LDS addresses of a whole chunk are
    // materialised at once and the kernel spills.
    const int C = a.cols, RPC = a.rows, LS = a.row_ld;
    extern __shared__ float2 smem[];
    float2* slot = a.scratch + (long long)blockIdx.x * L;

    // column-chunk lane: column g of the chunk, lane lc within the column
    const int g = threadIdx.x % C, lc = threadIdx.x / C;
    const TileInter ctile{smem + g, C};
    // row-chunk lane: row gr of the chunk, lane lr within the row
    const int gr = threadIdx.x / FR::T, lr = threadIdx.x - gr * FR::T;
    const bool row_ok = gr < RPC;
    const TileContig rtile{smem + (row_ok ? gr : 0) * LS};

    for (long long t = blockIdx.x; t < a.ntrans; t += gridDim.x) {
        const float* in = a.in_p ? a.in_p[t] : a.in + t * a.in_st;
        float* out = a.out_p ? a.out_p[t] : a.out + t * a.out_st;

        // ---- columns: input -> slot (twiddled)
#pragma unroll 1
        for (int grp = 0; grp < a.col_chunks; ++grp) {
            const int n2 = grp * C + g;
            const float2* tw0 = fresh(a.colc.p[0].tw);
            auto cols = [&](const auto& ld, const auto& st) {
                pass_first<16, FC::NB0>(ctile, lc, FC::T, FC::m0, tw0, ld);
                __syncthreads();
                stock_last<15, FC::NB1>(ctile, lc, FC::T, FC::s1, st);
            };
            if constexpr (OPK == FS1_INV) {
                const StSlotTw st{slot, B, n2, fresh(a.tfs)};
                if constexpr (MOD == MOD_FO)
                    cols(LdCol<LdFoBins>{{(const float2*)in, L}, B, n2}, st);
                else
                    cols(LdCol<LdCx<true>>{{(const float2*)in}, B, n2}, st);
            } else {
                if constexpr (MOD == MOD_FO)
                    cols(LdColFoMod{in, L, B, n2, fresh(a.trow)}, StSlotTw{slot, B, n2, fresh(a.tfs_fo)});
                else
                    cols(LdCol<LdCx<false>>{{(const float2*)in}, B, n2}, StSlotTw{slot, B, n2, fresh(a.tfs)});
            }
            __syncthreads();
        }

        // ---- rows
#pragma unroll 1
        for (int rc = 0; rc < a.row_chunks; ++rc) {
            const int k1_0 = rc * RPC;
            const int k1 = k1_0 + (row_ok ? gr : 0);
            const int lt = row_ok ? lr : kIdle;
            const LdSlot ld{slot + k1 * B};
            const float2* tw0 = fresh(a.rowc.p[0].tw);
            if constexpr (OPK == FS1_CONV) {
                // DIF -> x K -> DIT back into the slot row
                const float* kern = a.kern_p ? a.kern_p[t] : a.kern + t * a.kern_st;
                const BotConvC<(BOT & 1) != 0, (BOT & 2) != 0> bot{(const float2*)kern + k1 * B};
                pass_first<16, FR::NB0>(rtile, lt, FR::T, FR::m0, tw0, ld);
                __syncthreads();
                dif_bottom<16, FR::NB1>(rtile, lt, FR::T, FR::s1, bot);
                __syncthreads();
                dit_last<16, FR::NB0>(rtile, lt, FR::T, FR::m0, fresh(a.rowc.p[0].tw),
                                      StSlotRowTw{slot + k1 * B, fresh(a.tfs) + k1 * B});
            } else {
                pass_first<16, FR::NB0>(rtile, lt, FR::T, FR::m0, tw0, ld);
                __syncthreads();
                stock_mid<16, FR::NB1>(rtile, lt, FR::T, 1, FR::s1, fresh(a.rowc.p[1].tw));
                __syncthreads();
                // transposed store: consecutive lanes write consecutive bins
                // f = k1_0 + gg + A k2 (runs of RPC values)
                for (int idx = threadIdx.x; idx < RPC * B; idx += blockDim.x) {
                    const int gg = idx % RPC, k2 = idx / RPC;
                    const float2 v = smem[gg * LS + lds_pad(k2)];
                    const int f = k1_0 + gg + A * k2;
                    if constexpr (OPK == FS1_FWD) {
                        if constexpr (MOD == MOD_FO)
                            StFoBins{(float2*)out, L}(f, v);
                        else
                            StCx<false>{(float2*)out}(f, v);
                    } else {
                        if constexpr (MOD == MOD_FO)
                            StFoDemod{out, L, fresh(a.tmod)}(f, v);
                        else
                            StCx<true>{(float2*)out}(f, v);
                    }
                }
            }
            __syncthreads();
        }

        // ---- closing columns of a convolution: slot -> output
        if constexpr (OPK == FS1_CONV) {
#pragma unroll 1
            for (int grp = 0; grp < a.col_chunks; ++grp) {
                const int n2 = grp * C + g;
                const LdCol<LdSlot> ld{{slot}, B, n2};
                pass_first<16, FC::NB0>(ctile, lc, FC::T, FC::m0, fresh(a.colc.p[0].tw), ld);
                __syncthreads();
                if constexpr (MOD == MOD_FO)
                    stock_last<15, FC::NB1>(ctile, lc, FC::T, FC::s1, StCol<StFoDemod>{{out, L, fresh(a.tmod)}, B, n2});
                else
                    stock_last<15, FC::NB1>(ctile, lc, FC::T, FC::s1, StCol<StCx<true>>{{(float2*)out}, B, n2});
                __syncthreads();
            }
        }
    }
}

// ------------------------------------------------------------------ host
bool fs1_matches(const ChainD& colc, const ChainD& rowc, int colT, int rowT)
{
    return Fix2<16, 15>::matches(colc, colT) && Fix2<16, 16>::matches(rowc, rowT);
}

int fs1_lds_bytes()
{
    const int cols = (1024 / Fix2<16, 15>::T) * Fix2<16, 15>::n * 8;
    const int rows = 60 * lds_len(Fix2<16, 16>::n) * 8;  // 60 rows per chunk
    return cols > rows ? cols : rows;
}

// One workgroup per CU (the kernel's LDS admits no second one).
long long fs1_grid(long long ntrans)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return ntrans < cus ? ntrans : cus;
}

template <int MOD, int OPK>
static void* fs1_fn(int bot)
{
    if constexpr (OPK == FS1_CONV) {
        switch (bot) {
        case 0: return (void*)k_fs1<MOD, OPK, 0>;
        case 1: return (void*)k_fs1<MOD, OPK, 1>;
        case 2: return (void*)k_fs1<MOD, OPK, 2>;
        default: return (void*)k_fs1<MOD, OPK, 3>;
        }
    }
    return (void*)k_fs1<MOD, OPK, 0>;
}

hipError_t launch_fs1(int mod, int opk, const Fs1Args& args, long long grid, hipStream_t st)
{
    if (args.ntrans <= 0 || grid <= 0)
        return hipSuccess;
    Fs1Args a = args;
    a.cols = 1024 / Fix2<16, 15>::T;
    a.col_chunks = Fix2<16, 16>::n / a.cols;
    a.rows = 60;
    a.row_chunks = Fix2<16, 15>::n / a.rows;
    a.row_ld = lds_len(Fix2<16, 16>::n);
    const int bot = (a.conj_kern ? 1 : 0) | (!a.kern_p && a.kern_st == 0 ? 2 : 0);
    void* fn = nullptr;
    if (mod == MOD_FO)
        fn = opk == FS1_FWD ? fs1_fn<MOD_FO, FS1_FWD>(bot) : opk == FS1_INV ? fs1_fn<MOD_FO, FS1_INV>(bot)
                                                                            : fs1_fn<MOD_FO, FS1_CONV>(bot);
    else
        fn = opk == FS1_FWD ? fs1_fn<MOD_CX, FS1_FWD>(bot) : opk == FS1_INV ? fs1_fn<MOD_CX, FS1_INV>(bot)
                                                                            : fs1_fn<MOD_CX, FS1_CONV>(bot);
    const int lds = fs1_lds_bytes();
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    void* kargs[] = {(void*)&a};
    const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)grid), dim3(1024), kargs, lds, st);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace fsa
