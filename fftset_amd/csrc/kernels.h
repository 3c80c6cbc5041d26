// kernels.h -- kernel argument blocks and host-side launchers.
#pragma once

#include "engine.h"
#include "plan_desc.h"

namespace fsa {

// One launch of the single-kernel LDS engine over a uniform batch.
struct LdsArgs {
    ChainD c;
    const float2* tmod;  // FREQOFFSETREAL modulation: tmod[e] = W_{4L}^e, e < L
    const float2* tfo;   // FREQOFFSETREAL first / closing pass: tfo[k m + u] = W_{4L}^{u (1 + 4k)}
    int L;               // complex bins (== c.n)
    long long ntrans;  // transforms in the batch
    int G, T;          // transforms per workgroup, lanes per transform
    const float* in;
    float* out;
    const float* kern;
    long long in_st, out_st, kern_st;  // floats between consecutive transforms
    float* spec;        // convolution: the signal's spectrum out (optional), stride spec_st
    long long spec_st;
};

hipError_t launch_lds(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st);

// Four-step (n = A*B) steps for transforms larger than LDS.
enum ColLoad : int { CL_FOMOD = 0, CL_FOBINS = 1, CL_CX = 2, CL_CXCONJ = 3, CL_FOMODC = 4 };
enum ColStore : int { CS_WORKTW = 0, CS_FODEMOD = 1, CS_CXCONJ = 2 };
enum RowKind : int { RK_STOCK = 0, RK_CONV = 1, RK_KERN = 2 };
enum RowStore : int { RS_FOBINS = 0, RS_CX = 1, RS_FODEMOD = 2, RS_CXCONJ = 3, RS_NONE = 4 };

struct FsArgs {
    ChainD c;            // column chain (length A) or row chain (length B)
    const float2* tmod;  // FREQOFFSETREAL modulation: tmod[e] = W_{4L}^e
    const float2* tfs;   // tfs[k1*B + n2] = W_L^{n2 k1}
    int L, A, B;         // n = L = A*B
    long long ntrans;
    int C;       // columns (or rows) per workgroup
    int T;       // lanes per column (or row) transform
    int groups;  // workgroups per transform
    int conj_kern;
    const float* in;
    float* out;
    const float* kern;
    long long in_st, out_st, kern_st;
    float* spec;  // row step of a convolution: the signal's spectrum out (optional)
    long long spec_st;
    // Per-transform pointers (ragged batches): when set, transform t reads
    // in_p[t] / kern_p[t] and writes out_p[t], and the strides are unused.
    const float* const* in_p;
    float* const* out_p;
    const float* const* kern_p;
    // FREQOFFSETREAL column loads with folded modulation (CL_FOMODC):
    // trow[n1] = W_{4A}^{n1}; the column factor W_{4L}^{n2} is in tfs
    const float2* trow;
};

hipError_t launch_cols(int load, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st);
hipError_t launch_rows(int kind, int store, const FsArgs& a, const ShapeD& sh, hipStream_t st);

// ------------------------------------------------------------ ragged batch
// A ragged batch is a table of tiles over a table of classes (one class per
// (plan, operation)); one workgroup runs one tile. Tiles are sorted by cost,
// largest first, so the hardware's in-order dispatch balances the load.
enum TileKind : int { TK_LDS = 0, TK_COL = 1, TK_ROW = 2, TK_COL2 = 3 };

struct RClass {
    int mod, op;  // Mod, Op
    int L, A, B;  // A, B > 0 for four-step classes
    int T, G;     // LDS tiles: lanes per transform, transforms per tile
    int colT, colC, colGroups;  // four-step column tiles
    int rowT, rowR, rowGroups;  // four-step row tiles
    ChainD c;   // LDS chain or column chain (length A)
    ChainD c2;  // row chain (length B)
    const float2* tmod;
    const float2* tfo;  // FREQOFFSETREAL LDS classes: folded first / closing pass table
    const float2* tfs;
    const float* const* in;  // per-transform buffers of this class
    float* const* out;
    const float* const* kern;
    float* const* work;
};

struct RTile {
    int cls;
    int kind;   // TileKind
    int job;    // first transform (TK_LDS) or the transform (four-step)
    int count;  // transforms (TK_LDS) or the column / row group
};

struct RaggedArgs {
    const RClass* classes;
    const RTile* tiles;
    int ntiles;
    int* next;  // persistent launch: tile counter (zeroed by the launcher); null: one tile per workgroup
    int four_step;  // every tile is a four-step tile (else every tile is an LDS tile)
};

// Launches tiles [0, ntiles) of `a`; every class must fit `threads` lanes.
hipError_t launch_ragged(const RaggedArgs& a, int threads, int lds_bytes, hipStream_t st);

// Per-pass baseline (kernels_pass.hip): step `step` of every transform of
// one class, one workgroup per transform, data in `work` between steps.
int pass_steps(int op, int np);
hipError_t launch_pass(const RClass* d_cls, int step, int ntrans, int T, float2* work, hipStream_t st);

// Per-size-launch baseline with the reference's pass structure
// (kernels_refpass.hip): one launch per sweep over every transform of a class.
enum RpKind : int { RPK_DIF = 0, RPK_DIT = 1, RPK_STOCK = 2, RPK_MULCONJ = 3, RPK_INTER = 4, RPK_DEINTER = 5 };
// Where a sweep reads or writes: the work row, or the transform's input /
// output (FREQOFFSETREAL: modulated / demodulated; COMPLEX: plain / conj).
enum RpIo : int { RP_WORK = 0, RP_IN_FO = 1, RP_IN_CX = 2, RP_OUT_FO = 3, RP_OUT_CX = 4, RP_IN_OUT_CX = 5 };

struct RpArgs {
    int L;           // complex bins
    int R, m, s;     // the pass (in-place view; Stockham: blocks per lane)
    int V;           // sub-transforms (lanes) per transform
    int conj_kern;   // product with conj(K) (signal adjoint)
    long long ntrans;
    const float2* tw;    // tw[(k-1) m + j] = W_{Rm}^{jk}; null when m == 1
    const float2* tmod;  // FREQOFFSETREAL modulation W_{4L}^e
    const float* const* in;  // per-transform input pointers (device array)
    float* const* out;       // per-transform output pointers
    float2* work;    // ntrans * L: in-place passes, Stockham source
    float2* work2;   // ntrans * L: Stockham destination
    const float2* kref;  // ntrans * L: kernel spectra in this structure's order
    int slot[8];     // lane slot holding residue l (forward reorder)
    int resid[8];    // residue held by lane slot q (inverse de-interleave)
};

hipError_t launch_refpass(int kind, int R, int io, const RpArgs& a, hipStream_t st);
// kref-style gather at set-up: a.work[t][p] = ((float2*)a.in[t])[perm[p]].
hipError_t launch_refpass_gather(const RpArgs& a, const int* perm, hipStream_t st);

// ------------------------------------------------------- block convolution
// Overlap-save FIR filtering of long real signals (see kernels_block.hip).
struct BlockArgs {
    ChainD c;
    const float2* tmod;
    const float2* tfo;   // folded first / closing pass table of the plan
    int L;               // complex bins; frames are N = 2L samples
    long long ntrans;    // channels * nblocks frames
    int G, T;
    const float* in;     // channel c starts at in + c * in_st
    float* out;          // channel c starts at out + c * out_st
    const float* kern;   // shared kernel spectrum of the taps (scaled 1/L)
    long long in_st, out_st;
    long long n;         // samples per channel
    long long nblocks;   // blocks per channel
    long long b;         // new samples per block (N - k + 1)
    int skip;            // k - 1
    long long lo;        // first readable sample: 0, or -(k-1) with history
};

hipError_t launch_block(const BlockArgs& a, const ShapeD& sh, hipStream_t st);
hipError_t launch_taps(float* dst, const float* taps, int k, int n, float scale, hipStream_t st);

// Grid for a persistent kernel: workgroups resident at once, capped by the
// number of tiles.
long long persistent_grid(const void* fn, int threads, int lds_bytes, long long tiles);
int persistent_waves();
// Kernel selection (select.cpp): FFTSET_SELECT_PERSISTENT and
// FFTSET_SELECT_GENERIC of fftset_kernel_select, and a counter of launches
// of specialised (compile-time chain) and generic (runtime chain) kernels.
bool persistent_mode();
bool generic_forced();
bool four_step_fused();
bool four_step_split();
void count_launch(bool specialised);

// out[r*cols + c] = W_N^{((a0 + r) * c) mod N}, computed in FP64.
hipError_t launch_twiddles(float2* out, long long N, int a0, int rs, int rows, int cols, hipStream_t st);

}  // namespace fsa
