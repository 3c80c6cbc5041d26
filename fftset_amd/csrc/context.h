// context.h -- host-side objects behind the opaque C types.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "fftset.h"
#include "kernels.h"
#include "plan_desc.h"

struct fftset_modulation {
    int kind;  // fsa::MOD_FO or fsa::MOD_CX
};

namespace fsa {

// A twiddle table in HBM: dev[r*cols + c] = W_N^{((a0 + r) c) mod N},
// generated in FP64 on the GPU. Tables are keyed by (N, a0, rows, cols) and
// shared by every plan that needs the same one -- a pass of sub-length N and
// radix R = rows+1 is the same table in every plan containing it (the
// reference shares its per-pass tables the same way, by suffix reuse:
// src/fftset_vec.c:1064-1120).
struct Table {
    long long N;
    int a0, rs, rows, cols;
    float2* dev;
};

// Four-step decomposition n = A * B for transforms larger than LDS.
// Column step: B transforms of length A (stride B); row step: A transforms
// of length B (contiguous), with W_n^{n2 k1} applied between them.
struct FourStep {
    unsigned A, B;
    ChainD colc, rowc;    // chains of the column (length A) and row (length B) transforms
    ShapeD cols, rows;    // launch shapes of the column and row kernels
    const float2* tfs;    // tfs[k1*B + n2] = W_n^{n2 k1}
    // FREQOFFSETREAL: the modulation W_{4n}^{n1 B + n2} of the column step
    // split into a row factor (input) and a column factor folded into the
    // output twiddles: trow_fo[n1] = W_{4A}^{n1}, tfs_fo[k1*B + n2] =
    // W_n^{n2 k1} W_{4n}^{n2} = W_{4n}^{n2 (4 k1 + 1)}
    const float2* tfs_fo = nullptr;
    const float2* trow_fo = nullptr;
    // The 240 x 256 split with chains [16, 15] and [16, 16]: the CU-resident
    // kernel (kernels_fsr_*.hip) runs it, with the closing columns' first
    // pass table W_240^{u k} at (k-1)*16 + u (chain [15, 16]).
    bool fsr = false;
    const float2* tw_close = nullptr;
    // An A x 256 split with row chain [16, 16] and a compiled column length:
    // the column-register kernel (kernels_rcol_*.hip) runs every operation,
    // with no intermediate in HBM.
    bool rcol = false;
};

// Autotuned choices for one (modulation, size): radix chain and packing of
// the LDS engine, or the four-step split. Empty / zero fields keep the
// defaults.
struct TuneEntry {
    int mod;
    unsigned L;
    std::vector<int> chain;
    int G;       // transforms per workgroup (LDS engine) or columns per column tile (four-step)
    unsigned A;  // four-step column length
    float ms;    // measured score of the tuned operations, for reports
};

struct Ctx {
    int device = 0;
    std::vector<TuneEntry> tune;
    std::vector<Table> tables;
    size_t table_bytes = 0;
    fftset_fft* first = nullptr;  // plans sorted by descending size (src/fftset.c:97-129)
    std::mutex mu;  // guards the scratch list
    // Library-owned work space for calls that pass no work buffer, one per
    // stream: calls on different streams never share a buffer. A call holds
    // its stream's lock from taking the buffer until its launches are queued,
    // so calls from several host threads on one stream do not interleave
    // their steps over the same buffer.
    struct Scratch {
        hipStream_t stream = nullptr;
        float* buf = nullptr;
        size_t floats = 0;
        std::vector<float*> retired;  // outgrown buffers, freed by fftset_destroy
        std::mutex m;
    };
    std::vector<std::unique_ptr<Scratch>> scratch;
    std::mutex stage_mu;          // guards stage
    float* stage = nullptr;       // device copies of host buffers passed to the drop-in calls
    float* noalloc = nullptr;     // the device's static work array (fftset_fft_execute_noalloc)
    size_t stage_floats = 0;
};

enum Engine : int { ENG_LDS = 0, ENG_FOUR_STEP = 1 };

}  // namespace fsa

struct fftset_fft {
    // Search keys, as in the reference (src/fftset_modulation.h:32-37).
    unsigned lfft;
    const fftset_modulation* modulator;
    fftset_fft* next;

    fsa::Ctx* ctx;
    int mod;
    int engine;
    fsa::ChainD chain;  // LDS engine chain (ENG_LDS)
    fsa::ShapeD shape;  // LDS engine launch shape for large batches
    fsa::FourStep fs;   // ENG_FOUR_STEP
    const float2* tmod;  // FREQOFFSETREAL modulation W_{4L}^e, e < L (null for COMPLEX)
    // FREQOFFSETREAL LDS plans with two or more passes: the first pass's
    // folded modulation table W_{4L}^{u (1 + 4k)} (k < R0, u < L / R0)
    const float2* tfo = nullptr;
};

namespace fsa {
// Helpers shared by the C-ABI files.
hipStream_t thread_stream(int device);
int use_device(const Ctx* ctx);
void set_error(int code);
const float2* get_table(Ctx* ctx, long long N, int a0, int rows, int cols, hipStream_t st, int rs = 1);
// Frees the tables created after the arena held `mark` of them (the caller
// has synchronised every stream that used them and no plan refers to them).
void release_tables_after(Ctx* ctx, size_t mark);
bool attach_twiddles(Ctx* ctx, ChainD* c, hipStream_t st);
// Four-step split of L whose column and row tiles fit max_threads lanes;
// prefer_A > 0 forces that column length when it is valid.
// cols_C: columns per column tile (0 = default).
int build_four_step(Ctx* ctx, unsigned L, int max_threads, FourStep* out, hipStream_t st, unsigned prefer_A = 0,
                    int cols_C = 0);
// Plans p (mod, lfft, ctx set) with an optional radix chain / packing
// (LDS engine) or split (four-step) instead of the defaults.
int plan_lds_with(fftset_fft* p, Ctx* ctx, hipStream_t st, const std::vector<int>* chain, int G);
int plan_four_step_with(fftset_fft* p, Ctx* ctx, hipStream_t st, unsigned A, int cols_C = 0);
// True when operation `op` of four-step plan p keeps its intermediate in a
// work region (false for plans the column-register kernel runs).
bool four_step_work(const fftset_fft* p, int op);
const TuneEntry* find_tune(const Ctx* ctx, int mod, unsigned L);
int load_tune_file(Ctx* ctx, const char* path);

// Runs `op` over a uniform batch on `stream` (asynchronous). Returns an
// FFTSET_* code. work2 (optional, count * 2L floats): the second work
// region of a four-step CONV_GRAD_KERNEL, which otherwise comes from the
// library's scratch.
int execute(const fftset_fft* fft, int op, long long count, float* out, long long out_st, const float* in,
            long long in_st, const float* kern, long long kern_st, float* work, long long work_st,
            hipStream_t stream, float* work2 = nullptr);
// Device address of the static work array of fftset_fft_execute_noalloc on
// `device` (which must be current); looked up once per device.
int noalloc_work(int device, float** out);
}  // namespace fsa
