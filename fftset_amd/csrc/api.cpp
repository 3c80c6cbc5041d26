// api.cpp -- the fftset C ABI on top of the gfx950 kernels.
//
// Mirrors the reference's L3 layer (src/fftset.c): a set object owning plans
// and twiddle memory, a plan cache keyed by (size, modulation), and four
// execution calls. Execution is a kernel launch on a per-thread HIP stream
// followed by a synchronisation, so a drop-in call returns with its result
// complete, as the reference's CPU call does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

#include "context.h"
#include "fs1.h"
#include "fsr.h"
#include "rcol.h"
#include "planner.h"

static const fftset_modulation g_mod_fo = {fsa::MOD_FO};
static const fftset_modulation g_mod_cx = {fsa::MOD_CX};

extern "C" {
const struct fftset_modulation* FFTSET_MODULATION_FREQ_OFFSET_REAL = &g_mod_fo;
const struct fftset_modulation* FFTSET_MODULATION_COMPLEX = &g_mod_cx;
}

namespace fsa {

static thread_local int t_last_error = FFTSET_OK;

void set_error(int code)
{
    if (code != FFTSET_OK)
        t_last_error = code;
}

// One non-blocking-free (default-flag) stream per thread and device. Default
// flags keep it ordered with respect to the legacy default stream.
hipStream_t thread_stream(int device)
{
    constexpr int kMaxDev = 64;
    static thread_local hipStream_t streams[kMaxDev] = {};
    if (device < 0 || device >= kMaxDev)
        return nullptr;
    if (!streams[device]) {
        if (hipStreamCreateWithFlags(&streams[device], hipStreamDefault) != hipSuccess)
            streams[device] = nullptr;
    }
    return streams[device];
}

int use_device(const Ctx* ctx)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess)
        return FFTSET_E_DEVICE;
    if (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess)
        return FFTSET_E_DEVICE;
    return FFTSET_OK;
}

// Finds or creates the table (N, a0, rs, rows, cols).
const float2* get_table(Ctx* ctx, long long N, int a0, int rows, int cols, hipStream_t st, int rs)
{
    for (const Table& t : ctx->tables)
        if (t.N == N && t.a0 == a0 && t.rs == rs && t.rows == rows && t.cols == cols)
            return t.dev;
    Table t{N, a0, rs, rows, cols, nullptr};
    const size_t bytes = sizeof(float2) * (size_t)rows * (size_t)cols;
    if (hipMalloc((void**)&t.dev, bytes) != hipSuccess)
        return nullptr;
    if (launch_twiddles(t.dev, N, a0, rs, rows, cols, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(t.dev);
        return nullptr;
    }
    ctx->tables.push_back(t);
    ctx->table_bytes += bytes;
    return t.dev;
}

void release_tables_after(Ctx* ctx, size_t mark)
{
    for (size_t i = mark; i < ctx->tables.size(); ++i) {
        const Table& t = ctx->tables[i];
        ctx->table_bytes -= sizeof(float2) * (size_t)t.rows * (size_t)t.cols;
        (void)hipFree(t.dev);
    }
    if (ctx->tables.size() > mark)
        ctx->tables.resize(mark);
}

// Attaches the per-pass tables W_{R m}^{jk} ([k-1][j] layout) to a chain.
bool attach_twiddles(Ctx* ctx, ChainD* c, hipStream_t st)
{
    for (int i = 0; i < c->np; ++i) {
        PassD& p = c->p[i];
        if (p.m <= 1) {
            p.tw = nullptr;
            continue;
        }
        p.tw = get_table(ctx, (long long)p.R * p.m, 1, p.R - 1, p.m, st);
        if (!p.tw)
            return false;
    }
    return true;
}

// Work space of at least `floats` for a call on stream `st`. On success
// `lease` holds the stream's scratch lock: the caller queues its launches
// before releasing it. A buffer that must grow is replaced and the old one
// retired, not freed (kernels queued earlier on the stream may still use it;
// fftset_destroy frees retired buffers). Growing inside graph capture is
// refused, since the allocation would not be part of the graph.
static int lease_scratch(Ctx* ctx, hipStream_t st, size_t floats, std::unique_lock<std::mutex>* lease, float** out)
{
    Ctx::Scratch* s = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (auto& c : ctx->scratch)
            if (c->stream == st)
                s = c.get();
        if (!s) {
            ctx->scratch.push_back(std::make_unique<Ctx::Scratch>());
            s = ctx->scratch.back().get();
            s->stream = st;
        }
    }
    std::unique_lock<std::mutex> lk(s->m);
    if (s->floats < floats) {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
            return FFTSET_E_INVALID;
        float* buf = nullptr;
        if (hipMalloc((void**)&buf, floats * sizeof(float)) != hipSuccess)
            return FFTSET_E_NOMEM;
        if (s->buf)
            s->retired.push_back(s->buf);
        s->buf = buf;
        s->floats = floats;
    }
    *out = s->buf;
    *lease = std::move(lk);
    return FFTSET_OK;
}

static int attach_modulation(fftset_fft* p, Ctx* ctx, hipStream_t st)
{
    p->tmod = nullptr;
    if (p->mod != MOD_FO)
        return FFTSET_OK;
    const unsigned L = p->lfft;
    p->tmod = get_table(ctx, 4LL * L, 1, 1, (int)L, st);
    return p->tmod ? FFTSET_OK : FFTSET_E_NOMEM;
}

int plan_lds_with(fftset_fft* p, Ctx* ctx, hipStream_t st, const std::vector<int>* chain, int G)
{
    const unsigned L = p->lfft;
    std::vector<int> radices = chain && !chain->empty() ? *chain : default_chain(L);
    if (!build_chain(radices, L, &p->chain))
        return FFTSET_E_UNSUPPORTED;
    if (!choose_shape(p->chain, 16, 0, &p->shape) && !choose_shape(p->chain, 32, 0, &p->shape))
        return FFTSET_E_UNSUPPORTED;
    if (G > 0) {
        // requested packing, kept within the lane and LDS limits
        const int max_threads = p->shape.E <= 16 ? 1024 : 512;
        const int lds_one = lds_len((int)L) * 8;
        while (G > 1 && (G * p->shape.T > max_threads || G * lds_one > kMaxLdsBytes)) --G;
        p->shape.G = G;
        p->shape.threads = ((G * p->shape.T + 63) / 64) * 64;
        p->shape.lds_bytes = G * lds_one;
    }
    if (!attach_twiddles(ctx, &p->chain, st))
        return FFTSET_E_NOMEM;
    int rc = attach_modulation(p, ctx, st);
    if (rc != FFTSET_OK)
        return rc;
    p->tfo = nullptr;
    if (p->mod == MOD_FO && p->chain.np >= 2) {
        // W_{4L}^{u (1 + 4k)}: rows k < R0, columns u < L / R0
        p->tfo = get_table(ctx, 4LL * L, 1, p->chain.p[0].R, p->chain.p[0].m, st, 4);
        if (!p->tfo)
            return FFTSET_E_NOMEM;
    }
    p->engine = ENG_LDS;
    return FFTSET_OK;
}

static int plan_lds(fftset_fft* p, Ctx* ctx, hipStream_t st)
{
    const TuneEntry* t = find_tune(ctx, p->mod, p->lfft);
    return plan_lds_with(p, ctx, st, t ? &t->chain : nullptr, t ? t->G : 0);
}

static int run_lds(const fftset_fft* p, int op, long long count, float* out, long long out_st, const float* in,
                   long long in_st, const float* kern, long long kern_st, hipStream_t st, float* spec,
                   long long spec_st)
{
    ShapeD sh = p->shape;
    if (count < sh.G && !choose_shape(p->chain, sh.E, count, &sh))
        return FFTSET_E_UNSUPPORTED;
    LdsArgs a{};
    a.c = p->chain;
    a.tmod = p->tmod;
    a.tfo = p->tfo;
    a.L = (int)p->lfft;
    a.ntrans = count;
    a.G = sh.G;
    a.T = sh.T;
    a.in = in;
    a.out = out;
    a.kern = kern;
    a.in_st = in_st;
    a.out_st = out_st;
    a.kern_st = kern_st;
    a.spec = spec;
    a.spec_st = spec_st;
    return launch_lds(p->mod, op, a, sh, st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
}

// ------------------------------------------------------------- four-step
// Column tiles: C adjacent columns per workgroup, so every global access is
// a run of C values. 32 measured best on MI355X at 61440 = 240 x 256
// (wider runs, still two or more workgroups per CU).
static bool col_shape(const ChainD& c, int max_threads, int C0, ShapeD* out)
{
    ShapeD one;
    if (!choose_shape(c, 16, 1, &one))
        return false;
    int C = C0 > 0 ? C0 : 32;
    while (C > 1 && (C * one.T > max_threads || (long long)C * c.n * 8 > kMaxLdsBytes)) C /= 2;
    if (C * one.T > max_threads || (long long)C * c.n * 8 > kMaxLdsBytes)
        return false;
    out->E = 16;
    out->T = one.T;
    out->G = C;
    out->threads = ((C * one.T + 63) / 64) * 64;
    out->lds_bytes = C * c.n * 8;
    return true;
}

static bool row_shape(const ChainD& c, int rows_avail, int max_threads, ShapeD* out)
{
    if (!choose_shape(c, 16, rows_avail, out) || out->E != 16)
        return false;
    while (out->G > 1 && out->G * out->T > max_threads) {
        out->G--;
        out->lds_bytes = out->G * lds_len(c.n) * 8;
        out->threads = ((out->G * out->T + 63) / 64) * 64;
    }
    return out->G * out->T <= max_threads;
}

int build_four_step(Ctx* ctx, unsigned L, int max_threads, FourStep* out, hipStream_t st, unsigned prefer_A,
                    int cols_C)
{
    // the most balanced split L = A * B with A <= B that both kernels accept
    FourStep best{};
    double best_score = 1e30;
    for (unsigned A = 2; A < L; ++A) {
        if (L % A)
            continue;
        const unsigned B = L / A;
        if (A > B && A != prefer_A)
            continue;
        FourStep fs{};
        fs.A = A;
        fs.B = B;
        if (!fits_lds(A) || !fits_lds(B))
            continue;
        if (!build_chain(default_chain(A), A, &fs.colc) || !build_chain(default_chain(B), B, &fs.rowc))
            continue;
        if (!col_shape(fs.colc, max_threads, cols_C, &fs.cols) ||
            !row_shape(fs.rowc, (int)A, max_threads, &fs.rows))
            continue;
        double score = std::abs(std::log2((double)B) - std::log2((double)A));
        if (prefer_A && A == prefer_A)
            score = -1.0;
        if (score < best_score) {
            best_score = score;
            best = fs;
        }
    }
    if (best_score > 1e29)
        return FFTSET_E_UNSUPPORTED;
    if (!attach_twiddles(ctx, &best.colc, st) || !attach_twiddles(ctx, &best.rowc, st))
        return FFTSET_E_NOMEM;
    best.tfs = get_table(ctx, L, 0, (int)best.A, (int)best.B, st);
    if (!best.tfs)
        return FFTSET_E_NOMEM;
    *out = best;
    return FFTSET_OK;
}

int plan_four_step_with(fftset_fft* p, Ctx* ctx, hipStream_t st, unsigned A, int cols_C)
{
    int rc = build_four_step(ctx, p->lfft, 1024, &p->fs, st, A, cols_C);
    if (rc != FFTSET_OK)
        return rc;
    rc = attach_modulation(p, ctx, st);
    if (rc != FFTSET_OK)
        return rc;
    if (p->mod == MOD_FO) {
        const long long L = p->lfft;
        const int A = (int)p->fs.A, B = (int)p->fs.B;
        p->fs.tfs_fo = get_table(ctx, 4 * L, 1, A, B, st, 4);  // x = (1 + 4 k1) n2
        p->fs.trow_fo = get_table(ctx, 4LL * A, 1, 1, A, st);  // x = n1
        if (!p->fs.tfs_fo || !p->fs.trow_fo)
            return FFTSET_E_NOMEM;
    }
    if (fsr_matches(p->fs.colc, p->fs.rowc)) {
        // the CU-resident kernel's closing columns: first pass of [15, 16]
        p->fs.tw_close = get_table(ctx, (long long)p->fs.A, 1, 14, 16, st);
        if (!p->fs.tw_close)
            return FFTSET_E_NOMEM;
        p->fs.fsr = true;
    }
    p->fs.rcol = rcol_matches(p->fs.colc, p->fs.rowc);
    p->engine = ENG_FOUR_STEP;
    return FFTSET_OK;
}

static int plan_four_step(fftset_fft* p, Ctx* ctx, hipStream_t st)
{
    const TuneEntry* t = find_tune(ctx, p->mod, p->lfft);
    return plan_four_step_with(p, ctx, st, t ? t->A : 0, t && t->chain.empty() ? t->G : 0);
}

// True when (plan, op) runs the CU-resident kernel under the current
// kernel selection.
static bool use_fsr(const fftset_fft* p, int op)
{
    return p->engine == ENG_FOUR_STEP && p->fs.fsr && fsr_op(p->mod, op) && !generic_forced() &&
           !persistent_mode() && !four_step_fused() && !four_step_split();
}

// True when (plan, op) runs the column-register kernel under the current
// kernel selection (every operation of such a plan does).
static bool use_rcol(const fftset_fft* p, int /*op*/)
{
    return p->engine == ENG_FOUR_STEP && p->fs.rcol && !generic_forced() && !persistent_mode() &&
           !four_step_fused() && !four_step_split();
}

bool four_step_work(const fftset_fft* p, int op) { return p->engine == ENG_FOUR_STEP && !use_rcol(p, op); }

static int run_rcol(const fftset_fft* p, int op, long long count, float* out, long long out_st, const float* in,
                    long long in_st, const float* kern, long long kern_st, hipStream_t st, float* spec,
                    long long spec_st)
{
    const FourStep& fs = p->fs;
    RcolArgs a{};
    a.tw_row = fs.rowc.p[0].tw;
    a.tfs = fs.tfs;
    a.tfs_in = p->mod == MOD_FO ? fs.tfs_fo : fs.tfs;
    a.tmod = p->tmod;
    a.ntrans = count;
    a.conj_kern = op == OP_GRAD_IN;
    a.in = in;
    a.in_st = in_st;
    a.out = out;
    a.out_st = out_st;
    a.kern = kern;
    a.kern_st = kern_st;
    a.spec = spec;
    a.spec_st = spec_st;
    const int opk = op == OP_FWD ? RC_FWD : op == OP_INV ? RC_INV : op == OP_KERN ? RC_KERN : RC_CONV;
    count_launch(true);
    return launch_rcol((int)fs.A, p->mod, opk, a, st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
}

static int run_fsr(const fftset_fft* p, int op, long long count, float* out, long long out_st, const float* in,
                   long long in_st, const float* kern, long long kern_st, float* spill, long long spill_st,
                   hipStream_t st, float* spec, long long spec_st)
{
    const FourStep& fs = p->fs;
    FsrArgs a{};
    a.tw_col = fs.colc.p[0].tw;
    a.tw_close = fs.tw_close;
    a.tw_row = fs.rowc.p[0].tw;
    a.tfs = fs.tfs;
    a.tfs_in = p->mod == MOD_FO ? fs.tfs_fo : fs.tfs;
    a.trow = fs.trow_fo;
    a.tmod = p->tmod;
    a.ntrans = count;
    a.conj_kern = op == OP_GRAD_IN;
    a.in = in;
    a.in_st = in_st;
    a.out = out;
    a.out_st = out_st;
    a.kern = kern;
    a.kern_st = kern_st;
    a.spec = spec;
    a.spec_st = spec_st;
    a.spill = (float2*)spill;
    a.spill_st = spill_st / 2;
    const int opk = op == OP_FWD ? FSR_FWD : op == OP_INV ? FSR_INV : op == OP_KERN ? FSR_KERN : FSR_CONV;
    count_launch(true);
    return launch_fsr(p->mod, opk, a, st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
}

static int run_four_step(const fftset_fft* p, int op, long long count, float* out, long long out_st,
                         const float* in, long long in_st, const float* kern, long long kern_st, float* work,
                         long long work_st, hipStream_t st, float* spec,
                         long long spec_st)
{
    const FourStep& fs = p->fs;
    const bool fo = p->mod == MOD_FO;
    FsArgs base{};
    base.tmod = p->tmod;
    base.tfs = fs.tfs;
    base.L = (int)p->lfft;
    base.A = (int)fs.A;
    base.B = (int)fs.B;
    base.ntrans = count;
    base.conj_kern = op == OP_GRAD_IN;
    FsArgs ca = base, ra = base;
    if (fo) {  // column step of the forward direction: folded modulation
        ca.tfs = fs.tfs_fo;
        ca.trow = fs.trow_fo;
    }
    ca.c = fs.colc;
    ca.C = fs.cols.G;
    ca.T = fs.cols.T;
    ca.groups = (int)((fs.B + ca.C - 1) / ca.C);
    ra.c = fs.rowc;
    ra.C = fs.rows.G;
    ra.T = fs.rows.T;
    ra.groups = (int)((fs.A + ra.C - 1) / ra.C);
    auto io = [](FsArgs& a, const float* i, long long ist, float* o, long long ost) {
        a.in = i;
        a.in_st = ist;
        a.out = o;
        a.out_st = ost;
    };
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    if (use_rcol(p, op))  // no work rows: the intermediate stays on the CU
        return run_rcol(p, op, count, out, out_st, in, in_st, kern, kern_st, st, spec, spec_st);
    if (op != OP_KERN && !spec && four_step_fused() && !persistent_mode() && !generic_forced() &&
        fs1_matches(fs.colc, fs.rowc, fs.cols.T, fs.rows.T)) {
        // one workgroup per transform; work holds its scratch slots (at most
        // count of them, 2L floats each)
        Fs1Args f{};
        f.colc = fs.colc;
        f.rowc = fs.rowc;
        f.tmod = p->tmod;
        f.tfs = fs.tfs;
        f.tfs_fo = fs.tfs_fo;
        f.trow = fs.trow_fo;
        f.ntrans = count;
        f.conj_kern = op == OP_GRAD_IN;
        f.in = in;
        f.in_st = in_st;
        f.out = out;
        f.out_st = out_st;
        f.kern = kern;
        f.kern_st = kern_st;
        f.scratch = (float2*)work;
        count_launch(true);
        const int opk = op == OP_FWD ? FS1_FWD : op == OP_INV ? FS1_INV : FS1_CONV;
        return ok(launch_fs1(p->mod, opk, f, fs1_grid(count), st)) ? FFTSET_OK : FFTSET_E_DEVICE;
    }
    if (use_fsr(p, op))  // work (at least kFsrSpill complex values per transform) holds the spill rows
        return run_fsr(p, op, count, out, out_st, in, in_st, kern, kern_st, work, work_st, st, spec, spec_st);
    switch (op) {
    case OP_FWD:
        io(ca, in, in_st, work, work_st);
        io(ra, work, work_st, out, out_st);
        if (!ok(launch_cols(fo ? CL_FOMODC : CL_CX, CS_WORKTW, ca, fs.cols, st)))
            return FFTSET_E_DEVICE;
        return ok(launch_rows(RK_STOCK, fo ? RS_FOBINS : RS_CX, ra, fs.rows, st)) ? FFTSET_OK : FFTSET_E_DEVICE;
    case OP_INV:
        io(ca, in, in_st, work, work_st);
        io(ra, work, work_st, out, out_st);
        ca.tfs = fs.tfs;  // no modulation on the inverse's input side
        if (!ok(launch_cols(fo ? CL_FOBINS : CL_CXCONJ, CS_WORKTW, ca, fs.cols, st)))
            return FFTSET_E_DEVICE;
        return ok(launch_rows(RK_STOCK, fo ? RS_FODEMOD : RS_CXCONJ, ra, fs.rows, st)) ? FFTSET_OK
                                                                                       : FFTSET_E_DEVICE;
    case OP_KERN:
        io(ca, in, in_st, out, out_st);
        io(ra, out, out_st, out, out_st);
        if (!ok(launch_cols(fo ? CL_FOMODC : CL_CX, CS_WORKTW, ca, fs.cols, st)))
            return FFTSET_E_DEVICE;
        return ok(launch_rows(RK_KERN, RS_NONE, ra, fs.rows, st)) ? FFTSET_OK : FFTSET_E_DEVICE;
    case OP_CONV:
    case OP_GRAD_IN: {
        io(ca, in, in_st, work, work_st);
        io(ra, work, work_st, work, work_st);
        ra.kern = kern;
        ra.kern_st = kern_st;
        ra.spec = spec;  // the signal's spectrum, written by the row step's bottom
        ra.spec_st = spec_st;
        FsArgs cb = ca;
        io(cb, work, work_st, out, out_st);
        if (!ok(launch_cols(fo ? CL_FOMODC : CL_CX, CS_WORKTW, ca, fs.cols, st)))
            return FFTSET_E_DEVICE;
        if (!ok(launch_rows(RK_CONV, RS_NONE, ra, fs.rows, st)))
            return FFTSET_E_DEVICE;
        return ok(launch_cols(CL_CX, fo ? CS_FODEMOD : CS_CXCONJ, cb, fs.cols, st)) ? FFTSET_OK : FFTSET_E_DEVICE;
    }
    }
    return FFTSET_E_INVALID;
}

static int run_op(const fftset_fft* p, int op, long long count, float* out, long long out_st, const float* in,
                  long long in_st, const float* kern, long long kern_st, float* work, long long work_st,
                  hipStream_t st, float* spec = nullptr, long long spec_st = 0)
{
    if (p->engine == ENG_LDS)
        return run_lds(p, op, count, out, out_st, in, in_st, kern, kern_st, st, spec, spec_st);
    const long long len2 = 2LL * p->lfft;
    std::unique_lock<std::mutex> lease;
    if (op != OP_KERN && !work && four_step_work(p, op)) {
        int rc = lease_scratch(p->ctx, st, (size_t)(count * len2), &lease, &work);
        if (rc != FFTSET_OK)
            return rc;
        work_st = len2;
    } else if (op == OP_KERN && !work && use_fsr(p, op)) {
        // the resident kernel's spill rows (get_kernel has no work argument)
        const long long sp = 2LL * kFsrSpill;
        int rc = lease_scratch(p->ctx, st, (size_t)(count * sp), &lease, &work);
        if (rc != FFTSET_OK)
            return rc;
        work_st = sp;
    }
    return run_four_step(p, op, count, out, out_st, in, in_st, kern, kern_st, work, work_st, st, spec, spec_st);
}

int execute(const fftset_fft* p, int op, long long count, float* out, long long out_st, const float* in,
            long long in_st, const float* kern, long long kern_st, float* work, long long work_st,
            hipStream_t st, float* work2)
{
    if (!p || count < 0 || op < 0 || op > FFTSET_OP_CONV_SAVE_SPECTRUM)
        return FFTSET_E_INVALID;
    if (count == 0)
        return FFTSET_OK;
    const long long len2 = 2LL * p->lfft;
    if (!out || !in)
        return FFTSET_E_INVALID;
    if ((count > 1 && (out_st < len2 || in_st < len2)) || out_st < 0 || in_st < 0)
        return FFTSET_E_INVALID;
    const bool needs_kern = op == FFTSET_OP_CONV || op == FFTSET_OP_CONV_GRAD_INPUT ||
                            op == FFTSET_OP_CONV_GRAD_KERNEL || op == FFTSET_OP_CONV_SAVE_SPECTRUM;
    if (needs_kern && (!kern || (kern_st != 0 && kern_st < len2 && count > 1)))
        return FFTSET_E_INVALID;
    if (((uintptr_t)out | (uintptr_t)in | (uintptr_t)kern | (uintptr_t)work) & 7)
        return FFTSET_E_INVALID;

    // get_kernel needs work only for the resident kernel's spill rows
    const long long need_st = op == FFTSET_OP_GET_KERNEL ? 2LL * kFsrSpill : len2;
    if (work && count > 1 && work_st < need_st)
        return FFTSET_E_INVALID;
    switch (op) {
    case FFTSET_OP_FORWARD: return run_op(p, OP_FWD, count, out, out_st, in, in_st, nullptr, 0, work, work_st, st);
    case FFTSET_OP_INVERSE: return run_op(p, OP_INV, count, out, out_st, in, in_st, nullptr, 0, work, work_st, st);
    case FFTSET_OP_GET_KERNEL:
        return run_op(p, OP_KERN, count, out, out_st, in, in_st, nullptr, 0, work, work_st, st);
    case FFTSET_OP_CONV:
        return run_op(p, OP_CONV, count, out, out_st, in, in_st, kern, kern_st, work, work_st, st);
    case FFTSET_OP_CONV_GRAD_INPUT:
        return run_op(p, OP_GRAD_IN, count, out, out_st, in, in_st, kern, kern_st, work, work_st, st);
    case FFTSET_OP_CONV_SAVE_SPECTRUM:
        // work receives the signal's spectrum; a four-step intermediate
        // comes from the library's scratch
        if (!work)
            return FFTSET_E_INVALID;
        return run_op(p, OP_CONV, count, out, out_st, in, in_st, kern, kern_st, nullptr, 0, st, work,
                      count > 1 ? work_st : len2);
    case FFTSET_OP_CONV_GRAD_KERNEL: {
        // spectrum of the signal (in) into a scratch region w, then the
        // signal-adjoint pass of the output gradient (kern) against it.
        // Four-step plans need a second region for their intermediate.
        const bool four = four_step_work(p, OP_GRAD_IN);
        float* w = work;
        long long wst = work_st;
        float* w2 = four ? work2 : nullptr;
        std::unique_lock<std::mutex> lease;  // held until both steps are queued
        if (!w || (four && !w2)) {
            const size_t need = (size_t)(count * len2) * ((!w && four) ? 2 : 1);
            float* s = nullptr;
            int rc = lease_scratch(p->ctx, st, need, &lease, &s);
            if (rc != FFTSET_OK)
                return rc;
            if (!w) {
                w = s;
                wst = len2;
                s += count * len2;
            }
            w2 = four ? s : nullptr;
        }
        // (a four-step get_kernel takes its spill rows from w2, which the
        // second step uses only after it)
        int rc = run_op(p, OP_KERN, count, w, wst, in, in_st, nullptr, 0, w2, len2, st);
        if (rc != FFTSET_OK)
            return rc;
        return run_op(p, OP_GRAD_IN, count, out, out_st, kern, kern_st, w, wst, w2, len2, st);
    }
    }
    return FFTSET_E_INVALID;
}

// True for memory the device cannot address as its own: pageable or pinned
// host memory. The drop-in calls accept such buffers, as the reference's
// callers pass, and stage them through device memory.
static bool is_host_ptr(const void* ptr)
{
    if (!ptr)
        return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
        (void)hipGetLastError();  // unknown to the runtime: plain host memory
        return true;
    }
    return a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged;
}

// One drop-in call with some buffers in host memory: copies them through the
// set's staging buffer (4 x 2L floats) around the same execution.
static int staged_call(const fftset_fft* p, int op, float* out, const float* in, const float* kern, float* work,
                       hipStream_t st)
{
    Ctx* ctx = p->ctx;
    const long long len2 = 2LL * p->lfft;
    const size_t bytes = (size_t)len2 * sizeof(float);
    std::lock_guard<std::mutex> g(ctx->stage_mu);
    if (ctx->stage_floats < (size_t)(4 * len2)) {
        if (ctx->stage) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(ctx->stage);
        }
        ctx->stage = nullptr;
        ctx->stage_floats = 0;
        if (hipMalloc((void**)&ctx->stage, (size_t)(4 * len2) * sizeof(float)) != hipSuccess)
            return FFTSET_E_NOMEM;
        ctx->stage_floats = (size_t)(4 * len2);
    }
    float* s = ctx->stage;
    const bool h_out = is_host_ptr(out), h_in = is_host_ptr(in), h_k = is_host_ptr(kern), h_w = is_host_ptr(work);
    float* d_out = h_out ? s : out;
    const float* d_in = h_in ? s + len2 : in;
    const float* d_k = h_k ? s + 2 * len2 : kern;
    float* d_w = h_w ? s + 3 * len2 : work;
    if (h_in && hipMemcpyAsync((void*)d_in, in, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        return FFTSET_E_DEVICE;
    if (h_k && hipMemcpyAsync((void*)d_k, kern, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        return FFTSET_E_DEVICE;
    int rc = execute(p, op, 1, d_out, len2, d_in, len2, d_k, len2, d_w, len2, st);
    if (rc == FFTSET_OK && h_out && hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = FFTSET_E_DEVICE;
    if (hipStreamSynchronize(st) != hipSuccess && rc == FFTSET_OK)
        rc = FFTSET_E_DEVICE;
    return rc;
}

static void sync_call(const fftset_fft* p, int op, float* out, const float* in, const float* kern, float* work)
{
    if (!p) {
        set_error(FFTSET_E_INVALID);
        return;
    }
    int rc = use_device(p->ctx);
    if (rc != FFTSET_OK) {
        set_error(rc);
        return;
    }
    hipStream_t st = thread_stream(p->ctx->device);
    if (!st) {
        set_error(FFTSET_E_DEVICE);
        return;
    }
    if (is_host_ptr(out) || is_host_ptr(in) || is_host_ptr(kern) || is_host_ptr(work)) {
        set_error(staged_call(p, op, out, in, kern, work, st));
        return;
    }
    const long long len2 = 2LL * p->lfft;
    rc = execute(p, op, 1, out, len2, in, len2, kern, len2, work, len2, st);
    if (hipStreamSynchronize(st) != hipSuccess && rc == FFTSET_OK)
        rc = FFTSET_E_DEVICE;
    set_error(rc);
}

}  // namespace fsa

using namespace fsa;

extern "C" {

int fftset_init(struct fftset* fc)
{
    if (!fc)
        return -1;
    std::memset(fc, 0, sizeof(*fc));
    Ctx* ctx = new (std::nothrow) Ctx();
    if (!ctx)
        return -1;
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        delete ctx;
        return -1;
    }
    if (const char* tf = getenv("FFTSET_AMD_TUNE_FILE"))
        (void)load_tune_file(ctx, tf);
    // the no-allocation entry point's work array on this device (a failed
    // lookup disables only that entry point: it then reports FFTSET_E_DEVICE)
    (void)noalloc_work(ctx->device, &ctx->noalloc);
    fc->impl = ctx;
    return 0;
}

void fftset_destroy(struct fftset* fc)
{
    if (!fc || !fc->impl)
        return;
    Ctx* ctx = (Ctx*)fc->impl;
    (void)use_device(ctx);
    (void)hipDeviceSynchronize();
    for (fftset_fft* p = ctx->first; p;) {
        fftset_fft* n = p->next;
        delete p;
        p = n;
    }
    for (Table& t : ctx->tables)
        (void)hipFree(t.dev);
    for (auto& c : ctx->scratch) {
        if (c->buf)
            (void)hipFree(c->buf);
        for (float* r : c->retired)
            (void)hipFree(r);
    }
    if (ctx->stage)
        (void)hipFree(ctx->stage);
    delete ctx;
    fc->impl = nullptr;
}

const struct fftset_fft* fftset_create_fft(struct fftset* fc, const struct fftset_modulation* modulation,
                                           unsigned complex_bins)
{
    if (!fc || !fc->impl || (modulation != &g_mod_fo && modulation != &g_mod_cx)) {
        set_error(FFTSET_E_INVALID);
        return nullptr;
    }
    Ctx* ctx = (Ctx*)fc->impl;
    // search the sorted list (src/fftset.c:103-108)
    for (fftset_fft* p = ctx->first; p; p = p->next) {
        if (p->lfft == complex_bins && p->modulator == modulation)
            return p;
        if (p->lfft < complex_bins)
            break;
    }
    int a, b, c;
    if (!factor235(complex_bins, &a, &b, &c)) {
        set_error(FFTSET_E_UNSUPPORTED);
        return nullptr;
    }
    if (use_device(ctx) != FFTSET_OK) {
        set_error(FFTSET_E_DEVICE);
        return nullptr;
    }
    hipStream_t st = thread_stream(ctx->device);
    fftset_fft* p = new (std::nothrow) fftset_fft();
    if (!p || !st) {
        delete p;
        set_error(p ? FFTSET_E_DEVICE : FFTSET_E_NOMEM);
        return nullptr;
    }
    p->lfft = complex_bins;
    p->modulator = modulation;
    p->ctx = ctx;
    p->mod = modulation->kind;
    const TuneEntry* tuned = find_tune(ctx, p->mod, complex_bins);
    const bool tuned_four_step = tuned && tuned->A && tuned->chain.empty();
    int rc = fits_lds(complex_bins) && !tuned_four_step ? plan_lds(p, ctx, st) : FFTSET_E_UNSUPPORTED;
    if (rc == FFTSET_E_UNSUPPORTED)
        rc = plan_four_step(p, ctx, st);
    if (rc != FFTSET_OK) {
        delete p;
        set_error(rc);
        return nullptr;
    }
    // insert, keeping descending size (src/fftset.c:122-127)
    fftset_fft** ipos = &ctx->first;
    while (*ipos && complex_bins < (*ipos)->lfft) ipos = &(*ipos)->next;
    p->next = *ipos;
    *ipos = p;
    return p;
}

void fftset_fft_forward(const struct fftset_fft* fft, float* out, const float* in, float* work)
{
    sync_call(fft, FFTSET_OP_FORWARD, out, in, nullptr, work);
}

void fftset_fft_inverse(const struct fftset_fft* fft, float* out, const float* in, float* work)
{
    sync_call(fft, FFTSET_OP_INVERSE, out, in, nullptr, work);
}

void fftset_fft_conv_get_kernel(const struct fftset_fft* fft, float* out, const float* in)
{
    sync_call(fft, FFTSET_OP_GET_KERNEL, out, in, nullptr, nullptr);
}

void fftset_fft_conv(const struct fftset_fft* fft, float* out, const float* in, const float* kern, float* work)
{
    sync_call(fft, FFTSET_OP_CONV, out, in, kern, work);
}

void fftset_fft_conv_grad_input(const struct fftset_fft* fft, float* grad_in, const float* grad_out,
                                const float* kern, float* work)
{
    sync_call(fft, FFTSET_OP_CONV_GRAD_INPUT, grad_in, grad_out, kern, work);
}

void fftset_fft_conv_grad_kernel(const struct fftset_fft* fft, float* grad_kern, const float* grad_out,
                                 const float* in, float* work)
{
    // execute() reads the signal through `in` and the output gradient
    // through `kern` for this operation
    sync_call(fft, FFTSET_OP_CONV_GRAD_KERNEL, grad_kern, in, grad_out, work);
}

int fftset_fft_execute(const struct fftset_fft* fft, int op, size_t count, float* out, size_t out_stride,
                       const float* in, size_t in_stride, const float* kern, size_t kern_stride, float* work,
                       size_t work_stride, void* stream)
{
    if (!fft)
        return FFTSET_E_INVALID;
    int rc = use_device(fft->ctx);
    if (rc == FFTSET_OK)
        rc = execute(fft, op, (long long)count, out, (long long)out_stride, in, (long long)in_stride, kern,
                     (long long)kern_stride, work, (long long)work_stride, (hipStream_t)stream);
    set_error(rc);
    return rc;
}

int fftset_last_error(void)
{
    int e = t_last_error;
    t_last_error = FFTSET_OK;
    return e;
}

unsigned fftset_fft_complex_bins(const struct fftset_fft* fft) { return fft ? fft->lfft : 0; }

unsigned long long fftset_table_bytes(const struct fftset* fc)
{
    return (fc && fc->impl) ? (unsigned long long)((const Ctx*)fc->impl)->table_bytes : 0;
}

unsigned fftset_fft_radices(const struct fftset_fft* fft, unsigned* radices, unsigned max)
{
    if (!fft || !radices)
        return 0;
    unsigned n = 0;
    auto put = [&](const ChainD& c) {
        for (int i = 0; i < c.np && n < max; ++i) radices[n++] = (unsigned)c.p[i].R;
    };
    if (fft->engine == ENG_LDS) {
        put(fft->chain);
    } else {
        put(fft->fs.colc);
        if (n < max)
            radices[n++] = 0;
        put(fft->fs.rowc);
    }
    return n;
}

// src/fftset.c:131-167, bit-exact.
static unsigned round_up_factorisation(unsigned min)
{
    unsigned length = 1;
    while (length < min) {
        if (length * 2 >= min)
            length *= 2;
        else if (length * 3 >= min)
            length *= 3;
        else
            length *= 4;
    }
    return length;
}

unsigned fftset_recommend_conv_length(unsigned kernel_length, unsigned max_block_size)
{
    const unsigned kMultiple = 32;  // real-length multiple (src/fftset.c:34)
    const unsigned target = 8 * kernel_length;
    const unsigned block = max_block_size > target ? max_block_size : target;
    const unsigned min_real = kernel_length + block - 1;
    return (kMultiple / 2) * round_up_factorisation((min_real + kMultiple - 1) / kMultiple);
}

}  // extern "C"
