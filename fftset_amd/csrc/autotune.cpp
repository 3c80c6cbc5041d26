// autotune.cpp -- measured choice of radix chain, packing and four-step split.
//
// The reference picks its pass chain from a static cost model
// (src/fftset_vec.c:1004-1062). Here the candidates are timed on the device:
// for a size that fits LDS, register-radix chains (candidate_chains) crossed
// with transforms-per-workgroup packings; for a larger size, the four-step
// splits up to 16:1, each with 16, 32 or 64 columns per column tile
// (stored in the G field of the table). The score is the time of the
// requested operations (forward and/or convolution, one launch each) over a
// batch of about 512 MiB (large enough that launch overhead and clock
// ramp-up do not decide the ranking), best of 3 trials.
// The winner is recorded per (modulation, size) and applies to plans
// created afterwards: the kernel
// spectrum layout depends on the chain, so a plan never changes once made.
// Tables persist as text, one entry per line:
//   <mod> <L> <G> <A> <ms> <n> <r1> ... <rn>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "context.h"
#include "planner.h"
#include "rcol.h"

namespace fsa {

const TuneEntry* find_tune(const Ctx* ctx, int mod, unsigned L)
{
    for (const TuneEntry& t : ctx->tune)
        if (t.mod == mod && t.L == L)
            return &t;
    return nullptr;
}

static void put_tune(Ctx* ctx, const TuneEntry& e)
{
    for (TuneEntry& t : ctx->tune)
        if (t.mod == e.mod && t.L == e.L) {
            t = e;
            return;
        }
    ctx->tune.push_back(e);
}

int load_tune_file(Ctx* ctx, const char* path)
{
    FILE* f = fopen(path, "r");
    if (!f)
        return FFTSET_E_INVALID;
    char line[1024];
    int rc = FFTSET_OK;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n')
            continue;
        TuneEntry e{};
        int n = 0, used = 0;
        if (sscanf(line, "%d %u %d %u %f %d%n", &e.mod, &e.L, &e.G, &e.A, &e.ms, &n, &used) != 6 || n < 0 ||
            n > kMaxPasses || (e.mod != MOD_FO && e.mod != MOD_CX)) {
            rc = FFTSET_E_INVALID;
            continue;
        }
        const char* q = line + used;
        unsigned prod = 1;
        for (int i = 0; i < n; ++i) {
            int r = 0, k = 0;
            if (sscanf(q, "%d%n", &r, &k) != 1 || r < 2 || r > 16) {
                rc = FFTSET_E_INVALID;
                break;
            }
            e.chain.push_back(r);
            prod *= (unsigned)r;
            q += k;
        }
        if ((int)e.chain.size() != n || (n > 0 && prod != e.L)) {
            rc = FFTSET_E_INVALID;
            continue;
        }
        put_tune(ctx, e);
    }
    fclose(f);
    return rc;
}

namespace {

struct Cand {
    std::vector<int> chain;
    int G;
    unsigned A;
};

struct Buffers {
    float *in = nullptr, *out = nullptr, *kern = nullptr, *work = nullptr;
    ~Buffers()
    {
        for (float* p : {in, out, kern, work})
            if (p)
                (void)hipFree(p);
    }
};

// Times the operations in `ops` (bit FFTSET_OP_*) of a temporary plan built
// for `c`.
static int score(Ctx* ctx, int mod, unsigned L, const Cand& c, Buffers& b, long long count, unsigned ops,
                 hipStream_t st, float* ms)
{
    fftset_fft p{};
    p.lfft = L;
    p.ctx = ctx;
    p.mod = mod;
    int rc = c.A ? plan_four_step_with(&p, ctx, st, c.A, c.G) : plan_lds_with(&p, ctx, st, &c.chain, c.G);
    if (rc != FFTSET_OK)
        return rc;
    if (c.A && (p.fs.A != c.A || (c.G && p.fs.cols.G != c.G)))
        return FFTSET_E_UNSUPPORTED;
    const long long n2 = 2LL * L;
    auto run = [&]() {
        int r = FFTSET_OK;
        for (int op = FFTSET_OP_FORWARD; op <= FFTSET_OP_CONV_GRAD_INPUT && r == FFTSET_OK; ++op)
            if (ops & (1u << op))
                r = execute(&p, op, count, b.out, n2, b.in, n2, b.kern, n2, b.work, n2, st);
        return r;
    };
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess)
        return FFTSET_E_DEVICE;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return FFTSET_E_DEVICE;
    }
    rc = run();  // warm-up (first launch of each kernel)
    if (rc == FFTSET_OK)
        rc = run();
    const int reps = 3;
    float best = 1e30f;
    for (int trial = 0; trial < 3 && rc == FFTSET_OK; ++trial) {
        if (hipEventRecord(e0, st) != hipSuccess)
            rc = FFTSET_E_DEVICE;
        for (int i = 0; i < reps && rc == FFTSET_OK; ++i) rc = run();
        if (rc == FFTSET_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
            rc = FFTSET_E_DEVICE;
        float t = 0.f;
        if (rc == FFTSET_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            rc = FFTSET_E_DEVICE;
        best = std::min(best, t / reps);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return rc;
}

}  // namespace
}  // namespace fsa

using namespace fsa;

extern "C" {

int fftset_autotune(struct fftset* fc, const struct fftset_modulation* modulation, unsigned complex_bins,
                    unsigned ops, float* best_ms)
{
    if (ops == 0)
        ops = (1u << FFTSET_OP_FORWARD) | (1u << FFTSET_OP_CONV);
    if (ops & ~((1u << FFTSET_OP_FORWARD) | (1u << FFTSET_OP_INVERSE) | (1u << FFTSET_OP_GET_KERNEL) |
                (1u << FFTSET_OP_CONV) | (1u << FFTSET_OP_CONV_GRAD_INPUT))) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    if (!fc || !fc->impl || !modulation || (modulation->kind != MOD_FO && modulation->kind != MOD_CX)) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    Ctx* ctx = (Ctx*)fc->impl;
    const int mod = modulation->kind;
    int a, b2, c5;
    if (!factor235(complex_bins, &a, &b2, &c5)) {
        set_error(FFTSET_E_UNSUPPORTED);
        return FFTSET_E_UNSUPPORTED;
    }
    for (fftset_fft* p = ctx->first; p; p = p->next)
        if (p->lfft == complex_bins && p->mod == mod) {
            // plans are immutable: their kernel spectra depend on the chain
            set_error(FFTSET_E_INVALID);
            return FFTSET_E_INVALID;
        }
    int rc = use_device(ctx);
    hipStream_t st = thread_stream(ctx->device);
    // Candidate plans take their tables from the set's arena; the ones made
    // only for candidates are freed on every way out (plans created after
    // tuning build the winner's tables anew).
    struct CandidateTables {
        Ctx* ctx;
        size_t mark;
        hipStream_t st;
        ~CandidateTables()
        {
            if (st)
                (void)hipStreamSynchronize(st);
            release_tables_after(ctx, mark);
        }
    } cand_tables{ctx, ctx->tables.size(), st};
    if (rc != FFTSET_OK || !st) {
        set_error(FFTSET_E_DEVICE);
        return FFTSET_E_DEVICE;
    }

    // candidates
    std::vector<Cand> cands;
    if (fits_lds(complex_bins)) {
        for (auto& ch : candidate_chains(complex_bins, 5, 10)) {
            ChainD cd;
            ShapeD sh;
            if (!build_chain(ch, complex_bins, &cd) || !choose_shape(cd, 16, 0, &sh))
                continue;
            const int g0 = sh.G;
            for (int G : {g0, g0 / 2, g0 * 2}) {
                if (G < 1 || G * sh.T > 1024 || (long long)G * lds_len((int)complex_bins) * 8 > kMaxLdsBytes)
                    continue;
                bool dup = false;
                for (auto& c : cands) dup |= c.chain == ch && c.G == G;
                if (!dup)
                    cands.push_back({ch, G, 0});
            }
        }
    }
    // an LDS size that is A x 256 for a column length the column-register
    // kernel is compiled for: that split too (it keeps the transform in
    // VGPRs, so two workgroups share a CU where the LDS engine fits one)
    if (!cands.empty() && complex_bins % 256 == 0 && rcol_length(complex_bins / 256))
        cands.push_back({{}, 0, complex_bins / 256});
    if (cands.empty()) {
        std::vector<std::pair<double, unsigned>> splits;
        for (unsigned A = 2; A < complex_bins; ++A)
            if (complex_bins % A == 0 && fits_lds(A) && fits_lds(complex_bins / A))
                splits.push_back({std::abs(std::log2((double)(complex_bins / A)) - std::log2((double)A)), A});
        std::sort(splits.begin(), splits.end());
        // every split up to 16:1 either way (short columns allow wide,
        // better-coalesced column tiles), each with three tile widths
        for (size_t i = 0; i < splits.size(); ++i)
            if (splits[i].first <= 4.0)
                for (int C : {16, 32, 64}) cands.push_back({{}, C, splits[i].second});
    }

    // buffers: about 512 MiB of transforms
    const long long n2 = 2LL * complex_bins;
    const long long count = std::max<long long>(256, (512LL << 20) / (n2 * 4));
    Buffers b;
    const size_t bytes = (size_t)(count * n2) * sizeof(float);
    if (hipMalloc((void**)&b.in, bytes) != hipSuccess || hipMalloc((void**)&b.out, bytes) != hipSuccess ||
        hipMalloc((void**)&b.kern, bytes) != hipSuccess || hipMalloc((void**)&b.work, bytes) != hipSuccess) {
        set_error(FFTSET_E_NOMEM);
        return FFTSET_E_NOMEM;
    }
    (void)hipMemsetAsync(b.in, 0, bytes, st);
    (void)hipMemsetAsync(b.kern, 0, bytes, st);

    TuneEntry best{};
    best.ms = 1e30f;
    best.mod = mod;
    best.L = complex_bins;
    for (const Cand& c : cands) {
        float ms = 0.f;
        if (score(ctx, mod, complex_bins, c, b, count, ops, st, &ms) != FFTSET_OK)
            continue;
        const float norm = ms * (64.f * (1 << 20)) / (float)(count * n2 * 4);  // reported per 64 MiB
        if (norm < best.ms) {
            best.ms = norm;
            best.chain = c.chain;
            best.G = c.G;
            best.A = c.A;
        }
    }
    (void)hipStreamSynchronize(st);
    if (best.ms > 1e29f) {
        set_error(FFTSET_E_UNSUPPORTED);
        return FFTSET_E_UNSUPPORTED;
    }
    put_tune(ctx, best);
    if (best_ms)
        *best_ms = best.ms;
    return FFTSET_OK;
}

int fftset_tune_save(const struct fftset* fc, const char* path)
{
    if (!fc || !fc->impl || !path)
        return FFTSET_E_INVALID;
    const Ctx* ctx = (const Ctx*)fc->impl;
    FILE* f = fopen(path, "w");
    if (!f)
        return FFTSET_E_INVALID;
    fprintf(f, "# fftset_amd tune table: mod(0=FREQOFFSETREAL,1=COMPLEX) L G A ms_per_64MiB n radices...\n");
    for (const TuneEntry& t : ctx->tune) {
        fprintf(f, "%d %u %d %u %.5f %d", t.mod, t.L, t.G, t.A, t.ms, (int)t.chain.size());
        for (int r : t.chain) fprintf(f, " %d", r);
        fprintf(f, "\n");
    }
    return fclose(f) == 0 ? FFTSET_OK : FFTSET_E_INVALID;
}

int fftset_tune_load(struct fftset* fc, const char* path)
{
    if (!fc || !fc->impl || !path)
        return FFTSET_E_INVALID;
    return load_tune_file((Ctx*)fc->impl, path);
}

}  // extern "C"
