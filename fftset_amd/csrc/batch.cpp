// batch.cpp -- ragged batches: thousands of transforms of mixed sizes,
// modulations and operations run as one graph launch.
//
// Commit groups the transforms into classes (plan, operation), picks for
// each class either LDS-resident tiles (G transforms per workgroup) or a
// four-step decomposition, builds the tile table sorted largest first, and
// captures the launches into a hipGraph:
//   LDS tiles: one launch of the ragged kernel per workgroup size;
//   four-step classes: columns -> rows -> closing columns (convolutions),
//     each step one launch of the four-step kernels per class, reading the
//     class's per-transform pointer arrays (FFTSET_BATCH_FOUR_STEP_TILES:
//     one launch of the ragged kernel per step over all four-step classes).
// The same kernels launched once per (class, step), or once per transform,
// give the per-size-launch baselines.
//
// A kernel-gradient job (FFTSET_OP_CONV_GRAD_KERNEL) is expanded at commit
// into a get_kernel of its signal into a spectrum row the batch owns
// (stage 0) and a signal-adjoint pass of its output gradient against that
// row (stage 1); stage 1 runs once all of stage 0 has finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "context.h"
#include "fs1.h"
#include "fsr.h"
#include "rcol.h"
#include "planner.h"
#include "refpass.h"

namespace fsa {

struct BatchJob {
    const fftset_fft* p;
    int op;
    float* out;
    const float* in;
    const float* kern;
    int stage = 0;          // 1: runs after every stage-0 job
    bool internal = false;  // part of an expanded kernel-gradient job
    bool expanded = false;  // a kernel-gradient job, run by its two internal jobs
};

// A kernel-gradient job as added (expanded at commit).
constexpr int kOpGradKernelJob = 100;

struct BatchClass {
    const fftset_fft* p;
    int op;
    int stage = 0;
    bool internal = false;
    int spec_src = -1;      // stage-1 kernel-gradient class: the class writing its spectra
    std::vector<int> jobs;  // indices into the batch's job list
    RClass rc;              // device descriptor (pointers patched at upload)
    bool four_step;
    FourStep fs;
    int threads;            // workgroup size of this class's LDS tiles
    int col_threads, row_threads;  // four-step: column and row tiles
    std::vector<int> group_idx;    // this class's entries of group_ranges
    long long pass_off = 0;        // per-pass baseline: offset of its work tiles (float2)
    float* work_base = nullptr;    // four-step: its first work row in the pool (host view)
    bool fsr_ok = false;           // a 240 x 256 four-step class the CU-resident kernel can run
    bool rcol_ok = false;          // an A x 256 class the column-register kernel can run
};

// A contiguous run of tiles launched together with one workgroup size.
struct TileRange {
    int first, count;
    int threads = 256;
    int lds_bytes = 0;
    int counter = -1;  // persistent launches: index of its tile counter
    bool four_step = false;  // four-step tiles (else LDS tiles)
    int cls = -1, job = -1, step = -1;  // per-transform ranges: class, transform, four-step step
};

}  // namespace fsa

using namespace fsa;

struct fftset_batch {
    Ctx* ctx = nullptr;
    int threads = 256;
    bool committed = false;
    bool use_graph = true;
    std::vector<BatchJob> jobs;
    std::vector<BatchClass> classes;
    // device tables
    void* dev_mem = nullptr;
    RClass* d_classes = nullptr;
    RTile* d_tiles = nullptr;        // fused order: phase 1, 2, 3, each sorted largest first
    RTile* d_group_tiles = nullptr;  // per (class, phase), for the per-group baseline
    RTile* d_job_tiles = nullptr;    // per transform, for the per-transform baseline
    float* work_pool = nullptr;
    std::vector<TileRange> lds_r;    // LDS-resident tiles, one range per workgroup size
    std::vector<TileRange> lds_r1;   // stage-1 LDS-resident tiles
    float* spec_pool = nullptr;      // spectra of expanded kernel-gradient jobs
    std::vector<TileRange> fs_r[3];  // four-step steps (columns, rows, closing columns)
    int sched = 0;                   // FFTSET_BATCH_SCHED_*
    bool persistent = false;         // FFTSET_BATCH_PERSISTENT
    bool fs_tiles = false;           // four-step steps as ragged tiles (else the four-step kernels)
    int* d_counters = nullptr;       // one tile counter per fused range
    std::vector<TileRange> group_ranges;
    std::vector<TileRange> job_ranges;
    float2* pass_work = nullptr;     // per-pass baseline work tiles (allocated on first use)
    std::vector<RefPassClass> refpass;  // reference-pass-structure baseline (set up on first use)
    int lds_bytes = 0;  // largest of any range
    long long alg_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    hipEvent_t mid = nullptr;  // after the largest four-step class's column step
};

namespace fsa {

static long long op_bytes(int op, long long L) { return (op == OP_CONV || op == OP_GRAD_IN ? 24 : 16) * L; }

static int map_op(int api_op)
{
    switch (api_op) {
    case FFTSET_OP_FORWARD: return OP_FWD;
    case FFTSET_OP_INVERSE: return OP_INV;
    case FFTSET_OP_GET_KERNEL: return OP_KERN;
    case FFTSET_OP_CONV: return OP_CONV;
    case FFTSET_OP_CONV_GRAD_INPUT: return OP_GRAD_IN;
    default: return -1;
    }
}

static int round_threads(int lanes, int min_threads)
{
    int w = min_threads;
    while (w < lanes && w < 1024) w *= 2;
    return lanes <= w ? w : -1;
}

// Tiles of a class use the plan's own engine, radix chain and lanes per
// transform, so every result (and every kernel-spectrum layout) matches the
// drop-in calls on the same plan. Only the packing (transforms per
// workgroup) and the workgroup size are chosen here.
static int plan_class(fftset_batch* b, BatchClass& k, hipStream_t)
{
    const fftset_fft* p = k.p;
    RClass& r = k.rc;
    std::memset(&r, 0, sizeof(r));
    r.mod = p->mod;
    r.op = k.op;
    r.L = (int)p->lfft;
    r.tmod = p->tmod;
    r.tfo = p->tfo;
    k.four_step = p->engine == ENG_FOUR_STEP;
    if (!k.four_step) {
        if (p->shape.E != 16)
            return FFTSET_E_UNSUPPORTED;
        // the plan's own packing (its autotuned G where the tune table has
        // one, as the uniform launch uses) in the smallest workgroup of 128
        // or more lanes that holds it: a batch of 20000 1440-bin COMPLEX
        // forward transforms, one per 128 lanes, took 147.0 us against 157.8
        // at two per 256. FFTSET_BATCH_THREADS_512: as many transforms as 512
        // lanes hold.
        const int T = p->shape.T;
        const bool wide = b->threads >= 512;
        const long long lds_one = (long long)lds_len(r.L) * 8;
        int G = wide ? std::max(1, round_threads(T, 512) / T) : std::max(1, p->shape.G);
        while (G > 1 && (G * lds_one > 64 * 1024 || G * T > 1024 || G > (int)k.jobs.size())) --G;
        k.threads = round_threads(G * T, wide ? 512 : 128);
        if (k.threads < 0)
            return FFTSET_E_UNSUPPORTED;
        r.T = T;
        r.G = G;
        r.c = p->chain;
        return FFTSET_OK;
    }
    k.fs = p->fs;
    k.fsr_ok = k.fs.fsr && !b->fs_tiles;  // four-step steps as ragged tiles keep the split launches
    k.rcol_ok = k.fs.rcol && !b->fs_tiles;
    if (k.fs.cols.E != 16 || k.fs.rows.E != 16)
        return FFTSET_E_UNSUPPORTED;
    r.A = (int)k.fs.A;
    r.B = (int)k.fs.B;
    r.c = k.fs.colc;
    r.c2 = k.fs.rowc;
    r.tfs = k.fs.tfs;
    r.colT = k.fs.cols.T;
    r.colC = k.fs.cols.G;
    r.colGroups = (r.B + r.colC - 1) / r.colC;
    r.rowT = k.fs.rows.T;
    r.rowR = k.fs.rows.G;
    r.rowGroups = (r.A + r.rowR - 1) / r.rowR;
    // column and row tiles run in workgroups of their own size
    k.col_threads = round_threads(r.colC * r.colT, b->threads);
    k.row_threads = round_threads(r.rowR * r.rowT, b->threads);
    k.threads = std::max(k.col_threads, k.row_threads);
    return k.col_threads < 0 || k.row_threads < 0 ? FFTSET_E_UNSUPPORTED : FFTSET_OK;
}

// Workgroup size and LDS bytes of one tile kind of a class.
static int tile_threads(const BatchClass& k, int kind)
{
    if (kind == TK_LDS)
        return k.threads;
    return kind == TK_ROW ? k.row_threads : k.col_threads;
}

static int tile_lds_bytes(const BatchClass& k, int kind)
{
    const RClass& r = k.rc;
    if (kind == TK_LDS)
        return r.G * lds_len(r.L) * 8;
    return kind == TK_ROW ? r.rowR * lds_len(r.B) * 8 : r.colC * r.A * 8;
}

static int class_lds_bytes(const BatchClass& k)
{
    if (!k.four_step)
        return tile_lds_bytes(k, TK_LDS);
    return std::max(tile_lds_bytes(k, TK_COL), tile_lds_bytes(k, TK_ROW));
}

struct CostTile {
    RTile t;
    long long cost;
};

static int commit(fftset_batch* b, hipStream_t st)
{
    // expand kernel-gradient jobs: all their get_kernel jobs, then all their
    // adjoint passes, appended after the jobs as added
    size_t spec_floats = 0;
    for (const BatchJob& jb : b->jobs)
        if (jb.op == kOpGradKernelJob)
            spec_floats += 2ull * jb.p->lfft;
    if (spec_floats) {
        if (hipMalloc((void**)&b->spec_pool, spec_floats * sizeof(float)) != hipSuccess)
            return FFTSET_E_NOMEM;
        const size_t n0 = b->jobs.size();
        std::vector<BatchJob> second;
        float* sp = b->spec_pool;
        for (size_t j = 0; j < n0; ++j) {
            BatchJob& jb = b->jobs[j];
            if (jb.op != kOpGradKernelJob)
                continue;
            jb.expanded = true;
            // execute()'s convention: in = the signal, kern = the output gradient
            BatchJob a{jb.p, OP_KERN, sp, jb.in, nullptr, 0, true, false};
            BatchJob c{jb.p, OP_GRAD_IN, jb.out, jb.kern, sp, 1, true, false};
            b->jobs.push_back(a);
            second.push_back(c);
            sp += 2ull * jb.p->lfft;
        }
        b->jobs.insert(b->jobs.end(), second.begin(), second.end());
    }
    // classes in first-seen order of (plan, op, stage, internal)
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const BatchJob& jb = b->jobs[j];
        if (jb.expanded) {
            b->alg_bytes += 24LL * jb.p->lfft;  // signal, output gradient, kernel gradient
            continue;
        }
        BatchClass* k = nullptr;
        for (auto& c : b->classes)
            if (c.p == jb.p && c.op == jb.op && c.stage == jb.stage && c.internal == jb.internal) {
                k = &c;
                break;
            }
        if (!k) {
            b->classes.push_back(BatchClass{});
            k = &b->classes.back();
            k->p = jb.p;
            k->op = jb.op;
            k->stage = jb.stage;
            k->internal = jb.internal;
        }
        k->jobs.push_back((int)j);
    }
    // a stage-1 class reads the spectra its stage-0 twin writes, in job order
    for (auto& k : b->classes)
        if (k.stage == 1)
            for (size_t c = 0; c < b->classes.size(); ++c)
                if (b->classes[c].internal && b->classes[c].stage == 0 && b->classes[c].p == k.p)
                    k.spec_src = (int)c;
    size_t work_floats = 0;
    for (auto& k : b->classes) {
        int rc = plan_class(b, k, st);
        if (rc != FFTSET_OK)
            return rc;
        if (class_lds_bytes(k) > kMaxLdsBytes)
            return FFTSET_E_UNSUPPORTED;
        if (k.four_step && (k.op != OP_KERN || k.fsr_ok))
            work_floats += 2ull * k.p->lfft * k.jobs.size();
        if (!k.internal)
            b->alg_bytes += op_bytes(k.op, k.p->lfft) * (long long)k.jobs.size();
    }

    // tiles: fused (by phase, largest first), per group, per transform
    std::vector<CostTile> ph[3], ph1;  // stage 0 by step; stage-1 LDS tiles
    std::vector<RTile> group, per_job;
    // per transform: one tile list per phase (each phase is its own launch)
    std::vector<std::vector<RTile>> job_tiles(3 * b->jobs.size());
    for (int ci = 0; ci < (int)b->classes.size(); ++ci) {
        const BatchClass& k = b->classes[ci];
        const RClass& r = k.rc;
        const int nj = (int)k.jobs.size();
        const long long per = op_bytes(k.op, r.L);
        TileRange gr[3];
        for (int q = 0; q < 3; ++q) gr[q] = {0, 0};
        std::vector<RTile> g3[3];
        if (!k.four_step) {
            for (int j = 0; j < nj; j += r.G) {
                RTile t{ci, TK_LDS, j, std::min(r.G, nj - j)};
                (k.stage ? ph1 : ph[0]).push_back({t, per * t.count});
                g3[0].push_back(t);
            }
            for (int j = 0; j < nj; ++j) job_tiles[3 * k.jobs[j]].push_back(RTile{ci, TK_LDS, j, 1});
        } else {
            const long long col_cost = per * r.colC / std::max(1, r.B);
            const long long row_cost = per * r.rowR / std::max(1, r.A);
            const bool conv = k.op == OP_CONV || k.op == OP_GRAD_IN;
            for (int j = 0; j < nj; ++j) {
                for (int q = 0; q < r.colGroups; ++q) {
                    RTile t{ci, TK_COL, j, q};
                    if (!k.stage)  // stage-1 four-step classes run their own launches
                        ph[0].push_back({t, col_cost});
                    g3[0].push_back(t);
                    job_tiles[3 * k.jobs[j]].push_back(t);
                }
            }
            for (int j = 0; j < nj; ++j) {
                for (int q = 0; q < r.rowGroups; ++q) {
                    RTile t{ci, TK_ROW, j, q};
                    if (!k.stage)
                        ph[1].push_back({t, row_cost});
                    g3[1].push_back(t);
                    job_tiles[3 * k.jobs[j] + 1].push_back(t);
                }
            }
            if (conv) {
                for (int j = 0; j < nj; ++j) {
                    for (int q = 0; q < r.colGroups; ++q) {
                        RTile t{ci, TK_COL2, j, q};
                        if (!k.stage)
                            ph[2].push_back({t, col_cost});
                        g3[2].push_back(t);
                        job_tiles[3 * k.jobs[j] + 2].push_back(t);
                    }
                }
            }
        }
        for (int q = 0; q < 3; ++q) {
            if (g3[q].empty())
                continue;
            const int kind = g3[q][0].kind;
            b->classes[ci].group_idx.push_back((int)b->group_ranges.size());
            b->group_ranges.push_back({(int)group.size(), (int)g3[q].size(), tile_threads(k, kind),
                                       tile_lds_bytes(k, kind), -1, kind != TK_LDS});
            group.insert(group.end(), g3[q].begin(), g3[q].end());
        }
    }
    // fused tables: LDS tiles (all in phase 0) and four-step tiles, each as
    // one range per workgroup size, keeping the largest-first order
    std::vector<RTile> fused;
    size_t placed = 0, wanted = 0;  // every tile must land in a range
    auto add_ranges = [&](const std::vector<CostTile>& src, bool want_lds, std::vector<TileRange>& dst) {
        for (auto& ct : src) wanted += (ct.t.kind == TK_LDS) == want_lds;
        for (int w = 128; w <= 1024; w *= 2) {
            TileRange r{(int)fused.size(), 0, w, 0, -1, !want_lds};
            for (auto& ct : src) {
                const BatchClass& k = b->classes[ct.t.cls];
                if (tile_threads(k, ct.t.kind) != w || (ct.t.kind == TK_LDS) != want_lds)
                    continue;
                fused.push_back(ct.t);
                placed++;
                r.count++;
                r.lds_bytes = std::max(r.lds_bytes, tile_lds_bytes(k, ct.t.kind));
            }
            if (r.count) {
                dst.push_back(r);
                b->lds_bytes = std::max(b->lds_bytes, r.lds_bytes);
            }
        }
    };
    for (int q = 0; q < 3; ++q)
        std::stable_sort(ph[q].begin(), ph[q].end(),
                         [](const CostTile& x, const CostTile& y) { return x.cost > y.cost; });
    std::stable_sort(ph1.begin(), ph1.end(), [](const CostTile& x, const CostTile& y) { return x.cost > y.cost; });
    add_ranges(ph[0], true, b->lds_r);
    add_ranges(ph1, true, b->lds_r1);
    for (int q = 0; q < 3; ++q) add_ranges(ph[q], false, b->fs_r[q]);
    if (placed != wanted)
        return FFTSET_E_UNSUPPORTED;
    for (size_t v = 0; v < job_tiles.size(); ++v) {
        const auto& tl = job_tiles[v];
        TileRange r{(int)per_job.size(), (int)tl.size(), 256, 0};
        if (!tl.empty()) {
            const BatchClass& k = b->classes[tl[0].cls];
            r.threads = tile_threads(k, tl[0].kind);
            r.lds_bytes = tile_lds_bytes(k, tl[0].kind);
            r.four_step = tl[0].kind != TK_LDS;
            r.cls = tl[0].cls;
            r.job = tl[0].job;
            r.step = (int)(v % 3);
        }
        b->job_ranges.push_back(r);
        per_job.insert(per_job.end(), tl.begin(), tl.end());
    }

    // tile counters of the fused ranges (persistent launches)
    int ncounters = 0;
    for (TileRange& r : b->lds_r) r.counter = ncounters++;
    for (TileRange& r : b->lds_r1) r.counter = ncounters++;
    for (int q = 0; q < 3; ++q)
        for (TileRange& r : b->fs_r[q]) r.counter = ncounters++;

    // device memory: classes, three tile tables, per-class pointer arrays.
    // Lay out every region first (16-byte aligned), then allocate.
    size_t off = 0;
    auto place = [&](size_t n) {
        size_t o = off;
        off += (n + 15) & ~size_t(15);
        return o;
    };
    const size_t o_cls = place(sizeof(RClass) * b->classes.size());
    const size_t o_f = place(sizeof(RTile) * fused.size());
    const size_t o_g = place(sizeof(RTile) * group.size());
    const size_t o_j = place(sizeof(RTile) * per_job.size());
    const size_t o_c = place(sizeof(int) * (size_t)std::max(1, ncounters));
    std::vector<size_t> o_ptr;  // four arrays per class: in, out, kern, work
    for (auto& k : b->classes)
        for (int q = 0; q < 4; ++q) o_ptr.push_back(place(sizeof(void*) * k.jobs.size()));
    const size_t bytes = off;
    if (hipMalloc(&b->dev_mem, bytes) != hipSuccess)
        return FFTSET_E_NOMEM;
    if (work_floats && hipMalloc((void**)&b->work_pool, work_floats * sizeof(float)) != hipSuccess)
        return FFTSET_E_NOMEM;
    std::vector<char> host(bytes, 0);
    char* dbase = (char*)b->dev_mem;
    float* wp = b->work_pool;
    std::vector<RClass> cls;
    for (size_t ci = 0; ci < b->classes.size(); ++ci) {
        const BatchClass& k = b->classes[ci];
        RClass r = k.rc;
        const size_t nj = k.jobs.size();
        const size_t o_in = o_ptr[4 * ci], o_out = o_ptr[4 * ci + 1];
        const size_t o_k = o_ptr[4 * ci + 2], o_w = o_ptr[4 * ci + 3];
        if (k.four_step && (k.op != OP_KERN || k.fsr_ok))
            b->classes[ci].work_base = wp;
        const void** hin = (const void**)(host.data() + o_in);
        void** hout = (void**)(host.data() + o_out);
        const void** hk = (const void**)(host.data() + o_k);
        void** hw = (void**)(host.data() + o_w);
        for (size_t j = 0; j < nj; ++j) {
            const BatchJob& jb = b->jobs[k.jobs[j]];
            hin[j] = jb.in;
            hout[j] = jb.out;
            hk[j] = jb.kern;
            hw[j] = nullptr;
            if (k.four_step && (k.op != OP_KERN || k.fsr_ok)) {
                hw[j] = wp;
                wp += 2ull * k.p->lfft;
            }
        }
        r.in = (const float* const*)(dbase + o_in);
        r.out = (float* const*)(dbase + o_out);
        r.kern = (const float* const*)(dbase + o_k);
        r.work = (float* const*)(dbase + o_w);
        cls.push_back(r);
        b->classes[ci].rc = r;  // the four-step launches take the device pointer arrays
    }
    std::memcpy(host.data() + o_cls, cls.data(), sizeof(RClass) * cls.size());
    std::memcpy(host.data() + o_f, fused.data(), sizeof(RTile) * fused.size());
    std::memcpy(host.data() + o_g, group.data(), sizeof(RTile) * group.size());
    std::memcpy(host.data() + o_j, per_job.data(), sizeof(RTile) * per_job.size());
    if (hipMemcpyAsync(b->dev_mem, host.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return FFTSET_E_DEVICE;
    b->d_classes = (RClass*)(dbase + o_cls);
    b->d_tiles = (RTile*)(dbase + o_f);
    b->d_group_tiles = (RTile*)(dbase + o_g);
    b->d_job_tiles = (RTile*)(dbase + o_j);
    b->d_counters = (int*)(dbase + o_c);
    return FFTSET_OK;
}

static int launch_range(const fftset_batch* b, const RTile* tiles, TileRange r, hipStream_t st)
{
    if (r.count <= 0)
        return FFTSET_OK;
    RaggedArgs a{b->d_classes, tiles + r.first, r.count,
                 b->persistent && r.counter >= 0 && tiles == b->d_tiles ? b->d_counters + r.counter : nullptr,
                 r.four_step ? 1 : 0};
    return launch_ragged(a, r.threads, r.lds_bytes, st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
}

static int launch_list(const fftset_batch* b, const std::vector<TileRange>& rs, hipStream_t st)
{
    for (const TileRange& r : rs) {
        int rc = launch_range(b, b->d_tiles, r, st);
        if (rc != FFTSET_OK)
            return rc;
    }
    return FFTSET_OK;
}

// Step q (0 columns, 1 rows, 2 closing columns) of a four-step class through
// the drop-in four-step kernels, which read the class's per-transform pointer
// arrays: the same kernels, tile shapes and data order as run_four_step.
// j0 / n: a sub-range of the class's transforms (n < 0: all of them).
// True when the class runs the one-workgroup-per-transform kernel.
static bool fs1_class(const BatchClass& k)
{
    return k.four_step && k.op != OP_KERN && four_step_fused() && !persistent_mode() && !generic_forced() &&
           fs1_matches(k.fs.colc, k.fs.rowc, k.fs.cols.T, k.fs.rows.T);
}

// True when the class runs the column-register kernel (one launch, step 0).
static bool rcol_class(const BatchClass& k)
{
    return k.rcol_ok && !generic_forced() && !persistent_mode() && !four_step_fused() && !four_step_split();
}

// True when the class runs the CU-resident kernel (one launch, step 0).
static bool fsr_class(const BatchClass& k)
{
    return k.fsr_ok && fsr_op(k.rc.mod, k.op) && !generic_forced() && !persistent_mode() && !four_step_fused() &&
           !four_step_split();
}

static int launch_fs_class(const BatchClass& k, int q, hipStream_t st, long long j0 = 0, long long n = -1)
{
    RClass r = k.rc;
    r.in += j0;
    r.out += j0;
    r.kern += j0;
    r.work += j0;
    const FourStep& fs = k.fs;
    const bool fo = r.mod == MOD_FO;
    const bool conv = r.op == OP_CONV || r.op == OP_GRAD_IN;
    if (q == 2 && !conv)
        return FFTSET_OK;
    FsArgs a{};
    a.tmod = r.tmod;
    a.tfs = r.tfs;
    a.L = r.L;
    a.A = r.A;
    a.B = r.B;
    a.ntrans = n < 0 ? (long long)k.jobs.size() - j0 : n;
    a.conj_kern = r.op == OP_GRAD_IN;
    if (fs1_class(k)) {
        // one workgroup per transform, all steps in one launch (step 0);
        // its scratch slots are the class's first work rows
        if (q != 0)
            return FFTSET_OK;
        Fs1Args f{};
        f.colc = fs.colc;
        f.rowc = fs.rowc;
        f.tmod = r.tmod;
        f.tfs = fs.tfs;
        f.tfs_fo = fs.tfs_fo;
        f.trow = fs.trow_fo;
        f.ntrans = a.ntrans;
        f.conj_kern = a.conj_kern;
        f.in_p = r.in;
        f.out_p = r.out;
        f.kern_p = conv ? r.kern : nullptr;
        f.scratch = (float2*)(k.work_base + 2LL * r.L * j0);
        count_launch(true);
        const int opk = r.op == OP_FWD ? FS1_FWD : r.op == OP_INV ? FS1_INV : FS1_CONV;
        return launch_fs1(r.mod, opk, f, fs1_grid(a.ntrans), st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
    }
    if (rcol_class(k)) {
        // columns in registers, rows through LDS: all steps in one launch
        if (q != 0)
            return FFTSET_OK;
        RcolArgs f{};
        f.tw_row = fs.rowc.p[0].tw;
        f.tfs = fs.tfs;
        f.tfs_in = fo ? fs.tfs_fo : fs.tfs;
        f.tmod = r.tmod;
        f.ntrans = a.ntrans;
        f.conj_kern = a.conj_kern;
        f.in_p = r.in;
        f.out_p = r.out;
        f.kern_p = conv ? r.kern : nullptr;
        count_launch(true);
        const int opk = r.op == OP_FWD ? RC_FWD : r.op == OP_INV ? RC_INV : r.op == OP_KERN ? RC_KERN : RC_CONV;
        return launch_rcol(r.A, r.mod, opk, f, st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
    }
    if (fsr_class(k)) {
        // the whole transform on one CU, all steps in one launch (step 0);
        // its spill rows are the start of each transform's work row
        if (q != 0)
            return FFTSET_OK;
        FsrArgs f{};
        f.tw_col = fs.colc.p[0].tw;
        f.tw_close = fs.tw_close;
        f.tw_row = fs.rowc.p[0].tw;
        f.tfs = fs.tfs;
        f.tfs_in = fo ? fs.tfs_fo : fs.tfs;
        f.trow = fs.trow_fo;
        f.tmod = r.tmod;
        f.ntrans = a.ntrans;
        f.conj_kern = a.conj_kern;
        f.in_p = r.in;
        f.out_p = r.out;
        f.kern_p = conv ? r.kern : nullptr;
        f.spill_p = r.work;
        count_launch(true);
        const int opk = r.op == OP_FWD ? FSR_FWD : r.op == OP_INV ? FSR_INV : r.op == OP_KERN ? FSR_KERN : FSR_CONV;
        return launch_fsr(r.mod, opk, f, st) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
    }
    const ShapeD& sh = q == 1 ? fs.rows : fs.cols;
    a.c = q == 1 ? fs.rowc : fs.colc;
    a.C = sh.G;
    a.T = sh.T;
    a.groups = q == 1 ? (r.A + a.C - 1) / a.C : (r.B + a.C - 1) / a.C;
    hipError_t e;
    if (q == 0) {
        a.in_p = r.in;
        a.out_p = r.op == OP_KERN ? r.out : r.work;
        const int load = r.op == OP_INV ? (fo ? CL_FOBINS : CL_CXCONJ) : (fo ? CL_FOMODC : CL_CX);
        if (load == CL_FOMODC) {  // folded modulation, as run_four_step
            a.tfs = fs.tfs_fo;
            a.trow = fs.trow_fo;
        }
        e = launch_cols(load, CS_WORKTW, a, sh, st);
    } else if (q == 1) {
        a.in_p = r.op == OP_KERN ? (const float* const*)r.out : (const float* const*)r.work;
        if (conv) {
            a.out_p = r.work;
            a.kern_p = r.kern;
            e = launch_rows(RK_CONV, RS_NONE, a, sh, st);
        } else if (r.op == OP_KERN) {
            a.out_p = r.out;
            e = launch_rows(RK_KERN, RS_NONE, a, sh, st);
        } else {
            a.out_p = r.out;
            const int store = r.op == OP_FWD ? (fo ? RS_FOBINS : RS_CX) : (fo ? RS_FODEMOD : RS_CXCONJ);
            e = launch_rows(RK_STOCK, store, a, sh, st);
        }
    } else {
        a.in_p = (const float* const*)r.work;
        a.out_p = r.out;
        e = launch_cols(CL_CX, fo ? CS_FODEMOD : CS_CXCONJ, a, sh, st);
    }
    return e == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
}

// Four-step step q of every four-step class, largest class first.
static int launch_fs_step(const fftset_batch* b, int q, hipStream_t st)
{
    if (b->fs_tiles)
        return launch_list(b, b->fs_r[q], st);
    for (const BatchClass& k : b->classes) {
        if (!k.four_step || k.stage)
            continue;
        if (int rc = launch_fs_class(k, q, st))
            return rc;
    }
    return FFTSET_OK;
}

// Four-step classes, largest (algorithmic bytes) first.
static std::vector<const BatchClass*> four_step_classes(const fftset_batch* b)
{
    std::vector<const BatchClass*> fsc;
    for (const BatchClass& k : b->classes)
        if (k.four_step && !k.stage)
            fsc.push_back(&k);
    std::stable_sort(fsc.begin(), fsc.end(), [](const BatchClass* x, const BatchClass* y) {
        return op_bytes(x->op, x->rc.L) * (long long)x->jobs.size() >
               op_bytes(y->op, y->rc.L) * (long long)y->jobs.size();
    });
    return fsc;
}

// Launch schedules (the graph records the same structure):
//  BRANCHES (default): the largest four-step class's chain (columns ->
//    rows -> closing columns) on the main stream, the other four-step
//    classes' chains on side branches from the start, and the LDS groups on
//    one more side branch (ascending workgroup size) once the main chain's
//    column step is done, so they share the GPU with the row and
//    closing-column steps rather than with the column steps. When the
//    largest class runs the CU-resident kernel (one launch, no column step),
//    every LDS group gets a side branch of its own from the start. With
//    FFTSET_BATCH_FOUR_STEP_TILES: the LDS groups on side branches from the
//    start, one chain of tile launches on the main stream;
//  PHASED: the LDS groups and the four-step column step run concurrently,
//    then the rows, then the closing columns;
//  SERIAL: every group one after the other, largest workgroups first.
// Without side streams (direct launches) every schedule runs serially.
static int launch_stage0(const fftset_batch* b, hipStream_t st, bool branches)
{
    const int sched = branches ? b->sched : FFTSET_BATCH_SCHED_SERIAL;
    if (sched == FFTSET_BATCH_SCHED_SERIAL) {
        for (int i = (int)b->lds_r.size() - 1; i >= 0; --i)
            if (int rc = launch_range(b, b->d_tiles, b->lds_r[i], st))
                return rc;
        for (int q = 0; q < 3; ++q)
            if (int rc = launch_fs_step(b, q, st))
                return rc;
        return FFTSET_OK;
    }
    if (hipEventRecord(b->fork, st) != hipSuccess)
        return FFTSET_E_DEVICE;
    bool forked[3] = {false, false, false};
    auto side = [&](int i) -> hipStream_t {
        if (!forked[i]) {
            if (hipStreamWaitEvent(b->side[i], b->fork, 0) != hipSuccess)
                return nullptr;
            forked[i] = true;
        }
        return b->side[i];
    };
    auto join_all = [&]() {
        for (int i = 0; i < 3; ++i)
            if (forked[i] && (hipEventRecord(b->join[i], b->side[i]) != hipSuccess ||
                              hipStreamWaitEvent(st, b->join[i], 0) != hipSuccess))
                return FFTSET_E_DEVICE;
        return FFTSET_OK;
    };
    // LDS groups from side branch `first`: one branch each (spread), or all
    // on that one branch in ascending workgroup size, after event `after`
    auto lds_from = [&](int first, hipEvent_t after, bool spread) {
        for (int i = 0; i < (int)b->lds_r.size(); ++i) {
            hipStream_t s = side(spread ? (first + i) % 3 : first);
            if (!s || ((spread || i == 0) && after && hipStreamWaitEvent(s, after, 0) != hipSuccess))
                return FFTSET_E_DEVICE;
            if (int rc = launch_range(b, b->d_tiles, b->lds_r[i], s))
                return rc;
        }
        return FFTSET_OK;
    };
    if (sched == FFTSET_BATCH_SCHED_PHASED || b->fs_tiles) {
        if (int rc = lds_from(0, nullptr, true))
            return rc;
        if (int rc = launch_fs_step(b, 0, st))
            return rc;
        if (sched == FFTSET_BATCH_SCHED_PHASED)
            if (int rc = join_all())
                return rc;
        for (int q = 1; q < 3; ++q)
            if (int rc = launch_fs_step(b, q, st))
                return rc;
        return join_all();
    }
    const std::vector<const BatchClass*> fsc = four_step_classes(b);
    for (size_t c = 1; c < fsc.size(); ++c) {
        hipStream_t s = side((int)((c - 1) % 3));
        if (!s)
            return FFTSET_E_DEVICE;
        for (int q = 0; q < 3; ++q)
            if (int rc = launch_fs_class(*fsc[c], q, s))
                return rc;
    }
    hipEvent_t after = nullptr;
    if (!fsc.empty()) {
        if (int rc = launch_fs_class(*fsc[0], 0, st))
            return rc;
        if (!fs1_class(*fsc[0]) && !fsr_class(*fsc[0]) && !rcol_class(*fsc[0])) {  // one launch: nothing to wait for
            if (hipEventRecord(b->mid, st) != hipSuccess)
                return FFTSET_E_DEVICE;
            after = b->mid;
        }
    }
    // the side branch after the other chains' (shared only beyond three
    // classes); beside CU-resident four-step kernels (nothing to wait for)
    // each LDS group takes a branch of its own: measured 1.5% faster
    const bool spread = !fsc.empty() && !after;
    if (int rc = lds_from(std::min<int>(std::max<int>((int)fsc.size() - 1, 0), 2), after, spread))
        return rc;
    if (!fsc.empty())
        for (int q = 1; q < 3; ++q)
            if (int rc = launch_fs_class(*fsc[0], q, st))
                return rc;
    return join_all();
}

// Stage 1 (the adjoint passes of kernel-gradient jobs) after all of stage
// 0: its LDS tiles, then its four-step classes, on the launch stream.
static int launch_fused(const fftset_batch* b, hipStream_t st, bool branches)
{
    if (int rc = launch_stage0(b, st, branches))
        return rc;
    if (int rc = launch_list(b, b->lds_r1, st))
        return rc;
    for (const BatchClass& k : b->classes) {
        if (!k.four_step || !k.stage)
            continue;
        for (int q = 0; q < 3; ++q)
            if (int rc = launch_fs_class(k, q, st))
                return rc;
    }
    return FFTSET_OK;
}

}  // namespace fsa

extern "C" {

struct fftset_batch* fftset_batch_create(struct fftset* fc)
{
    if (!fc || !fc->impl) {
        set_error(FFTSET_E_INVALID);
        return nullptr;
    }
    fftset_batch* b = new (std::nothrow) fftset_batch();
    if (!b) {
        set_error(FFTSET_E_NOMEM);
        return nullptr;
    }
    b->ctx = (Ctx*)fc->impl;
    return b;
}

int fftset_batch_add(struct fftset_batch* b, const struct fftset_fft* fft, int op, float* out, const float* in,
                     const float* kern)
{
    if (!b || !fft || b->committed || fft->ctx != b->ctx) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    const int iop = op == FFTSET_OP_CONV_GRAD_KERNEL ? kOpGradKernelJob : map_op(op);
    if (iop < 0) {
        set_error(FFTSET_E_UNSUPPORTED);
        return FFTSET_E_UNSUPPORTED;
    }
    const bool needs_kern = iop == OP_CONV || iop == OP_GRAD_IN || iop == kOpGradKernelJob;
    if (!out || !in || (needs_kern && !kern) || (((uintptr_t)out | (uintptr_t)in | (uintptr_t)kern) & 7) ||
        (iop == OP_KERN && (const float*)out == in)) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    b->jobs.push_back(BatchJob{fft, iop, out, in, needs_kern ? kern : nullptr});
    return (int)b->jobs.size() - 1;
}

int fftset_batch_commit(struct fftset_batch* b, unsigned flags)
{
    if (!b || b->committed) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    int rc = use_device(b->ctx);
    if (rc != FFTSET_OK) {
        set_error(rc);
        return rc;
    }
    b->threads = (flags & FFTSET_BATCH_THREADS_512) ? 512 : 256;
    b->sched = (flags & FFTSET_BATCH_SCHED_SERIAL) ? FFTSET_BATCH_SCHED_SERIAL
               : (flags & FFTSET_BATCH_SCHED_PHASED) ? FFTSET_BATCH_SCHED_PHASED
                                                     : 0;
    b->use_graph = !(flags & FFTSET_BATCH_NO_GRAPH);
    b->persistent = (flags & FFTSET_BATCH_PERSISTENT) != 0;
    b->fs_tiles = (flags & FFTSET_BATCH_FOUR_STEP_TILES) != 0;
    hipStream_t st = thread_stream(b->ctx->device);
    rc = commit(b, st);
    if (rc == FFTSET_OK && b->use_graph) {
        // capture the three phase launches once; launches replay the graph
        bool ok = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&b->side[0], hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&b->side[1], hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&b->side[2], hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&b->join[0], hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&b->join[1], hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&b->join[2], hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&b->mid, hipEventDisableTiming) == hipSuccess;
        if (!ok || hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = FFTSET_E_DEVICE;
        } else {
            int lrc = launch_fused(b, b->cap_stream, true);
            hipError_t e = hipStreamEndCapture(b->cap_stream, &b->graph);
            if (lrc != FFTSET_OK || e != hipSuccess ||
                hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
                rc = FFTSET_E_DEVICE;
        }
    }
    b->committed = rc == FFTSET_OK;
    set_error(rc);
    return rc;
}

int fftset_batch_launch(struct fftset_batch* b, void* stream)
{
    if (!b || !b->committed) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    int rc = use_device(b->ctx);
    if (rc == FFTSET_OK) {
        if (b->exec)
            rc = hipGraphLaunch(b->exec, (hipStream_t)stream) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
        else
            rc = launch_fused(b, (hipStream_t)stream, false);
    }
    set_error(rc);
    return rc;
}

int fftset_batch_launch_baseline(struct fftset_batch* b, int mode, void* stream)
{
    if (!b || !b->committed) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    int rc = use_device(b->ctx);
    if (rc != FFTSET_OK) {
        set_error(rc);
        return rc;
    }
    hipStream_t st = (hipStream_t)stream;
    if (mode == FFTSET_BASELINE_PER_GROUP) {
        // one launch per (class, step), class by class
        for (size_t ci = 0; ci < b->classes.size() && rc == FFTSET_OK; ++ci) {
            const BatchClass& k = b->classes[ci];
            if (k.four_step && !b->fs_tiles) {
                for (int q = 0; q < 3 && rc == FFTSET_OK; ++q) rc = launch_fs_class(k, q, st);
                continue;
            }
            for (int g : k.group_idx)
                if ((rc = launch_range(b, b->d_group_tiles, b->group_ranges[g], st)) != FFTSET_OK)
                    break;
        }
    } else if (mode == FFTSET_BASELINE_PER_PASS) {
        if (!b->pass_work) {
            long long n = 0;
            for (auto& k : b->classes)
                if (!k.four_step) {
                    k.pass_off = n;
                    n += (long long)k.rc.L * (long long)k.jobs.size();
                }
            if (n > 0 && hipMalloc((void**)&b->pass_work, (size_t)n * sizeof(float2)) != hipSuccess) {
                b->pass_work = nullptr;
                set_error(FFTSET_E_NOMEM);
                return FFTSET_E_NOMEM;
            }
        }
        for (size_t ci = 0; ci < b->classes.size() && rc == FFTSET_OK; ++ci) {
            const BatchClass& k = b->classes[ci];
            if (k.four_step) {
                if (!b->fs_tiles) {
                    for (int q = 0; q < 3 && rc == FFTSET_OK; ++q) rc = launch_fs_class(k, q, st);
                    continue;
                }
                for (int g : k.group_idx)
                    if ((rc = launch_range(b, b->d_group_tiles, b->group_ranges[g], st)) != FFTSET_OK)
                        break;
                continue;
            }
            const int steps = pass_steps(k.op, k.rc.c.np);
            for (int s = 0; s < steps && rc == FFTSET_OK; ++s)
                rc = launch_pass(b->d_classes + ci, s, (int)k.jobs.size(), k.rc.T, b->pass_work + k.pass_off, st) ==
                             hipSuccess
                         ? FFTSET_OK
                         : FFTSET_E_DEVICE;
        }
    } else if (mode == FFTSET_BASELINE_REFERENCE_PASSES) {
        if (b->refpass.empty()) {
            for (const BatchClass& k : b->classes) {
                RefPassClass r;
                float2* from = k.spec_src >= 0 ? b->refpass[k.spec_src].work : nullptr;
                rc = refpass_setup(b->ctx, k.p, k.op, k.rc, (long long)k.jobs.size(), &r, st, from);
                if (rc != FFTSET_OK)
                    break;
                b->refpass.push_back(r);
            }
            if (rc != FFTSET_OK) {
                for (RefPassClass& r : b->refpass) refpass_free(&r);
                b->refpass.clear();
                set_error(rc);
                return rc;
            }
        }
        for (const RefPassClass& r : b->refpass)
            if ((rc = refpass_run(r, st)) != FFTSET_OK)
                break;
    } else if (mode == FFTSET_BASELINE_PER_TRANSFORM) {
        // one launch per transform and step (job_ranges holds 3 per transform)
        for (const TileRange& r : b->job_ranges) {
            if (r.four_step && !b->fs_tiles)
                rc = launch_fs_class(b->classes[r.cls], r.step, st, r.job, 1);
            else
                rc = launch_range(b, b->d_job_tiles, r, st);
            if (rc != FFTSET_OK)
                break;
        }
    } else {
        rc = FFTSET_E_INVALID;
    }
    set_error(rc);
    return rc;
}

long long fftset_batch_bytes(const struct fftset_batch* b) { return b ? b->alg_bytes : 0; }

int fftset_batch_info(const struct fftset_batch* b, int* ntiles, int* nclasses, int* lds_bytes, int* four_step_classes)
{
    if (!b || !b->committed)
        return FFTSET_E_INVALID;
    if (ntiles) {
        int n = 0;
        for (auto& r : b->lds_r) n += r.count;
        for (auto& r : b->lds_r1) n += r.count;
        for (int q = 0; q < 3; ++q)
            for (auto& r : b->fs_r[q]) n += r.count;
        *ntiles = n;
    }
    if (nclasses)
        *nclasses = (int)b->classes.size();
    if (lds_bytes)
        *lds_bytes = b->lds_bytes;
    if (four_step_classes) {
        int n = 0;
        for (auto& k : b->classes) n += k.four_step;
        *four_step_classes = n;
    }
    return FFTSET_OK;
}

void fftset_batch_destroy(struct fftset_batch* b)
{
    if (!b)
        return;
    (void)use_device(b->ctx);
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    if (b->cap_stream)
        (void)hipStreamDestroy(b->cap_stream);
    for (int i = 0; i < 3; ++i) {
        if (b->side[i])
            (void)hipStreamDestroy(b->side[i]);
        if (b->join[i])
            (void)hipEventDestroy(b->join[i]);
    }
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    if (b->mid)
        (void)hipEventDestroy(b->mid);
    if (b->dev_mem || b->work_pool || b->pass_work || b->spec_pool || !b->refpass.empty())
        (void)hipDeviceSynchronize();
    for (RefPassClass& r : b->refpass) refpass_free(&r);
    if (b->dev_mem)
        (void)hipFree(b->dev_mem);
    if (b->work_pool)
        (void)hipFree(b->work_pool);
    if (b->pass_work)
        (void)hipFree(b->pass_work);
    if (b->spec_pool)
        (void)hipFree(b->spec_pool);
    delete b;
}

}  // extern "C"
