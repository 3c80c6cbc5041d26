// refpass.cpp -- the per-size-launch baseline with the reference's pass
// structure: planning (the reference's vector widths and min-cost chains)
// and the launch sequence of one batch class.
#include "refpass.h"

#include <vector>

#include "planner.h"

namespace fsa {

// ------------------------------------------------------------ the planner
// FFTSET_FLOAT_PASSES at one vector width: radices 2, 3, 4 at any position,
// 5, 6, 8, 16 only as the final pass (src/fftset_vec.c:944-971). A pass of
// radix r costs 100 + ((r + 2)(vw + 3) 10^4) / (r vw) in integer
// arithmetic (src/fftset_vec.c:1023); the search keeps the cheapest chain,
// replacing the best so far only when strictly cheaper, in table order
// (src/fftset_vec.c:1004-1062).
static const int kRefRadices[7] = {2, 3, 4, 5, 6, 8, 16};

static unsigned ref_cost(unsigned r, unsigned vw) { return 100u + ((r + 2u) * (vw + 3u) * 10000u) / (r * vw); }

static bool ref_search(unsigned length, unsigned vw, unsigned* cost, std::vector<int>* chain)
{
    *cost = 0;
    chain->clear();
    for (int r : kRefRadices) {
        const unsigned c = ref_cost((unsigned)r, vw);
        if (length == (unsigned)r) {
            if (*cost == 0 || *cost > c) {
                *cost = c;
                *chain = {r};
            }
            continue;
        }
        if (r > 4 || length % (unsigned)r)
            continue;
        unsigned sub_cost;
        std::vector<int> sub;
        if (!ref_search(length / (unsigned)r, vw, &sub_cost, &sub))
            continue;
        if (*cost == 0 || *cost > sub_cost + c) {
            *cost = sub_cost + c;
            chain->assign(1, r);
            chain->insert(chain->end(), sub.begin(), sub.end());
        }
    }
    return *cost != 0;
}

std::vector<int> reference_chain(unsigned length, unsigned vw)
{
    unsigned cost;
    std::vector<int> chain;
    if (length < 2 || vw == 0 || !ref_search(length, vw, &cost, &chain))
        chain.clear();
    return chain;
}

bool reference_structure(int mod, unsigned L, RefStructure* out)
{
    RefStructure s;
    if (mod == MOD_FO) {
        // src/fftset_mod_freqoffsetreal.c:1014-1023 (v8f), :1073 (v4f), :1119 (v1f)
        if (L % 32 == 0) {
            s.vw = 8;
            s.split = {4, 2};
        } else if (L % 16 == 0) {
            s.vw = 4;
            s.split = {4};
        }
    } else if (L > 4 && L % 4 == 0) {  // src/fftset_mod_cplx.c:322-331
        s.vw = 4;
        s.split = {4};
    }
    s.inner = reference_chain(L / s.vw, s.vw);
    if (s.inner.empty())
        return false;
    *out = s;
    return true;
}

// ------------------------------------------------------ spectrum orders
// Frequency held at position p after in-place DIF passes with `radices`
// over length n: the first pass's output k heads sub-block k (positions
// k m + p', m = n / R), whose sub-transform yields bins k + R k'.
static long long dif_freq(const std::vector<int>& radices, long long n, long long p)
{
    long long f = 0, mult = 1;
    for (int R : radices) {
        const long long m = n / R;
        f += (p / m) * mult;
        p %= m;
        mult *= R;
        n = m;
    }
    return f;
}

static std::vector<int> radices_of(const ChainD& c)
{
    std::vector<int> r;
    for (int i = 0; i < c.np; ++i) r.push_back(c.p[i].R);
    return r;
}

// Frequency of element `idx` of one of the plan's kernel spectra: the LDS
// engine stores value k of last-pass block b at k s + b (slot-major, block
// b at DIF position b R + k); four-step plans store row k1 the same way with
// the row chain, holding bins k1 + A k2.
static bool lib_kern_freqs(const fftset_fft* p, std::vector<long long>* freq)
{
    const long long L = p->lfft;
    freq->assign((size_t)L, -1);
    if (p->engine == ENG_LDS) {
        const ChainD& c = p->chain;
        const std::vector<int> r = radices_of(c);
        const int R = c.p[c.np - 1].R;
        const long long s = L / R;
        for (long long idx = 0; idx < L; ++idx) {
            const long long k = idx / s, b = idx % s;
            (*freq)[(size_t)idx] = dif_freq(r, L, b * R + k);
        }
        return true;
    }
    const FourStep& fs = p->fs;
    const std::vector<int> r = radices_of(fs.rowc);
    const long long A = fs.A, B = fs.B;
    const int R = fs.rowc.p[fs.rowc.np - 1].R;
    const long long s = B / R;
    for (long long idx = 0; idx < L; ++idx) {
        const long long k1 = idx / B, q = idx % B;
        const long long k = q / s, b = q % s;
        (*freq)[(size_t)idx] = k1 + A * dif_freq(r, B, b * R + k);
    }
    return true;
}

// ------------------------------------------------------------ set-up / run
int refpass_setup(Ctx* ctx, const fftset_fft* p, int op, const RClass& rc, long long ntrans, RefPassClass* out,
                  hipStream_t st, float2* kref_from)
{
    RefStructure rs;
    const unsigned L = p->lfft;
    if (!reference_structure(p->mod, L, &rs))
        return FFTSET_E_UNSUPPORTED;
    RefPassClass k;
    k.mod = p->mod;
    k.op = op;
    k.L = (int)L;
    k.ntrans = ntrans;
    k.V = rs.vw;
    k.nsplit = (int)rs.split.size();
    std::vector<int> full = rs.split;
    full.insert(full.end(), rs.inner.begin(), rs.inner.end());
    if (!build_chain(full, L, &k.c) || !attach_twiddles(ctx, &k.c, st))
        return FFTSET_E_NOMEM;
    k.tmod = p->tmod;
    k.in = rc.in;
    k.out = rc.out;
    // lane slot q of the split holds the bins f = V k' + resid[q]: slot
    // q = 2 r + s of the width-8 split ([4, 2]) holds residue r + 4 s
    for (unsigned q = 0; q < k.V; ++q) {
        const unsigned l = k.V == 8 ? (q / 2) + 4 * (q % 2) : q;
        k.resid[q] = (int)l;
        k.slot[l] = (int)q;
    }
    const size_t bytes = sizeof(float2) * (size_t)ntrans * L;
    const bool conv = op == OP_CONV || op == OP_GRAD_IN;
    const bool stock = op == OP_FWD || op == OP_INV;
    if (hipMalloc((void**)&k.work, bytes) != hipSuccess)
        return FFTSET_E_NOMEM;
    if (stock && hipMalloc((void**)&k.work2, bytes) != hipSuccess) {
        refpass_free(&k);
        return FFTSET_E_NOMEM;
    }
    if (conv && kref_from) {
        k.kref = kref_from;
        k.own_kref = false;
    } else if (conv) {
        // the class's kernel spectra (library order) in this structure's
        // DIF order: kref[t][pos] = K_t[perm[pos]]
        std::vector<long long> lib;
        lib_kern_freqs(p, &lib);
        std::vector<int> where((size_t)L, -1), perm((size_t)L);
        for (long long idx = 0; idx < (long long)L; ++idx) where[(size_t)lib[(size_t)idx]] = (int)idx;
        for (long long pos = 0; pos < (long long)L; ++pos) perm[(size_t)pos] = where[(size_t)dif_freq(full, L, pos)];
        int* d_perm = nullptr;
        RpArgs a{};
        a.L = (int)L;
        a.ntrans = ntrans;
        a.in = (const float* const*)rc.kern;
        if (hipMalloc((void**)&k.kref, bytes) != hipSuccess || hipMalloc((void**)&d_perm, sizeof(int) * L) != hipSuccess) {
            refpass_free(&k);
            return FFTSET_E_NOMEM;
        }
        a.work = k.kref;
        const bool ok = hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * L, hipMemcpyHostToDevice, st) == hipSuccess &&
                        launch_refpass_gather(a, d_perm, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        (void)hipFree(d_perm);
        if (!ok) {
            refpass_free(&k);
            return FFTSET_E_DEVICE;
        }
    }
    *out = k;
    return FFTSET_OK;
}

void refpass_free(RefPassClass* k)
{
    if (!k->own_kref)
        k->kref = nullptr;
    for (float2** b : {&k->work, &k->work2, &k->kref}) {
        if (*b)
            (void)hipFree(*b);
        *b = nullptr;
    }
}

// The sweeps, one launch each (see kernels_refpass.hip):
//   forward:      split DIF passes (the first reads the input, modulated) or a
//                 modulation sweep at width 1, the inner chain's Stockham
//                 passes in every lane, the reorder sweep;
//   inverse:      the de-interleave sweep, Stockham passes, the split's DIT
//                 passes (the last writes the output, demodulated or
//                 conjugated) or a demodulation / conjugation sweep;
//   get_kernel:   DIF passes (the spectrum stays in the baseline's own order);
//   convolution:  DIF passes, the product sweep, DIT passes, as the inverse's
//                 last sweeps.
int refpass_run(const RefPassClass& k, hipStream_t st)
{
    const bool fo = k.mod == MOD_FO;
    const int np = k.c.np;
    RpArgs a{};
    a.L = k.L;
    a.V = (int)k.V;
    a.ntrans = k.ntrans;
    a.tmod = k.tmod;
    a.in = k.in;
    a.out = k.out;
    a.work = k.work;
    a.work2 = k.work2;
    a.kref = k.kref;
    a.conj_kern = k.op == OP_GRAD_IN;
    for (int i = 0; i < 8; ++i) {
        a.slot[i] = k.slot[i];
        a.resid[i] = k.resid[i];
    }
    auto pass = [&](int i) {
        a.R = k.c.p[i].R;
        a.m = k.c.p[i].m;
        a.s = k.c.p[i].s;
        a.tw = k.c.p[i].tw;
    };
    auto go = [&](int kind, int R, int io) { return launch_refpass(kind, R, io, a, st) == hipSuccess; };
    const int in_io = fo ? RP_IN_FO : RP_IN_CX;
    const int out_io = fo ? RP_OUT_FO : RP_OUT_CX;
    // whole-transform sweeps of radix 1 (modulation, demodulation, conj)
    auto sweep1 = [&](int kind, int io) {
        a.R = 1;
        a.m = k.L;
        a.s = 1;
        a.tw = nullptr;
        return go(kind, 1, io);
    };
    // the first DIF passes: from the input, modulated
    auto dif_all = [&]() {
        int first = 0;
        if (k.nsplit == 0 && fo) {
            if (!sweep1(RPK_DIF, RP_IN_FO))
                return false;
        } else {
            pass(0);
            if (!go(RPK_DIF, a.R, in_io))
                return false;
            first = 1;
        }
        for (int i = first; i < np; ++i) {
            pass(i);
            if (!go(RPK_DIF, a.R, RP_WORK))
                return false;
        }
        return true;
    };
    bool ok = true;
    switch (k.op) {
    case OP_FWD: {
        const bool direct = k.V == 1 && !fo;  // width-1 COMPLEX: no first or last sweep
        if (k.nsplit > 0) {
            pass(0);
            ok = go(RPK_DIF, a.R, in_io);
            for (int i = 1; ok && i < k.nsplit; ++i) {
                pass(i);
                ok = go(RPK_DIF, a.R, RP_WORK);
            }
        } else if (fo) {
            ok = sweep1(RPK_DIF, RP_IN_FO);
        }
        for (int i = k.nsplit; ok && i < np; ++i) {
            pass(i);
            a.s = k.c.p[i].s / (int)k.V;  // blocks within one lane
            int io = RP_WORK;
            if (direct)
                io = np == 1 ? RP_IN_OUT_CX : i == 0 ? RP_IN_CX : i == np - 1 ? RP_OUT_CX : RP_WORK;
            ok = go(RPK_STOCK, a.R, io);
            if (io != RP_IN_CX && io != RP_IN_OUT_CX)
                std::swap(a.work, a.work2);
            else
                a.work = k.work2, a.work2 = k.work;
        }
        if (ok && !direct)
            ok = go(RPK_INTER, 0, out_io);
        break;
    }
    case OP_INV: {
        ok = go(RPK_DEINTER, 0, in_io);
        for (int i = k.nsplit; ok && i < np; ++i) {
            pass(i);
            a.s = k.c.p[i].s / (int)k.V;
            ok = go(RPK_STOCK, a.R, RP_WORK);
            std::swap(a.work, a.work2);
        }
        if (k.nsplit > 0) {
            for (int i = k.nsplit - 1; ok && i >= 0; --i) {
                pass(i);
                ok = go(RPK_DIT, a.R, i == 0 ? out_io : RP_WORK);
            }
        } else if (ok) {
            ok = sweep1(RPK_DIT, out_io);
        }
        break;
    }
    case OP_KERN: ok = dif_all(); break;
    case OP_CONV:
    case OP_GRAD_IN: {
        ok = dif_all() && go(RPK_MULCONJ, 0, RP_WORK);
        for (int i = np - 1; ok && i >= 0; --i) {
            pass(i);
            ok = go(RPK_DIT, a.R, (i == 0 && k.nsplit > 0) ? out_io : RP_WORK);
        }
        if (ok && k.nsplit == 0)
            ok = sweep1(RPK_DIT, out_io);
        break;
    }
    default: return FFTSET_E_INVALID;
    }
    return ok ? FFTSET_OK : FFTSET_E_DEVICE;
}

}  // namespace fsa

extern "C" unsigned fftset_baseline_chain(const struct fftset_modulation* modulation, unsigned complex_bins,
                                          unsigned* radices, unsigned max)
{
    using namespace fsa;
    if (!modulation || !radices)
        return 0;
    int a, b, c;
    RefStructure rs;
    if (!factor235(complex_bins, &a, &b, &c) || !reference_structure(modulation->kind, complex_bins, &rs))
        return 0;
    unsigned n = 0;
    for (int r : rs.split)
        if (n < max)
            radices[n++] = (unsigned)r;
    if (n < max)
        radices[n++] = 0;
    for (int r : rs.inner)
        if (n < max)
            radices[n++] = (unsigned)r;
    return n;
}
