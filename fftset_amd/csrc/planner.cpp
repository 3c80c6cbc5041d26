// planner.cpp -- radix chains and launch shapes (host side).
//
// The reference searches a min-cost chain over CPU vector passes
// (src/fftset_vec.c:1004-1062). On gfx950 the cost that matters is the number
// of LDS round trips and the register footprint of a butterfly, so the
// default chain takes the largest register radices first; the autotuner may
// replace it per size.
#include "planner.h"

#include <algorithm>

namespace fsa {

bool factor235(unsigned n, int* a, int* b, int* c)
{
    if (n < 2)
        return false;
    int x = 0, y = 0, z = 0;
    while (n % 2 == 0) { n /= 2; ++x; }
    while (n % 3 == 0) { n /= 3; ++y; }
    while (n % 5 == 0) { n /= 5; ++z; }
    if (n != 1 || z > 1)
        return false;
    *a = x;
    *b = y;
    *c = z;
    return true;
}

std::vector<int> default_chain(unsigned n)
{
    int a, b, c;
    std::vector<int> r;
    if (!factor235(n, &a, &b, &c))
        return r;
    if (c) {
        if (b) { r.push_back(15); --b; }
        else if (a) { r.push_back(10); --a; }
        else r.push_back(5);
    }
    while (b >= 2) { r.push_back(9); b -= 2; }
    if (b == 1) {
        if (a >= 2) { r.push_back(12); a -= 2; }
        else if (a == 1) { r.push_back(6); --a; }
        else r.push_back(3);
    }
    int sixteens = a / 4, rest = a % 4;
    if (rest == 1 && sixteens > 0) {
        // 16*2 -> 8*4: avoids a radix-2 pass
        --sixteens;
        r.push_back(8);
        r.push_back(4);
        rest = 0;
    }
    for (int i = 0; i < sixteens; ++i) r.push_back(16);
    if (rest == 3) r.push_back(8);
    else if (rest == 2) r.push_back(4);
    else if (rest == 1) r.push_back(2);
    std::sort(r.begin(), r.end(), [](int x, int y) { return x > y; });
    return r;
}

static void chain_multisets(unsigned n, int start, int max_passes, std::vector<int>& cur,
                            std::vector<std::vector<int>>& out)
{
    static const int kRadices[] = {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};
    if (n == 1) {
        if (!cur.empty())
            out.push_back(cur);
        return;
    }
    if ((int)cur.size() >= max_passes)
        return;
    for (int i = start; i < (int)(sizeof(kRadices) / sizeof(kRadices[0])); ++i) {
        const int r = kRadices[i];
        if (n % r)
            continue;
        // at most one radix-2 pass, one radix-5 family pass (5 | n once)
        if (r == 2 && std::count(cur.begin(), cur.end(), 2))
            continue;
        cur.push_back(r);
        chain_multisets(n / r, i, max_passes, cur, out);
        cur.pop_back();
    }
}

std::vector<std::vector<int>> candidate_chains(unsigned n, int max_passes, int limit)
{
    std::vector<std::vector<int>> sets, out;
    std::vector<int> cur;
    int a, b, c;
    if (!factor235(n, &a, &b, &c))
        return out;
    out.push_back(default_chain(n));
    chain_multisets(n, 0, max_passes, cur, sets);
    // fewest passes first, then largest leading radix
    std::stable_sort(sets.begin(), sets.end(), [](const std::vector<int>& x, const std::vector<int>& y) {
        return x.size() < y.size();
    });
    for (auto& m : sets) {
        std::vector<int> desc = m, asc = m;
        std::sort(desc.begin(), desc.end(), [](int x, int y) { return x > y; });
        std::sort(asc.begin(), asc.end());
        for (auto* v : {&desc, &asc}) {
            if ((int)out.size() >= limit)
                return out;
            if (std::find(out.begin(), out.end(), *v) == out.end())
                out.push_back(*v);
        }
    }
    return out;
}

bool build_chain(const std::vector<int>& radices, unsigned n, ChainD* out)
{
    if (radices.empty() || (int)radices.size() > kMaxPasses)
        return false;
    unsigned len = n, s = 1;
    out->n = (int)n;
    out->np = (int)radices.size();
    for (size_t i = 0; i < radices.size(); ++i) {
        const unsigned R = (unsigned)radices[i];
        if (len % R)
            return false;
        PassD& p = out->p[i];
        p.R = (int)R;
        p.m = (int)(len / R);
        p.s = (int)s;
        p.pad_ = 0;
        p.tw = nullptr;
        s *= R;
        len /= R;
    }
    return len == 1;
}

bool fits_lds(unsigned n)
{
    return n >= 2 && (long long)(n + (n >> 4) + 1) * 8 <= kMaxLdsBytes && n <= 16384;
}

bool choose_shape(const ChainD& c, int E, long long batch_hint, ShapeD* out)
{
    int T = 1;
    for (int i = 0; i < c.np; ++i) {
        const int R = c.p[i].R;
        const int nb = std::max(1, E / R);
        const int nbf = c.n / R;
        T = std::max(T, (nbf + nb - 1) / nb);
    }
    const int max_threads = E <= 16 ? 1024 : 512;
    if (T > max_threads)
        return false;
    const long long lds_one = (long long)(c.n + (c.n >> 4) + 1) * 8;
    if (lds_one > kMaxLdsBytes)
        return false;
    // Aim for ~256 lanes per workgroup and <= 64 KiB of LDS so several
    // workgroups share a CU.
    int G = std::max(1, 256 / T);
    while (G > 1 && (long long)G * lds_one > 64 * 1024) --G;
    if (batch_hint > 0 && G > batch_hint) G = (int)batch_hint;
    int threads = G * T;
    threads = ((threads + 63) / 64) * 64;
    if (threads > max_threads)
        return false;
    out->E = E;
    out->T = T;
    out->G = G;
    out->threads = threads;
    out->lds_bytes = (int)(G * lds_one);
    return true;
}

}  // namespace fsa
