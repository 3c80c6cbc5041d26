// select.cpp -- process-wide kernel selection and launch counters.
//
// Chains with a compile-time specialisation (kernels_fixed.hip, the Fix2
// four-step chains, the block-convolution frame kernel) run their
// specialised kernel unless FFTSET_SELECT_GENERIC is set; FFTSET_SELECT_
// PERSISTENT switches the LDS and four-step launches to resident grids with
// register prefetch; FFTSET_SELECT_FOUR_STEP_FUSED runs 61440-bin transforms
// with one workgroup per transform and a scratch slot, FFTSET_SELECT_FOUR_
// STEP_SPLIT with column and row launches, instead of the CU-resident kernel.
// The environment (FFTSET_AMD_GENERIC, FFTSET_AMD_PERSIST) gives the initial
// value, read once when the library loads; launches read an atomic, never
// the environment.
#include <atomic>
#include <cstdlib>

#include "fftset.h"
#include "kernels.h"

namespace fsa {

static unsigned initial_select()
{
    unsigned f = 0;
    const char* g = getenv("FFTSET_AMD_GENERIC");
    if (g && atoi(g) > 0)
        f |= FFTSET_SELECT_GENERIC;
    const char* p = getenv("FFTSET_AMD_PERSIST");
    if (p && atoi(p) > 0)
        f |= FFTSET_SELECT_PERSISTENT;
    return f;
}

static std::atomic<unsigned> g_select{initial_select()};
static std::atomic<unsigned long long> g_specialised{0}, g_generic{0};

bool persistent_mode() { return (g_select.load(std::memory_order_relaxed) & FFTSET_SELECT_PERSISTENT) != 0; }

bool generic_forced() { return (g_select.load(std::memory_order_relaxed) & FFTSET_SELECT_GENERIC) != 0; }

bool four_step_fused()
{
    return (g_select.load(std::memory_order_relaxed) & FFTSET_SELECT_FOUR_STEP_FUSED) != 0;
}

bool four_step_split()
{
    return (g_select.load(std::memory_order_relaxed) & FFTSET_SELECT_FOUR_STEP_SPLIT) != 0;
}

void count_launch(bool specialised)
{
    (specialised ? g_specialised : g_generic).fetch_add(1, std::memory_order_relaxed);
}

}  // namespace fsa

extern "C" {

unsigned fftset_kernel_select(unsigned flags)
{
    return fsa::g_select.exchange(flags &
                                  (FFTSET_SELECT_GENERIC | FFTSET_SELECT_PERSISTENT | FFTSET_SELECT_FOUR_STEP_FUSED |
                                   FFTSET_SELECT_FOUR_STEP_SPLIT));
}

void fftset_kernel_counts(unsigned long long* specialised, unsigned long long* generic)
{
    const unsigned long long s = fsa::g_specialised.exchange(0), g = fsa::g_generic.exchange(0);
    if (specialised)
        *specialised = s;
    if (generic)
        *generic = g;
}

}  // extern "C"
