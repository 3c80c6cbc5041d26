// kernels_fsr_cx.hip -- COMPLEX instantiations of the resident 61440-bin
// kernel (fsr_impl.h), and its host entry points.
#include "fsr_impl.h"

namespace fsa {
namespace fsr {

template <>
hipError_t launch_fsr_mod<MOD_CX>(int opk, const FsrArgs& a, hipStream_t st)
{
    const int bot = (a.conj_kern ? 1 : 0) | (!a.kern_p && a.kern_st == 0 ? 2 : 0);
    void* fn = nullptr;
    switch (opk) {
    case FSR_FWD: fn = (void*)k_fsr<MOD_CX, FSR_FWD, 0>; break;
    case FSR_INV: fn = (void*)k_fsr<MOD_CX, FSR_INV, 0>; break;
    case FSR_KERN: fn = (void*)k_fsr<MOD_CX, FSR_KERN, 0>; break;
    case FSR_CONV:
        fn = bot == 0 ? (void*)k_fsr<MOD_CX, FSR_CONV, 0>
           : bot == 1 ? (void*)k_fsr<MOD_CX, FSR_CONV, 1>
           : bot == 2 ? (void*)k_fsr<MOD_CX, FSR_CONV, 2>
                      : (void*)k_fsr<MOD_CX, FSR_CONV, 3>;
        break;
    default: return hipErrorInvalidValue;
    }
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLds);
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    FsrArgs args = a;
    void* kargs[] = {(void*)&args};
    const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)a.ntrans), dim3(kThreads), kargs, kLds, st);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace fsr

bool fsr_matches(const ChainD& colc, const ChainD& rowc)
{
    return colc.n == fsr::kA && colc.np == 2 && colc.p[0].R == 16 && colc.p[1].R == 15 && rowc.n == fsr::kB &&
           rowc.np == 2 && rowc.p[0].R == 16 && rowc.p[1].R == 16;
}

bool fsr_op(int mod, int op) { return !(mod == MOD_FO && (op == OP_FWD || op == OP_INV)); }

hipError_t launch_fsr(int mod, int opk, const FsrArgs& a, hipStream_t st)
{
    if (a.ntrans <= 0)
        return hipSuccess;
    if (a.ntrans > 0x7fffffffLL)
        return hipErrorInvalidValue;
    return mod == MOD_FO ? fsr::launch_fsr_mod<MOD_FO>(opk, a, st) : fsr::launch_fsr_mod<MOD_CX>(opk, a, st);
}

}  // namespace fsa
