// kernels_fsr_fo.hip -- FREQOFFSETREAL instantiations of the resident
// 61440-bin kernel (fsr_impl.h): get_kernel and the convolutions.
#include "fsr_impl.h"

namespace fsa {
namespace fsr {

template <>
hipError_t launch_fsr_mod<MOD_FO>(int opk, const FsrArgs& a, hipStream_t st)
{
    const int bot = (a.conj_kern ? 1 : 0) | (!a.kern_p && a.kern_st == 0 ? 2 : 0);
    void* fn = nullptr;
    if (opk == FSR_KERN)
        fn = (void*)k_fsr<MOD_FO, FSR_KERN, 0>;
    else if (opk == FSR_CONV)
        fn = bot == 0 ? (void*)k_fsr<MOD_FO, FSR_CONV, 0>
           : bot == 1 ? (void*)k_fsr<MOD_FO, FSR_CONV, 1>
           : bot == 2 ? (void*)k_fsr<MOD_FO, FSR_CONV, 2>
                      : (void*)k_fsr<MOD_FO, FSR_CONV, 3>;
    else
        return hipErrorInvalidValue;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLds);
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    FsrArgs args = a;
    void* kargs[] = {(void*)&args};
    const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)a.ntrans), dim3(kThreads), kargs, kLds, st);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace fsr
}  // namespace fsa
