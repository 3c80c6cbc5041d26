// kernels_rcol_fo.hip -- FREQOFFSETREAL instantiations of the
// column-register four-step kernel (rcol_impl.h).
#include "rcol_impl.h"

namespace fsa {
namespace rcol {

hipError_t launch_fo(int A, int opk, const RcolArgs& a, hipStream_t st)
{
    return A == 60 ? launch_mod<60, MOD_FO>(opk, a, st)
         : A == 18 ? launch_mod<18, MOD_FO>(opk, a, st)
         : A == 16 ? launch_mod<16, MOD_FO>(opk, a, st)
         : A == 15 ? launch_mod<15, MOD_FO>(opk, a, st)
         : A == 8  ? launch_mod<8, MOD_FO>(opk, a, st)
         : A == 3  ? launch_mod<3, MOD_FO>(opk, a, st)
                   : hipErrorInvalidValue;
}

}  // namespace rcol
}  // namespace fsa
