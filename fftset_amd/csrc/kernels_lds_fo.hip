// kernels_lds_fo.hip -- LDS-resident transforms of the FREQOFFSETREAL modulation.
#include "lds_impl.h"

namespace fsa {

hipError_t launch_lds_fo(int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E == 16)
        return launch_op<MOD_FO, 16>(op, a, sh, st);
    if (sh.E == 32)
        return launch_op<MOD_FO, 32>(op, a, sh, st);
    return hipErrorInvalidValue;
}

}  // namespace fsa
