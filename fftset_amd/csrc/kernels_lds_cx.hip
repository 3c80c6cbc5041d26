// kernels_lds_cx.hip -- LDS-resident transforms of the COMPLEX modulation.
#include "lds_impl.h"

namespace fsa {

hipError_t launch_lds_cx(int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (sh.E == 16)
        return launch_op<MOD_CX, 16>(op, a, sh, st);
    if (sh.E == 32)
        return launch_op<MOD_CX, 32>(op, a, sh, st);
    return hipErrorInvalidValue;
}

}  // namespace fsa
