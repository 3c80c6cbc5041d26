// kernels_rcol_cx.hip -- COMPLEX instantiations of the column-register
// four-step kernel (rcol_impl.h), and its host entry points.
#include "rcol_impl.h"

namespace fsa {
namespace rcol {
// FREQOFFSETREAL instantiations: kernels_rcol_fo.hip
hipError_t launch_fo(int A, int opk, const RcolArgs& a, hipStream_t st);
}  // namespace rcol

bool rcol_length(unsigned A) { return A == 60 || A == 18 || A == 16 || A == 15 || A == 8 || A == 3; }

bool rcol_matches(const ChainD& colc, const ChainD& rowc)
{
    return rcol_length((unsigned)colc.n) && rowc.n == rcol::kB && rowc.np == 2 && rowc.p[0].R == 16 && rowc.p[1].R == 16;
}

hipError_t launch_rcol(int A, int mod, int opk, const RcolArgs& a, hipStream_t st)
{
    if (a.ntrans <= 0)
        return hipSuccess;
    if (a.ntrans > 0x7fffffffLL || !rcol_length((unsigned)A))
        return hipErrorInvalidValue;
    if (mod == MOD_FO)
        return rcol::launch_fo(A, opk, a, st);
    return A == 60   ? rcol::launch_mod<60, MOD_CX>(opk, a, st)
         : A == 18 ? rcol::launch_mod<18, MOD_CX>(opk, a, st)
         : A == 16 ? rcol::launch_mod<16, MOD_CX>(opk, a, st)
         : A == 15 ? rcol::launch_mod<15, MOD_CX>(opk, a, st)
         : A == 8  ? rcol::launch_mod<8, MOD_CX>(opk, a, st)
                   : rcol::launch_mod<3, MOD_CX>(opk, a, st);
}

}  // namespace fsa
