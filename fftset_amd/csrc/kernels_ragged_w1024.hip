// kernels_ragged_w1024.hip -- the ragged-batch kernel with 1024-lane workgroups.
#include "ragged_impl.h"

namespace fsa {

hipError_t launch_ragged_1024(const RaggedArgs& a, int lds_bytes, hipStream_t st)
{
    return launch_w<1024>(a, lds_bytes, st);
}

}  // namespace fsa
