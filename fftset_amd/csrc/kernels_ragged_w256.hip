// kernels_ragged_w256.hip -- the ragged-batch kernel with 256-lane workgroups.
#include "ragged_impl.h"

namespace fsa {

hipError_t launch_ragged_256(const RaggedArgs& a, int lds_bytes, hipStream_t st)
{
    return launch_w<256>(a, lds_bytes, st);
}

}  // namespace fsa
