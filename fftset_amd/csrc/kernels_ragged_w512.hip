// kernels_ragged_w512.hip -- the ragged-batch kernel with 512-lane workgroups.
#include "ragged_impl.h"

namespace fsa {

hipError_t launch_ragged_512(const RaggedArgs& a, int lds_bytes, hipStream_t st)
{
    return launch_w<512>(a, lds_bytes, st);
}

}  // namespace fsa
