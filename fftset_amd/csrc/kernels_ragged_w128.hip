// kernels_ragged_w128.hip -- the ragged-batch kernel with 128-lane workgroups.
#include "ragged_impl.h"

namespace fsa {

hipError_t launch_ragged_128(const RaggedArgs& a, int lds_bytes, hipStream_t st)
{
    return launch_w<128>(a, lds_bytes, st);
}

}  // namespace fsa
