// kernels_ragged.hip -- launcher of the ragged-batch kernel (ragged_impl.h);
// each workgroup size is compiled in its own translation unit.
#include "kernels.h"

namespace fsa {

hipError_t launch_ragged_128(const RaggedArgs& a, int lds_bytes, hipStream_t st);
hipError_t launch_ragged_256(const RaggedArgs& a, int lds_bytes, hipStream_t st);
hipError_t launch_ragged_512(const RaggedArgs& a, int lds_bytes, hipStream_t st);
hipError_t launch_ragged_1024(const RaggedArgs& a, int lds_bytes, hipStream_t st);

hipError_t launch_ragged(const RaggedArgs& a, int threads, int lds_bytes, hipStream_t st)
{
    if (a.ntiles <= 0)
        return hipSuccess;
    switch (threads) {
    case 128: return launch_ragged_128(a, lds_bytes, st);
    case 256: return launch_ragged_256(a, lds_bytes, st);
    case 512: return launch_ragged_512(a, lds_bytes, st);
    case 1024: return launch_ragged_1024(a, lds_bytes, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace fsa
