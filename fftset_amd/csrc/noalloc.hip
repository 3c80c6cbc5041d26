// noalloc.hip -- fftset_fft_execute_noalloc: the stream-ordered batch call
// for callers that cannot allocate device memory. Its work space is a static
// device array instead of library-allocated (hipMalloc) scratch; the
// kernels and their results are those of fftset_fft_execute.
#include <hip/hip_runtime.h>

#include <mutex>

#include "context.h"
#include "fsr.h"

namespace fsa {

// FFTSET_NOALLOC_WORK_FLOATS floats (64 MiB) of work space on every GPU the
// library runs on, held until the process ends. In host code this name is
// a host-side stand-in, not a device address: host code uses only the
// address noalloc_work() looks up on each device.
static __device__ __align__(256) float scratch[FFTSET_NOALLOC_WORK_FLOATS];

constexpr int kNoallocDevices = 64;
static std::mutex g_mu;
static float* g_addr[kNoallocDevices];

// Looked up once per device, with that device current (fftset_init's).
int noalloc_work(int device, float** out)
{
    *out = nullptr;
    if (device < 0 || device >= kNoallocDevices)
        return FFTSET_E_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return FFTSET_E_DEVICE;
        }
        g_addr[device] = (float*)p;
    }
    *out = g_addr[device];
    return FFTSET_OK;
}

// Floats of work space per transform that fftset_fft_execute would
// allocate for this operation.
static size_t work_per_transform(const fftset_fft* p, int op)
{
    const size_t len2 = 2ull * p->lfft;
    const bool four = four_step_work(p, op == FFTSET_OP_GET_KERNEL ? OP_KERN : OP_CONV);
    switch (op) {
    case FFTSET_OP_GET_KERNEL: return four && p->fs.fsr ? 2ull * kFsrSpill : 0;  // the resident kernel's spill
    case FFTSET_OP_CONV_GRAD_KERNEL: return len2 * (four ? 2 : 1);
    default: return four ? len2 : 0;
    }
}

}  // namespace fsa

using namespace fsa;

extern "C" int fftset_fft_execute_noalloc(const struct fftset_fft* fft, int op, size_t count, float* out,
                                          size_t out_stride, const float* in, size_t in_stride, const float* kern,
                                          size_t kern_stride, void* stream)
{
    // FFTSET_OP_CONV_SAVE_SPECTRUM takes its spectrum output in the work
    // argument, which this call does not have
    if (!fft || op < FFTSET_OP_FORWARD || op > FFTSET_OP_CONV_GRAD_KERNEL) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    int rc = use_device(fft->ctx);
    if (rc != FFTSET_OK) {
        set_error(rc);
        return rc;
    }
    float* w = fft->ctx->noalloc;
    if (!w) {
        set_error(FFTSET_E_DEVICE);
        return FFTSET_E_DEVICE;
    }
    const size_t per = work_per_transform(fft, op);
    if (per && count > FFTSET_NOALLOC_WORK_FLOATS / per) {
        set_error(FFTSET_E_TOO_LARGE);
        return FFTSET_E_TOO_LARGE;
    }
    // the regions fftset_fft_execute would take from its scratch: work for
    // every operation that has one, a second region after it for a four-step
    // CONV_GRAD_KERNEL
    const long long len2 = 2LL * fft->lfft;
    const long long wst = op == FFTSET_OP_GET_KERNEL ? (long long)per : len2;
    float* work = per ? w : nullptr;
    float* work2 = per ? w + (long long)count * len2 : nullptr;
    rc = execute(fft, op, (long long)count, out, (long long)out_stride, in, (long long)in_stride, kern,
                 (long long)kern_stride, work, per ? wst : 0, (hipStream_t)stream, work2);
    set_error(rc);
    return rc;
}
