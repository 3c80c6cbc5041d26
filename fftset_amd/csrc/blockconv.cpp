// blockconv.cpp -- block (overlap-save) FIR convolution of long real signals.
//
// fftset_recommend_conv_length (reference fftset.h:144-151, src/fftset.c:131-167)
// sizes the FREQOFFSETREAL transform for a kernel of k taps and a block
// size; this is the driver that uses it: every channel of a long signal is
// cut into blocks of b = 2L - k + 1 samples, and every (channel, block)
// frame is one convolution in a single launch (kernels_block.hip).
#include <hip/hip_runtime.h>

#include <new>

#include "context.h"
#include "planner.h"

struct fftset_blockconv {
    fsa::Ctx* ctx = nullptr;
    const fftset_fft* plan = nullptr;
    float* kern = nullptr;  // kernel spectrum of taps / L (2L floats)
    unsigned k = 0;         // taps
    long long b = 0;        // new samples per block
};

using namespace fsa;

extern "C" {

struct fftset_blockconv* fftset_blockconv_create(struct fftset* fc, const float* taps, unsigned kernel_length,
                                                 unsigned max_block_size)
{
    if (!fc || !fc->impl || !taps || kernel_length == 0) {
        set_error(FFTSET_E_INVALID);
        return nullptr;
    }
    Ctx* ctx = (Ctx*)fc->impl;
    const unsigned L = fftset_recommend_conv_length(kernel_length, max_block_size);
    const fftset_fft* p = fftset_create_fft(fc, FFTSET_MODULATION_FREQ_OFFSET_REAL, L);
    if (!p)
        return nullptr;  // error already set
    if (p->engine != ENG_LDS || p->shape.E != 16 || 2LL * L < (long long)kernel_length) {
        // the fused gather / scatter runs on the LDS-resident engine only
        set_error(FFTSET_E_UNSUPPORTED);
        return nullptr;
    }
    int rc = use_device(ctx);
    hipStream_t st = thread_stream(ctx->device);
    if (rc != FFTSET_OK || !st) {
        set_error(FFTSET_E_DEVICE);
        return nullptr;
    }
    fftset_blockconv* bc = new (std::nothrow) fftset_blockconv;
    if (!bc) {
        set_error(FFTSET_E_NOMEM);
        return nullptr;
    }
    bc->ctx = ctx;
    bc->plan = p;
    bc->k = kernel_length;
    const long long N = 2LL * L;
    bc->b = N - kernel_length + 1;
    float* staging = nullptr;
    if (hipMalloc((void**)&bc->kern, N * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&staging, N * sizeof(float)) != hipSuccess) {
        if (bc->kern)
            (void)hipFree(bc->kern);
        delete bc;
        set_error(FFTSET_E_NOMEM);
        return nullptr;
    }
    // kernel spectrum of the taps scaled by 1/L, so a frame's convolution
    // (L times the negacyclic one) is the plain convolution
    rc = launch_taps(staging, taps, (int)kernel_length, (int)N, 1.0f / (float)L, st) == hipSuccess
             ? FFTSET_OK
             : FFTSET_E_DEVICE;
    if (rc == FFTSET_OK)
        rc = execute(p, FFTSET_OP_GET_KERNEL, 1, bc->kern, N, staging, N, nullptr, 0, nullptr, 0, st);
    if (rc == FFTSET_OK && hipStreamSynchronize(st) != hipSuccess)
        rc = FFTSET_E_DEVICE;
    (void)hipFree(staging);
    if (rc != FFTSET_OK) {
        (void)hipFree(bc->kern);
        delete bc;
        set_error(rc);
        return nullptr;
    }
    set_error(FFTSET_OK);
    return bc;
}

int fftset_blockconv_run(struct fftset_blockconv* bc, float* out, size_t out_stride, const float* in,
                         size_t in_stride, size_t nsamples, size_t channels, unsigned flags, void* stream)
{
    if (!bc || !out || !in || (flags & ~FFTSET_BLOCKCONV_HISTORY) ||
        (channels > 1 && (out_stride < nsamples || in_stride < nsamples))) {
        set_error(FFTSET_E_INVALID);
        return FFTSET_E_INVALID;
    }
    if (nsamples == 0 || channels == 0) {
        set_error(FFTSET_OK);
        return FFTSET_OK;
    }
    int rc = use_device(bc->ctx);
    if (rc != FFTSET_OK) {
        set_error(rc);
        return rc;
    }
    const fftset_fft* p = bc->plan;
    BlockArgs a{};
    a.c = p->chain;
    a.tmod = p->tmod;
    a.tfo = p->tfo;
    a.L = (int)p->lfft;
    a.n = (long long)nsamples;
    a.b = bc->b;
    a.nblocks = (a.n + a.b - 1) / a.b;
    a.ntrans = (long long)channels * a.nblocks;
    a.in = in;
    a.out = out;
    a.kern = bc->kern;
    a.in_st = (long long)in_stride;
    a.out_st = (long long)out_stride;
    a.skip = (int)bc->k - 1;
    a.lo = (flags & FFTSET_BLOCKCONV_HISTORY) ? -(long long)a.skip : 0;
    ShapeD sh = p->shape;
    if (a.ntrans < sh.G && !choose_shape(p->chain, sh.E, a.ntrans, &sh)) {
        set_error(FFTSET_E_UNSUPPORTED);
        return FFTSET_E_UNSUPPORTED;
    }
    a.G = sh.G;
    a.T = sh.T;
    rc = launch_block(a, sh, (hipStream_t)stream) == hipSuccess ? FFTSET_OK : FFTSET_E_DEVICE;
    set_error(rc);
    return rc;
}

unsigned fftset_blockconv_block_size(const struct fftset_blockconv* bc) { return bc ? (unsigned)bc->b : 0; }

unsigned fftset_blockconv_complex_bins(const struct fftset_blockconv* bc) { return bc ? bc->plan->lfft : 0; }

void fftset_blockconv_destroy(struct fftset_blockconv* bc)
{
    if (!bc)
        return;
    (void)use_device(bc->ctx);
    (void)hipFree(bc->kern);  // waits for work using it
    delete bc;
}

}  // extern "C"
