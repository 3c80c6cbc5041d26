// kernels_lds.hip -- the FP64 twiddle generator and the launcher of the
// LDS-resident transforms (lds_impl.h, one translation unit per modulation).
#include <cstdlib>

#include "kernels.h"

namespace fsa {

hipError_t launch_lds_fo(int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st);
hipError_t launch_lds_cx(int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st);

// ---------------------------------------------------------------- twiddles
// out[r*cols + c] = W_N^{x}, x = ((a0 + rs*r) * c) mod N, W_N = exp(-2 pi i / N):
// evaluated in FP64 (exact integer reduction of the angle) and rounded once
// to FP32. This is the only FP64 arithmetic in the library.
__global__ void k_twiddle_fp64(float2* __restrict__ out, long long N, int a0, int rs, int rows, int cols)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)rows * cols)
        return;
    const long long r = idx / cols, c = idx - r * cols;
    long long x = ((a0 + (long long)rs * r) % N) * (c % N) % N;
    if ((4 * x) % N == 0) {
        const int q = (int)((4 * x) / N);  // quarter turns are exact
        const float cr[4] = {1.f, 0.f, -1.f, 0.f};
        const float ci[4] = {0.f, -1.f, 0.f, 1.f};
        out[idx] = make_float2(cr[q], ci[q]);
        return;
    }
    if (2 * x > N)
        x -= N;  // angle in (-pi, pi]
    double sn, cs;
    sincospi(-2.0 * (double)x / (double)N, &sn, &cs);
    out[idx] = make_float2((float)cs, (float)sn);
}

hipError_t launch_twiddles(float2* out, long long N, int a0, int rs, int rows, int cols, hipStream_t st)
{
    const long long n = (long long)rows * cols;
    const int threads = 256;
    (void)hipGetLastError();  // report this launch, not an earlier failed call
    hipLaunchKernelGGL(k_twiddle_fp64, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0, st, out, N,
                       a0, rs, rows, cols);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launcher
// Workgroups resident at once (occupancy x CUs), capped by the work.
long long persistent_grid(const void* fn, int threads, int lds_bytes, long long tiles)
{
    static thread_local int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds_bytes) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const long long resident = (long long)per_cu * cus * persistent_waves();
    return tiles < resident ? tiles : resident;
}

// Resident-grid multiplier (waves of workgroups per launch); 1 = exactly the
// resident set. Read once from FFTSET_AMD_GRID_WAVES.
int persistent_waves()
{
    static int w = 0;
    if (!w) {
        const char* e = getenv("FFTSET_AMD_GRID_WAVES");
        w = e ? atoi(e) : 1;
        if (w < 1)
            w = 1;
    }
    return w;
}
hipError_t launch_lds_fixed(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st);

// Chains with a compile-time specialisation (kernels_fixed.hip) use it,
// unless the kernel selection asks for the generic kernel.
hipError_t launch_lds(int mod, int op, const LdsArgs& a, const ShapeD& sh, hipStream_t st)
{
    if (!persistent_mode() && !generic_forced()) {
        const hipError_t e = launch_lds_fixed(mod, op, a, sh, st);
        if (e != hipErrorNotSupported) {
            count_launch(true);
            return e;
        }
    }
    count_launch(false);
    return mod == MOD_FO ? launch_lds_fo(op, a, sh, st) : launch_lds_cx(op, a, sh, st);
}

}  // namespace fsa
