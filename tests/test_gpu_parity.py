"""GPU parity: the HIP path through the C ABI against the oracle (the FP32
restatement of the reference, <= 2e-6 relative RMS per transform), against
FP64 closed forms (<= 1e-6 * log2(L) relative RMS), and against the
reference's own known-answer tests with its thresholds."""
import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from oracle import fp64
from oracle.oracle import MOD_COMPLEX as OCX, MOD_FREQ_OFFSET_REAL as OFO
from tests import kat
from tests.gpu_util import (gpu_batch, gpu_conv, gpu_forward, gpu_get_kernel, gpu_inverse, torch_cuda)

pytestmark = pytest.mark.gpu

REF_TOL = 2e-6      # relative RMS against the reference restatement


def fp64_tol(L):
    return 1e-6 * max(1.0, np.log2(L))


@pytest.fixture(scope="module")
def fs():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


@pytest.fixture(scope="module")
def tuned_fs():
    import os
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet(tune_file=os.path.join(os.path.dirname(fftset_amd.__file__), "tune_mi355x.txt"))
    yield s
    s.destroy()


@pytest.fixture(scope="module")
def lds_tuned_fs(tmp_path_factory):
    """A set planning 3840, 4096 and 2048 on the LDS chains the compile-time
    three-pass kernels are specialised for (the MI355X table plans these
    sizes on the column-register kernel instead)."""
    if torch_cuda() is None:
        pytest.skip("no GPU")
    path = tmp_path_factory.mktemp("lds_tune") / "tune.txt"
    path.write_text("1 3840 1 0 0 3 15 16 16\n0 4096 1 0 0 3 16 16 16\n0 2048 2 0 0 3 8 16 16\n")
    s = fftset_amd.FFTSet(tune_file=str(path))
    yield s
    s.destroy()


FO, CX = fftset_amd.FREQ_OFFSET_REAL, fftset_amd.COMPLEX


# ---------------------------------------------------------------- reference KATs
@pytest.mark.parametrize("L", kat.TEST_LENGTHS)
def test_kat_fo_prime_impulse(fs, L):
    f = fs.create_fft(FO, L)
    x, X = kat.fo_impulse_case(L)
    Y = gpu_forward(f, x)
    assert kat.abs_rms_bins(Y, X, L) <= kat.FWD_RMS
    assert kat.abs_rms_bins(gpu_inverse(f, Y) / L, x, L) <= kat.INV_RMS


@pytest.mark.parametrize("L", kat.TEST_LENGTHS)
def test_kat_cx_prime_impulse(fs, L):
    f = fs.create_fft(CX, L)
    x, X = kat.cx_impulse_case(L)
    Y = gpu_forward(f, x)
    assert kat.abs_rms_bins(Y, X, L) <= kat.FWD_RMS
    assert kat.abs_rms_bins(gpu_inverse(f, Y) / L, x, L) <= kat.INV_RMS


@pytest.mark.parametrize("L", kat.TEST_LENGTHS)
def test_kat_convolution(fs, L):
    f = fs.create_fft(FO, L)
    x, h, expect = kat.conv_case(L)
    assert kat.abs_rms_bins(gpu_conv(f, x, h), expect, L) <= kat.CONV_RMS


# ---------------------------------------------------------------- random parity
SIZES = [2, 3, 5, 6, 9, 10, 12, 15, 16, 18, 27, 45, 48, 64, 96, 120, 128, 240, 360, 512, 576, 768,
         960, 1440, 2048, 3840, 4096, 4608, 6144, 9216, 15360, 16384,
         # beyond one LDS-resident kernel: four-step through HBM
         12960, 18432, 20480, 61440, 65536]


@pytest.mark.parametrize("L", SIZES)
def test_complex_parity(fs, oracle, L):
    rng = np.random.default_rng(1000 + L)
    f = fs.create_fft(CX, L)
    o = oracle.plan(OCX, L)
    x = rng.standard_normal(2 * L).astype(np.float32)
    y = gpu_forward(f, x)
    assert fp64.rel_rms(y, o.forward(x)) <= REF_TOL
    assert fp64.rel_rms(y, fp64.cx_forward64(x)) <= fp64_tol(L)
    z = gpu_inverse(f, x)
    assert fp64.rel_rms(z, o.inverse(x)) <= REF_TOL
    assert fp64.rel_rms(z, fp64.cx_inverse64(x)) <= fp64_tol(L)
    h = rng.standard_normal(2 * L).astype(np.float32)
    c = gpu_conv(f, x, h)
    assert fp64.rel_rms(c, o.conv(x, o.get_kernel(h))) <= REF_TOL
    assert fp64.rel_rms(c, fp64.cx_conv64(x, h)) <= fp64_tol(L)


@pytest.mark.parametrize("L", SIZES)
def test_freq_offset_real_parity(fs, oracle, L):
    rng = np.random.default_rng(2000 + L)
    f = fs.create_fft(FO, L)
    o = oracle.plan(OFO, L)
    x = rng.uniform(-1, 1, 2 * L).astype(np.float32)
    X = gpu_forward(f, x)
    assert fp64.rel_rms(X, o.forward(x)) <= REF_TOL
    assert fp64.rel_rms(X, fp64.fo_forward64(x)) <= fp64_tol(L)
    Xr = rng.standard_normal(2 * L).astype(np.float32)
    y = gpu_inverse(f, Xr)
    assert fp64.rel_rms(y, o.inverse(Xr)) <= REF_TOL
    assert fp64.rel_rms(y, fp64.fo_inverse64(Xr)) <= fp64_tol(L)
    # full-block kernel: the convolution wraps, so it must be negacyclic
    h = (rng.standard_normal(2 * L) * np.exp(-np.arange(2 * L) / max(1.0, L / 4))).astype(np.float32)
    c = gpu_conv(f, x, h)
    assert fp64.rel_rms(c, o.conv(x, o.get_kernel(h))) <= REF_TOL
    assert fp64.rel_rms(c, fp64.fo_conv64(x, h)) <= fp64_tol(L)


@pytest.mark.parametrize("mod", [FO, CX])
@pytest.mark.parametrize("L", [4, 45, 512, 3840, 61440])
def test_in_place_aliasing(fs, mod, L):
    # input and output may alias for forward, inverse and conv
    from tests.gpu_util import dev, empty, host
    rng = np.random.default_rng(L)
    f = fs.create_fft(mod, L)
    x = rng.standard_normal(2 * L).astype(np.float32)
    h = rng.standard_normal(2 * L).astype(np.float32)
    ref_f, ref_i, ref_c = gpu_forward(f, x), gpu_inverse(f, x), gpu_conv(f, x, h)
    work, k = empty(2 * L), empty(2 * L)
    buf = dev(x)
    f.forward(buf, buf, work)
    assert np.array_equal(host(buf), ref_f)
    buf = dev(x)
    f.inverse(buf, buf, work)
    assert np.array_equal(host(buf), ref_i)
    f.conv_get_kernel(k, dev(h))
    buf = dev(x)
    f.conv(buf, buf, k, work)
    assert np.array_equal(host(buf), ref_c)


@pytest.mark.parametrize("mod,L", [(CX, 48), (FO, 120), (CX, 3840), (FO, 4096), (FO, 2048), (CX, 61440),
                                   (FO, 61440)])
def test_uniform_batch_matches_single_calls(fs, oracle, mod, L):
    rng = np.random.default_rng(7 * L)
    f = fs.create_fft(mod, L)
    o = oracle.plan(OFO if mod == FO else OCX, L)
    B = 37 if L < 20000 else 5
    x = rng.standard_normal((B, 2 * L)).astype(np.float32)
    h = rng.standard_normal((B, 2 * L)).astype(np.float32)
    y = gpu_batch(f, fftset_amd.OP_FORWARD, x)
    assert np.max(fp64.rel_rms(y, o.forward(x))) <= REF_TOL
    z = gpu_batch(f, fftset_amd.OP_INVERSE, x)
    assert np.max(fp64.rel_rms(z, o.inverse(x))) <= REF_TOL
    K = gpu_batch(f, fftset_amd.OP_GET_KERNEL, h)
    c = gpu_batch(f, fftset_amd.OP_CONV, x, K)
    assert np.max(fp64.rel_rms(c, o.conv(x, o.get_kernel(h)))) <= REF_TOL
    # bit-identical to the drop-in call for the same transform
    assert np.array_equal(y[B - 1], gpu_forward(f, x[B - 1]))


@pytest.mark.parametrize("L", [2, 5, 16, 120, 2048, 4096, 61440])
def test_conv_adjoints_fp64(fs, L):
    rng = np.random.default_rng(3 * L)
    f = fs.create_fft(FO, L)
    N = 2 * L
    x = rng.standard_normal(N).astype(np.float32)
    h = (rng.standard_normal(N) * np.exp(-np.arange(N) / max(1.0, L / 8))).astype(np.float32)
    g = rng.standard_normal(N).astype(np.float32)
    gx_ref, gh_ref = fp64.fo_conv_grads64(g, x, h)
    K = gpu_get_kernel(f, h)
    gx = gpu_batch(f, fftset_amd.OP_CONV_GRAD_INPUT, g[None], K[None])[0]
    gh = gpu_batch(f, fftset_amd.OP_CONV_GRAD_KERNEL, x[None], g[None])[0]
    assert fp64.rel_rms(gx, gx_ref) <= fp64_tol(L)
    assert fp64.rel_rms(gh, gh_ref) <= fp64_tol(L)


def test_unsupported_sizes_return_null(fs):
    for n in (0, 1, 7, 25, 50, 11 * 16):
        with pytest.raises(fftset_amd.FFTSetError) as e:
            fs.create_fft(FO, n)
        assert e.value.code == fftset_amd.E_UNSUPPORTED


def test_plan_cache_returns_same_plan(fs):
    lib = fftset_amd.load_library()
    a = lib.fftset_create_fft(fs.address, fftset_amd._modulation(FO), 96)
    b = lib.fftset_create_fft(fs.address, fftset_amd._modulation(FO), 96)
    c = lib.fftset_create_fft(fs.address, fftset_amd._modulation(CX), 96)
    assert a and a == b and c != a


# ------------------------------------------------ sizes beyond 65536 bins
LARGE = [131072, 245760, 1048576, 4194304]


@pytest.mark.parametrize("mod", [FO, CX])
@pytest.mark.parametrize("L", LARGE)
def test_large_sizes_fp64(fs, oracle, mod, L):
    # four-step with both factors LDS-resident (up to 16384 x 16384 bins)
    rng = np.random.default_rng(3000 + L)
    f = fs.create_fft(mod, L)
    radices = f.radices()
    assert 0 in radices  # column chain, 0, row chain
    x = rng.standard_normal(2 * L).astype(np.float32)
    h = (rng.standard_normal(2 * L) * np.exp(-np.arange(2 * L) / (L / 4))).astype(np.float32)
    y = gpu_forward(f, x)
    c = gpu_conv(f, x, h)
    if mod == FO:
        assert fp64.rel_rms(y, fp64.fo_forward64(x)) <= fp64_tol(L)
        assert fp64.rel_rms(c, fp64.fo_conv64(x, h)) <= fp64_tol(L)
    else:
        assert fp64.rel_rms(y, fp64.cx_forward64(x)) <= fp64_tol(L)
        assert fp64.rel_rms(c, fp64.cx_conv64(x, h)) <= fp64_tol(L)
    if L <= 1048576:
        o = oracle.plan(OFO if mod == FO else OCX, L)
        assert fp64.rel_rms(y, o.forward(x)) <= REF_TOL


def test_concurrent_streams_without_work_buffers(fs):
    # four-step calls that pass no work buffer use library scratch; calls on
    # two streams at once must not share it
    import torch
    L, n = 61440, 24
    f = fs.create_fft(FO, L)
    rng = np.random.default_rng(77)
    xs = [rng.standard_normal((n, 2 * L)).astype(np.float32) for _ in range(2)]
    hs = [(rng.standard_normal((n, 2 * L)) * np.exp(-np.arange(2 * L) / (L / 4))).astype(np.float32)
          for _ in range(2)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    for x, h, s in zip(xs, hs, streams):
        with torch.cuda.stream(s):
            dx, dh = torch.from_numpy(x).cuda(), torch.from_numpy(h).cuda()
            k = torch.empty_like(dx)
            y = torch.empty_like(dx)
            f.execute(fftset_amd.OP_GET_KERNEL, n, k, dh, stream=s)
            for _ in range(3):  # keep both streams busy at the same time
                f.execute(fftset_amd.OP_CONV, n, y, dx, kern=k, stream=s)
            outs.append((y, dx, dh, k))
    torch.cuda.synchronize()
    for (y, _, _, _), x, h in zip(outs, xs, hs):
        ref = fp64.fo_conv64(x[[0, n - 1]], h[[0, n - 1]])
        assert np.max(fp64.rel_rms(y.cpu().numpy()[[0, n - 1]], ref)) <= fp64_tol(L)


@pytest.mark.parametrize("L", [960, 61440])
def test_drop_in_calls_from_concurrent_threads(fs, L):
    # execution is thread-safe when out / work differ (the reference's
    # contract): eight host threads on one plan, each with its own buffers
    import threading
    import torch
    from tests.gpu_util import dev, empty, host
    f = fs.create_fft(FO, L)
    rng = np.random.default_rng(L)
    xs = [rng.standard_normal(2 * L).astype(np.float32) for _ in range(8)]
    hs = [rng.standard_normal(2 * L).astype(np.float32) for _ in range(8)]
    expect = [(gpu_forward(f, x), gpu_conv(f, x, h)) for x, h in zip(xs, hs)]
    bufs = [(dev(x), dev(h), empty(2 * L), empty(2 * L), empty(2 * L), empty(2 * L)) for x, h in zip(xs, hs)]
    torch.cuda.synchronize()
    errors = []

    def worker(i):
        xi, hi, out, k, y, work = bufs[i]
        try:
            for _ in range(10):
                f.forward(out, xi, work)
                f.conv_get_kernel(k, hi)
                f.conv(y, xi, k, work)
        except Exception as e:  # pragma: no cover
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors
    for (xi, hi, out, k, y, work), (ef, ec) in zip(bufs, expect):
        assert np.array_equal(host(out), ef)
        assert np.array_equal(host(y), ec)


def test_scratch_growth_from_threads_on_one_stream():
    # work-less four-step calls from several host threads on one stream, with
    # growing batch sizes: a thread that grows the library's scratch must not
    # free the buffer another thread was just handed (outgrown buffers are
    # kept until fftset_destroy)
    import threading
    import torch
    s = fftset_amd.FFTSet()
    try:
        L = 61440
        f = s.create_fft(CX, L)
        rng = np.random.default_rng(91)
        sizes = [1, 2, 3, 4, 5, 6, 7, 8]
        xs = [torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda() for n in sizes]
        ys = [torch.empty_like(x) for x in xs]
        stream = torch.cuda.Stream()
        torch.cuda.synchronize()
        errors = []

        def worker(i):
            try:
                for _ in range(3):
                    f.execute(fftset_amd.OP_FORWARD, sizes[i], ys[i], xs[i], stream=stream.cuda_stream)
            except Exception as e:  # pragma: no cover
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(sizes))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        torch.cuda.synchronize()
        assert not errors
        for x, y in zip(xs, ys):
            xh, yh = x.cpu().numpy(), y.cpu().numpy()
            for r in (0, len(xh) - 1):
                assert fp64.rel_rms(yh[r], fp64.cx_forward64(xh[r])) <= fp64_tol(L)
    finally:
        s.destroy()


# the refused call is the only work in the capture, so the graph is empty
@pytest.mark.filterwarnings("ignore:The CUDA Graph is empty")
def test_scratch_growth_refused_under_capture():
    # growing the library's scratch inside a graph capture is refused with
    # FFTSET_E_INVALID (the allocation would not be part of the graph)
    import torch
    s = fftset_amd.FFTSet()
    try:
        L = 61440
        f = s.create_fft(CX, L)
        x = torch.zeros(2, 2 * L, device="cuda")
        y = torch.empty_like(x)
        stream = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with pytest.raises(fftset_amd.FFTSetError) as e:
            with torch.cuda.graph(g, stream=stream):
                f.execute(fftset_amd.OP_FORWARD, 2, y, x, stream=stream.cuda_stream)
        assert e.value.code == fftset_amd.E_INVALID
        # outside capture the same call allocates and runs
        f.execute(fftset_amd.OP_FORWARD, 2, y, x)
        torch.cuda.synchronize()
        assert np.all(y.cpu().numpy() == 0)
    finally:
        s.destroy()


def test_allocation_failure_is_reported(fs):
    # a work-less four-step batch far beyond HBM: the scratch allocation fails
    # and is reported before any kernel is launched; the set stays usable
    import torch
    L = 61440
    f = fs.create_fft(CX, L)
    x = torch.zeros(2 * L, device="cuda")
    with pytest.raises(fftset_amd.FFTSetError) as e:
        f.execute(fftset_amd.OP_FORWARD, 1 << 26, x, x, out_stride=2 * L, in_stride=2 * L)
    assert e.value.code == fftset_amd.E_NOMEM
    rng = np.random.default_rng(5)
    v = rng.standard_normal(2 * L).astype(np.float32)
    assert fp64.rel_rms(gpu_forward(f, v), fp64.cx_forward64(v)) <= fp64_tol(L)


@pytest.mark.parametrize("lds,mod,L,op", [
    (True, CX, 3840, "fwd"), (True, CX, 3840, "inv"), (True, FO, 4096, "conv"), (True, FO, 4096, "kern"),
    (True, FO, 2048, "conv"), (True, FO, 2048, "grad"), (True, FO, 2048, "kern"),
    (False, CX, 61440, "fwd"), (False, CX, 61440, "inv"), (False, FO, 61440, "fwd"), (False, FO, 61440, "inv"),
    (False, FO, 61440, "conv"), (False, CX, 61440, "grad"), (False, FO, 61440, "kern"),
    (False, FO, 15360, "conv"), (False, FO, 15360, "kern"), (False, CX, 3840, "fwd"), (False, FO, 4096, "conv"),
    (False, FO, 2048, "grad"), (False, CX, 4608, "fwd")])
def test_specialised_kernels_match_generic(tuned_fs, lds_tuned_fs, lds, mod, L, op):
    # compile-time specialised chains (kernels_fixed.hip, on a set planning
    # their chains; the four-step kernels' [16,15] x [16,16] split of 61440;
    # the column-register kernel of the A x 256 splits) against the generic
    # runtime-chain kernels on the same plan and inputs. The specialised
    # chains are the MI355X tune table's, so the plans come from a set that
    # loaded it.
    import torch
    f = (lds_tuned_fs if lds else tuned_fs).create_fft(mod, L)
    if lds:
        expect = {3840: [15, 16, 16], 4096: [16, 16, 16], 2048: [8, 16, 16]}[L]
    else:
        expect = {61440: [16, 15, 0, 16, 16], 15360: [15, 4, 0, 16, 16], 3840: [15, 0, 16, 16],
                  4096: [16, 0, 16, 16], 2048: [8, 0, 16, 16], 4608: [9, 2, 0, 16, 16]}[L]
    assert f.radices() == expect
    n = 64
    rng = np.random.default_rng(L + len(op))
    x = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    h = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)

    def run():
        y = torch.empty_like(x)
        if op == "fwd":
            f.execute(fftset_amd.OP_FORWARD, n, y, x)
        elif op == "inv":
            f.execute(fftset_amd.OP_INVERSE, n, y, x)
        elif op == "kern":
            f.execute(fftset_amd.OP_GET_KERNEL, n, y, x)
        elif op == "conv":
            f.execute(fftset_amd.OP_CONV, n, y, x, kern=k)
        else:
            f.execute(fftset_amd.OP_CONV_GRAD_INPUT, n, y, x, kern=k)
        torch.cuda.synchronize()
        return y.cpu().numpy()

    fftset_amd.kernel_counts()
    fixed = run()
    spec, gen = fftset_amd.kernel_counts()
    # the compile-time chain was chosen (at least one specialised launch)
    assert spec >= 1, (spec, gen)
    prev = fftset_amd.kernel_select(fftset_amd.SELECT_GENERIC)
    try:
        generic = run()
        spec2, gen2 = fftset_amd.kernel_counts()
    finally:
        fftset_amd.kernel_select(prev)
    assert spec2 == 0 and gen2 >= 1, (spec2, gen2)
    assert np.all(np.isfinite(fixed))
    assert np.max(fp64.rel_rms(fixed, generic)) <= 1e-6


@pytest.mark.parametrize("mod,op", [(CX, "fwd"), (FO, "fwd"), (CX, "inv"), (FO, "inv"), (FO, "conv"),
                                    (CX, "conv"), (FO, "grad")])
def test_one_workgroup_four_step_equals_split_launches(fs, mod, op):
    # 61440 bins: the opt-in one-workgroup-per-transform kernel (intermediate
    # in a per-workgroup scratch slot) against separate column and row
    # launches (intermediate in the work buffer): the same passes,
    # twiddles and order, so bit-identical; more transforms than workgroups,
    # so slots are reused
    import torch
    L, n = 61440, 300
    f = fs.create_fft(mod, L)
    rng = np.random.default_rng(n + len(op))
    x = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    h = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    code = {"fwd": fftset_amd.OP_FORWARD, "inv": fftset_amd.OP_INVERSE, "conv": fftset_amd.OP_CONV,
            "grad": fftset_amd.OP_CONV_GRAD_INPUT}[op]

    def run():
        y = torch.full_like(x, float("nan"))
        f.execute(code, n, y, x, kern=k if op in ("conv", "grad") else None)
        torch.cuda.synchronize()
        return y.cpu().numpy()

    prev = fftset_amd.kernel_select(fftset_amd.SELECT_FOUR_STEP_SPLIT)
    try:
        split = run()
    finally:
        fftset_amd.kernel_select(prev)
    prev = fftset_amd.kernel_select(fftset_amd.SELECT_FOUR_STEP_FUSED)
    try:
        fused = run()
    finally:
        fftset_amd.kernel_select(prev)
    assert np.all(np.isfinite(fused))
    assert np.array_equal(fused, split)
    xs = x.cpu().numpy()[[0, n - 1]]
    if op == "fwd":
        ref = fp64.fo_forward64(xs) if mod == FO else fp64.cx_forward64(xs)
    elif op == "conv":
        hs = h.cpu().numpy()[[0, n - 1]]
        ref = fp64.fo_conv64(xs, hs) if mod == FO else fp64.cx_conv64(xs, hs)
    else:
        return
    assert np.max(fp64.rel_rms(fused[[0, n - 1]], ref)) <= fp64_tol(L)


@pytest.mark.parametrize("mod,L", [(FO, 8), (FO, 120), (CX, 360), (FO, 2048), (FO, 4096), (CX, 20480),
                                   (FO, 61440)])
def test_conv_save_spectrum(fs, mod, L):
    # FFTSET_OP_CONV_SAVE_SPECTRUM: the same output as CONV (the same kernel,
    # so bit-identical), the signal's spectrum as get_kernel gives it, and
    # the kernel gradient as CONV_GRAD_INPUT with that spectrum
    import torch
    f = fs.create_fft(mod, L)
    n = 6
    rng = np.random.default_rng(L + 11)
    x = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    h = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    g = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    K, X, Xk = (torch.empty_like(x) for _ in range(3))
    y0, y1, gh0, gh1 = (torch.full_like(x, float("nan")) for _ in range(4))
    f.execute(fftset_amd.OP_GET_KERNEL, n, K, h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, Xk, x)
    f.execute(fftset_amd.OP_CONV, n, y0, x, kern=K)
    f.execute(fftset_amd.OP_CONV_SAVE_SPECTRUM, n, y1, x, kern=K, work=X)
    f.execute(fftset_amd.OP_CONV_GRAD_KERNEL, n, gh0, x, kern=g)
    f.execute(fftset_amd.OP_CONV_GRAD_INPUT, n, gh1, g, kern=X)
    torch.cuda.synchronize()
    assert np.array_equal(y0.cpu().numpy(), y1.cpu().numpy())
    assert np.max(fp64.rel_rms(X.cpu().numpy(), Xk.cpu().numpy())) <= 1e-6
    gh = gh1.cpu().numpy()
    assert np.max(fp64.rel_rms(gh, gh0.cpu().numpy())) <= 1e-6
    if mod == FO:
        xs, hs, gs = x.cpu().numpy()[[0, n - 1]], h.cpu().numpy()[[0, n - 1]], g.cpu().numpy()[[0, n - 1]]
        _, gh_ref = fp64.fo_conv_grads64(gs, xs, hs)
        assert np.max(fp64.rel_rms(gh[[0, n - 1]], gh_ref)) <= fp64_tol(L)
    with pytest.raises(fftset_amd.FFTSetError):  # the spectrum output is required
        f.execute(fftset_amd.OP_CONV_SAVE_SPECTRUM, n, y1, x, kern=K)
