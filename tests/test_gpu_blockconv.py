"""Block (overlap-save) FIR convolution on the GPU: checked against FP64
direct convolution (np.convolve) and, for a small case, against the oracle's
FREQOFFSETREAL convolution applied block by block (the reference's own
convolution path composed as the driver composes it)."""
import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from oracle import fp64
from oracle.oracle import MOD_FREQ_OFFSET_REAL as OFO
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fs():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


def _taps(k, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(k) * np.exp(-np.arange(k) / max(1.0, k / 4))).astype(np.float32)


def _fir64(x, h):
    """y[n] = sum_{j<k} h[j] x[n-j] for n < len(x), FP64."""
    return np.convolve(x.astype(np.float64), h.astype(np.float64))[: len(x)]


@pytest.mark.parametrize("k,max_block,extra", [(1024, 0, 123), (7, 0, 5), (1, 0, 0), (300, 5000, 1), (2000, 0, 77)])
def test_blockconv_matches_fp64(fs, oracle, k, max_block, extra):
    h = _taps(k, k)
    bc = fs.blockconv(torch.from_numpy(h).cuda(), max_block)
    L = bc.complex_bins
    assert L == oracle.recommend_conv_length(k, max_block)
    b = bc.block_size
    assert b == 2 * L - k + 1
    C = 3
    n = 3 * b + extra
    rng = np.random.default_rng(k + 1)
    x = rng.uniform(-1, 1, (C, n)).astype(np.float32)
    y = torch.full((C, n), float("nan"), device="cuda")
    bc.run(y, torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    yv = y.cpu().numpy()
    assert np.all(np.isfinite(yv))
    for c in range(C):
        ref = _fir64(x[c], h)
        assert fp64.rel_rms(yv[c], ref) <= 1e-6 * np.log2(2 * L), c
    bc.destroy()


@pytest.mark.parametrize("case", ["partial_single_frame", "exact_blocks", "history_strided", "ragged_tail"])
def test_blockconv_specialised_frame_bounds(case):
    # the specialised [16,8,8,8] frame kernel reads frames inside the signal
    # without range checks and the first/last frames of a channel with them;
    # every case is checked against FP64 direct convolution, and the padding
    # past each channel's signal must stay untouched
    import os
    if torch_cuda() is None:
        pytest.skip("no GPU")
    tune = os.path.join(os.path.dirname(fftset_amd.__file__), "tune_mi355x.txt")
    s = fftset_amd.FFTSet(tune_file=tune)
    k = 1024
    h = _taps(k, 29)
    bc = s.blockconv(torch.from_numpy(h).cuda(), 0)
    b = bc.block_size
    assert bc.complex_bins == 8192
    n = {"partial_single_frame": b // 2, "exact_blocks": 2 * b, "history_strided": 2 * b + 77,
         "ragged_tail": 4 * b + 1}[case]
    C, pad = 3, 96
    rng = np.random.default_rng({"partial_single_frame": 1, "exact_blocks": 2, "history_strided": 3, "ragged_tail": 4}[case])
    x = rng.standard_normal((C, n)).astype(np.float32)
    stride = n + pad
    xs = torch.zeros((C, stride), device="cuda")
    xs[:, :n] = torch.from_numpy(x).cuda()
    y = torch.full((C, stride), float("nan"), device="cuda")
    fftset_amd.kernel_counts()
    if case == "history_strided":
        # two calls; the second reads the first part's last k-1 samples as history
        n1 = b + 333
        bc.run(y, xs, nsamples=n1, channels=C, in_stride=stride, out_stride=stride)
        bc.run(y.data_ptr() + 4 * n1, xs.data_ptr() + 4 * n1, nsamples=n - n1, channels=C, in_stride=stride,
               out_stride=stride, flags=fftset_amd.BLOCKCONV_HISTORY)
        assert fftset_amd.kernel_counts() == (2, 0)
    else:
        bc.run(y, xs, nsamples=n, channels=C, in_stride=stride, out_stride=stride)
        assert fftset_amd.kernel_counts() == (1, 0)
    torch.cuda.synchronize()
    yv = y.cpu().numpy()
    for c in range(C):
        assert fp64.rel_rms(yv[c, :n], _fir64(x[c], h)) <= 1e-6 * np.log2(2 * bc.complex_bins), (case, c)
        assert np.all(np.isnan(yv[c, n:])), (case, c)
    bc.destroy()
    s.destroy()


def test_blockconv_history_and_strides(fs):
    # a signal filtered in two calls, the second reading the first's tail as
    # history, equals one call over the whole signal
    k = 257
    h = _taps(k, 3)
    bc = fs.blockconv(torch.from_numpy(h).cuda(), 0)
    b = bc.block_size
    C, n1, n2 = 2, 2 * b + 17, b + 5
    pad = 64  # channel stride larger than the signal
    rng = np.random.default_rng(7)
    x = rng.standard_normal((C, n1 + n2)).astype(np.float32)
    xs = torch.zeros((C, n1 + n2 + pad), device="cuda")
    xs[:, : n1 + n2] = torch.from_numpy(x).cuda()
    y = torch.full((C, n1 + n2 + pad), float("nan"), device="cuda")
    stride = n1 + n2 + pad
    bc.run(y, xs, nsamples=n1, channels=C, in_stride=stride, out_stride=stride)
    # second part: in points at sample n1, the k-1 samples before it are history
    bc.run(y.data_ptr() + 4 * n1, xs.data_ptr() + 4 * n1, nsamples=n2, channels=C, in_stride=stride,
           out_stride=stride, flags=fftset_amd.BLOCKCONV_HISTORY)
    torch.cuda.synchronize()
    yv = y.cpu().numpy()
    for c in range(C):
        ref = _fir64(x[c], h)
        assert fp64.rel_rms(yv[c, : n1 + n2], ref) <= 1e-6 * np.log2(2 * bc.complex_bins)
        assert np.all(np.isnan(yv[c, n1 + n2:]))  # nothing written past the signal
    bc.destroy()


def test_blockconv_matches_oracle_blocks(fs, oracle):
    # overlap-save assembled from the oracle's FP32 FREQOFFSETREAL convolution
    k = 40
    h = _taps(k, 11)
    bc = fs.blockconv(torch.from_numpy(h).cuda(), 0)
    L, b = bc.complex_bins, bc.block_size
    N = 2 * L
    n = 4 * b + 9
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n).astype(np.float32)
    y = torch.empty(1, n, device="cuda")
    bc.run(y, torch.from_numpy(x).cuda().reshape(1, n))
    torch.cuda.synchronize()
    o = oracle.plan(OFO, L)
    hk = np.zeros(N, np.float32)
    hk[:k] = h / np.float32(L)
    K = o.get_kernel(hk)
    ref = np.zeros(n, np.float32)
    for i in range((n + b - 1) // b):
        s0 = i * b - (k - 1)
        frame = np.zeros(N, np.float32)
        lo, hi = max(0, s0), min(n, s0 + N)
        frame[lo - s0: hi - s0] = x[lo:hi]
        out = o.conv(frame, K)
        m = min(b, n - i * b)
        ref[i * b: i * b + m] = out[k - 1: k - 1 + m]
    assert fp64.rel_rms(y.cpu().numpy()[0], ref) <= 2e-6
    bc.destroy()


def test_blockconv_rejects_bad_use(fs):
    with pytest.raises(fftset_amd.FFTSetError):
        fs.blockconv(torch.zeros(5000, device="cuda"), 0)  # 2L = 65536: beyond the LDS engine
    bc = fs.blockconv(torch.ones(3, device="cuda"), 0)
    x = torch.zeros(2, 100, device="cuda")
    with pytest.raises(fftset_amd.FFTSetError):
        bc.run(x, x, nsamples=100, channels=2, in_stride=50, out_stride=100)  # stride < nsamples
    with pytest.raises(fftset_amd.FFTSetError):
        bc.run(x, x, flags=8)
    bc.destroy()


def test_blockconv_specialised_chain_matches_generic():
    # with the tuned table, 1024 taps give L = 8192 on the chain [16,8,8,8],
    # which runs a compile-time specialised frame kernel; it must agree with
    # the generic runtime-chain kernel and with FP64 direct convolution
    import os
    if torch_cuda() is None:
        pytest.skip("no GPU")
    tune = os.path.join(os.path.dirname(fftset_amd.__file__), "tune_mi355x.txt")
    s = fftset_amd.FFTSet(tune_file=tune)
    k = 1024
    h = _taps(k, 19)
    bc = s.blockconv(torch.from_numpy(h).cuda(), 0)
    L = bc.complex_bins
    assert L == 8192
    assert list(s.create_fft(fftset_amd.FREQ_OFFSET_REAL, L).radices()) == [16, 8, 8, 8]
    C, n = 3, 3 * bc.block_size + 321
    rng = np.random.default_rng(23)
    x = rng.standard_normal((C, n)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()

    def run():
        y = torch.full((C, n), float("nan"), device="cuda")
        bc.run(y, xd)
        torch.cuda.synchronize()
        return y.cpu().numpy()

    fftset_amd.kernel_counts()
    fixed = run()
    assert fftset_amd.kernel_counts() == (1, 0)  # the [16,8,8,8] frame kernel ran
    prev = fftset_amd.kernel_select(fftset_amd.SELECT_GENERIC)
    try:
        generic = run()
        assert fftset_amd.kernel_counts() == (0, 1)
    finally:
        fftset_amd.kernel_select(prev)
    assert np.all(np.isfinite(fixed))
    for c in range(C):
        assert fp64.rel_rms(fixed[c], generic[c]) <= 1e-6, c
        assert fp64.rel_rms(fixed[c], _fir64(x[c], h)) <= 1e-6 * np.log2(2 * L), c
    bc.destroy()
    s.destroy()
