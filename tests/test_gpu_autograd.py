"""The torch.autograd harness over the C-ABI adjoints, at the differentiable
convolution configuration: FREQOFFSETREAL complex_bins=2048, 4096 channels,
x ~ N(0,1), h ~ N(0,1) exp(-t/256), loss = sum(w * conv(x, h)), w ~ N(0,1).
Gradients are checked against an FP64 reference gradient (every channel)
and against FP64 central finite differences (eps = 1e-3) at 256 sampled
coordinates per input."""
import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from oracle import fp64
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fs():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


def _inputs(L, C, seed):
    rng = np.random.default_rng(seed)
    N = 2 * L
    x = rng.standard_normal((C, N)).astype(np.float32)
    h = (rng.standard_normal((C, N)) * np.exp(-np.arange(N) / 256.0)).astype(np.float32)
    w = rng.standard_normal((C, N)).astype(np.float32)
    return x, h, w


def _fd(f64_loss_row, base, idx, eps=1e-3):
    e = np.zeros_like(base)
    e[idx] = eps
    return (f64_loss_row(base + e) - f64_loss_row(base - e)) / (2 * eps)


def test_differentiable_conv_config(fs):
    from fftset_amd.autograd import fft_conv
    L, C = 2048, 4096
    x, h, w = _inputs(L, C, 5)
    f = fs.create_fft(fftset_amd.FREQ_OFFSET_REAL, L)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    ht = torch.from_numpy(h).cuda().requires_grad_(True)
    wt = torch.from_numpy(w).cuda()
    y = fft_conv(xt, ht, f)
    loss = (wt * y).sum()
    loss.backward()
    gx = xt.grad.cpu().numpy()
    gh = ht.grad.cpu().numpy()
    yv = y.detach().cpu().numpy()

    # forward against FP64 on a sample of channels
    sample = np.arange(0, C, 97)
    assert np.max(fp64.rel_rms(yv[sample], fp64.fo_conv64(x[sample], h[sample]))) <= 1e-6 * np.log2(L)

    # FP64 reference gradient, every channel
    gx_ref, gh_ref = fp64.fo_conv_grads64(w, x, h)
    tol = 1e-6 * np.log2(L)
    assert np.max(fp64.rel_rms(gx, gx_ref)) <= tol
    assert np.max(fp64.rel_rms(gh, gh_ref)) <= tol

    # FP64 central differences of the bilinear loss at 256 coordinates per input
    rng = np.random.default_rng(9)
    for which, grad, gref in (("x", gx, gx_ref), ("h", gh, gh_ref)):
        chans = rng.integers(0, C, 256)
        coords = rng.integers(0, 2 * L, 256)
        for c, i in zip(chans, coords):
            xc, hc, wc = x[c].astype(np.float64), h[c].astype(np.float64), w[c].astype(np.float64)
            if which == "x":
                fd = _fd(lambda v: float(np.sum(wc * fp64.fo_conv64(v, hc))), xc, i)
            else:
                fd = _fd(lambda v: float(np.sum(wc * fp64.fo_conv64(xc, v))), hc, i)
            scale = np.sqrt(np.mean(gref[c] ** 2))
            # FD is exact for a bilinear loss up to FP64 rounding
            assert abs(fd - gref[c, i]) <= 1e-7 * scale + 1e-9
            assert abs(grad[c, i] - fd) <= 2e-5 * scale, (which, c, i)


def test_shared_kernel_and_chain_rule(fs):
    # one kernel for all channels, loss = sum(y^2): dL/dy = 2y through autograd
    from fftset_amd.autograd import fft_conv
    L, C = 120, 8
    x, h, _ = _inputs(L, C, 3)
    h1 = h[:1]
    f = fs.create_fft(fftset_amd.FREQ_OFFSET_REAL, L)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    ht = torch.from_numpy(h1).cuda().requires_grad_(True)
    y = fft_conv(xt, ht, f)
    (y * y).sum().backward()
    y64 = fp64.fo_conv64(x, np.broadcast_to(h1, x.shape))
    gx_ref, gh_ref = fp64.fo_conv_grads64(2 * y64, x, np.broadcast_to(h1, x.shape))
    assert np.max(fp64.rel_rms(xt.grad.cpu().numpy(), gx_ref)) <= 1e-5
    assert fp64.rel_rms(ht.grad.cpu().numpy()[0], gh_ref.sum(axis=0)) <= 1e-5


def test_harness_rejects_wrong_inputs(fs):
    from fftset_amd.autograd import fft_conv
    f = fs.create_fft(fftset_amd.FREQ_OFFSET_REAL, 16)
    with pytest.raises(ValueError):
        fft_conv(torch.zeros(3, 30, device="cuda"), torch.zeros(3, 32, device="cuda"), f)
    with pytest.raises(TypeError):
        fft_conv(torch.zeros(3, 32, device="cuda", dtype=torch.float64), torch.zeros(3, 32, device="cuda"), f)
