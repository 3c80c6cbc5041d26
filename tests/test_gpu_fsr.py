"""The CU-resident kernel for 61440-bin transforms (kernels_fsr_*.hip): the
whole 240 x 256 transform on one CU, one launch per batch. Checked against
the two-launch four-step path (FFTSET_SELECT_FOUR_STEP_SPLIT: same
decomposition and twiddle tables, different work split, so equal to
rounding), against FP64 closed forms, and for the kernel-spectrum layout the
two paths share: a spectrum from either path convolves correctly on the
other."""
import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from oracle import fp64
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

FO, CX = fftset_amd.FREQ_OFFSET_REAL, fftset_amd.COMPLEX
L = 61440
OPS = {"fwd": fftset_amd.OP_FORWARD, "inv": fftset_amd.OP_INVERSE, "kern": fftset_amd.OP_GET_KERNEL,
       "conv": fftset_amd.OP_CONV, "grad": fftset_amd.OP_CONV_GRAD_INPUT}


def fp64_tol(n):
    return 1e-6 * np.log2(n)


@pytest.fixture(scope="module")
def fs():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    h = torch.from_numpy((rng.standard_normal((n, 2 * L)) *
                          np.exp(-np.arange(2 * L) / 4096.0)).astype(np.float32)).cuda()
    return x, h


def _run(f, op, x, k=None):
    y = torch.full_like(x, float("nan"))
    f.execute(OPS[op], x.shape[0], y, x, kern=k)
    torch.cuda.synchronize()
    return y


def _split(fn):
    prev = fftset_amd.kernel_select(fftset_amd.SELECT_FOUR_STEP_SPLIT)
    try:
        return fn()
    finally:
        fftset_amd.kernel_select(prev)


@pytest.mark.parametrize("mod,op", [(CX, "fwd"), (CX, "inv"), (CX, "kern"), (FO, "kern"), (CX, "conv"),
                                    (FO, "conv"), (CX, "grad"), (FO, "grad")])
def test_resident_equals_split_launches(fs, mod, op):
    # 300 transforms: more than one workgroup per CU over the grid
    n = 300
    f = fs.create_fft(mod, L)
    assert f.radices() == [16, 15, 0, 16, 16]
    x, h = _data(n, 7 + len(op) + len(mod))
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    torch.cuda.synchronize()
    kk = k if op in ("conv", "grad") else None
    fftset_amd.kernel_counts()
    res = _run(f, op, x, kk)
    spec, gen = fftset_amd.kernel_counts()
    assert (spec, gen) == (1, 0), (spec, gen)  # one launch of the resident kernel
    split = _split(lambda: _run(f, op, x, kk))
    r, s = res.cpu().numpy(), split.cpu().numpy()
    assert np.all(np.isfinite(r))
    assert np.max(fp64.rel_rms(r, s)) <= 1e-6, np.max(fp64.rel_rms(r, s))
    # FP64 closed forms on the first and last transform
    xs = x.cpu().numpy()[[0, n - 1]]
    hs = h.cpu().numpy()[[0, n - 1]]
    if op == "fwd":
        ref = fp64.cx_forward64(xs)
    elif op == "inv":
        ref = fp64.cx_inverse64(xs)
    elif op == "conv":
        ref = fp64.fo_conv64(xs, hs) if mod == FO else fp64.cx_conv64(xs, hs)
    else:
        return
    assert np.max(fp64.rel_rms(r[[0, n - 1]], ref)) <= fp64_tol(L)


@pytest.mark.parametrize("mod", [FO, CX])
def test_spectra_interchange_with_split_launches(fs, mod):
    # the kernel spectrum is in the same slot-major layout on both paths
    n = 4
    f = fs.create_fft(mod, L)
    x, h = _data(n, 31 + len(mod))
    k_res = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k_res, h)
    k_split = torch.empty_like(h)
    _split(lambda: f.execute(fftset_amd.OP_GET_KERNEL, n, k_split, h))
    torch.cuda.synchronize()
    assert np.max(fp64.rel_rms(k_res.cpu().numpy(), k_split.cpu().numpy())) <= 1e-6
    ref = (fp64.fo_conv64 if mod == FO else fp64.cx_conv64)(x.cpu().numpy(), h.cpu().numpy())
    y1 = _run(f, "conv", x, k_split)                       # split spectrum, resident conv
    y2 = _split(lambda: _run(f, "conv", x, k_res))         # resident spectrum, split conv
    for y in (y1, y2):
        assert np.max(fp64.rel_rms(y.cpu().numpy(), ref)) <= fp64_tol(L)


def _at(n, off, seed, scale=None):
    # n transforms back to back starting `off` floats into a fresh buffer
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n * 2 * L + 64)
    if scale is not None:
        v[off:off + n * 2 * L] *= np.tile(scale, n)
    return torch.from_numpy(v.astype(np.float32)).cuda()[off:off + n * 2 * L].view(n, 2 * L)


@pytest.mark.parametrize("mod", [CX, FO])
@pytest.mark.parametrize("off_in,off_out", [(2, 2), (16, 16), (30, 30), (16, 0), (0, 26)])
def test_resident_misaligned_buffers(fs, mod, off_in, off_out):
    # buffers that start off a 128-byte line (the column chunks then start
    # at another column): every operation against the two-launch path and
    # FP64, with input and output offsets equal or different
    n = 3
    f = fs.create_fft(mod, L)
    decay = np.exp(-np.arange(2 * L) / 4096.0)
    x = _at(n, off_in, 70 + off_in)
    h = _at(n, off_in, 71 + off_in, decay)
    k = _at(n, off_in, 72)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    torch.cuda.synchronize()
    ks = _split(lambda: _run_into(f, "kern", _at(n, off_out, 73), h))
    assert np.max(fp64.rel_rms(k.cpu().numpy(), ks.cpu().numpy())) <= 1e-6
    xs, hs = x.cpu().numpy(), h.cpu().numpy()
    for op in ("fwd", "inv", "conv", "grad"):
        kk = k if op in ("conv", "grad") else None
        y = _run_into(f, op, _at(n, off_out, 74), x, kk)
        ys = _split(lambda: _run_into(f, op, _at(n, off_out, 75), x, kk))
        r = y.cpu().numpy()
        assert np.all(np.isfinite(r)), op
        assert np.max(fp64.rel_rms(r, ys.cpu().numpy())) <= 1e-6, op
        if op == "fwd":
            ref = fp64.cx_forward64(xs) if mod == CX else fp64.fo_forward64(xs)
        elif op == "inv":
            ref = fp64.cx_inverse64(xs) if mod == CX else fp64.fo_inverse64(xs)
        elif op == "conv":
            ref = fp64.fo_conv64(xs, hs) if mod == FO else fp64.cx_conv64(xs, hs)
        else:
            continue
        assert np.max(fp64.rel_rms(r, ref)) <= fp64_tol(L), op


def _run_into(f, op, y, x, k=None):
    y.fill_(float("nan"))
    f.execute(OPS[op], x.shape[0], y, x, kern=k)
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("mod,op", [(CX, "fwd"), (FO, "conv"), (CX, "kern")])
def test_resident_in_place(fs, mod, op):
    # out == in (the reference allows aliasing): the columns read the input
    # chunk by chunk, and the outputs are written only after every chunk
    n = 3
    f = fs.create_fft(mod, L)
    x, h = _data(n, 41)
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    ref = _run(f, op, x, k if op == "conv" else None).cpu().numpy()
    y = x.clone()
    f.execute(OPS[op], n, y, y, kern=k if op == "conv" else None)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), ref)


def test_resident_save_spectrum_and_kernel_gradient(fs):
    # CONV_SAVE_SPECTRUM: bit-identical output to CONV, the spectrum of get_kernel;
    # CONV_GRAD_KERNEL (get_kernel then the adjoint) against FP64
    n = 2
    f = fs.create_fft(FO, L)
    x, h = _data(n, 53)
    w = torch.from_numpy(np.random.default_rng(54).standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    y = _run(f, "conv", x, k)
    y2 = torch.empty_like(x)
    spec = torch.empty_like(x)
    f.execute(fftset_amd.OP_CONV_SAVE_SPECTRUM, n, y2, x, kern=k, work=spec)
    kx = torch.empty_like(x)
    f.execute(fftset_amd.OP_GET_KERNEL, n, kx, x)
    gh = torch.empty_like(x)
    f.execute(fftset_amd.OP_CONV_GRAD_KERNEL, n, gh, x, kern=w)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), y2.cpu().numpy())
    assert np.max(fp64.rel_rms(spec.cpu().numpy(), kx.cpu().numpy())) <= 1e-6
    _, gh_ref = fp64.fo_conv_grads64(w.cpu().numpy(), x.cpu().numpy(), h.cpu().numpy())
    assert np.max(fp64.rel_rms(gh.cpu().numpy(), gh_ref)) <= fp64_tol(L)


def test_resident_noalloc_get_kernel(fs):
    # the no-allocation entry point provides the resident get_kernel's spill rows
    n = 5
    f = fs.create_fft(CX, L)
    x, h = _data(n, 61)
    k1 = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k1, h)
    k2 = torch.full_like(h, float("nan"))
    f.execute_noalloc(fftset_amd.OP_GET_KERNEL, n, k2, h)
    torch.cuda.synchronize()
    assert np.array_equal(k1.cpu().numpy(), k2.cpu().numpy())
