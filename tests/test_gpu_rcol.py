"""The column-register four-step kernel (kernels_rcol_*.hip): 15360- and
4608-bin transforms as 60 x 256 and 18 x 256 splits whose columns each live
in one lane's VGPRs and whose rows pass through LDS in batches of 15 and 9. Checked against the
two-launch four-step path of the same split (FFTSET_SELECT_FOUR_STEP_SPLIT:
same decomposition, twiddle tables and spectrum layout, so equal to
rounding), against the LDS engine's plan of the same size (a different
chain: equal to rounding for every layout-free result), against FP64 closed
forms, and inside ragged batches against the drop-in calls."""
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
SIZES = [15360, 4608, 4096, 3840, 2048, 768]  # 60, 18, 16, 15, 8 and 3 x 256
OPS = {"fwd": fftset_amd.OP_FORWARD, "inv": fftset_amd.OP_INVERSE, "kern": fftset_amd.OP_GET_KERNEL,
       "conv": fftset_amd.OP_CONV, "grad": fftset_amd.OP_CONV_GRAD_INPUT}


def fp64_tol(n):
    return 1e-6 * np.log2(n)


def _decay(L):
    return np.exp(-np.arange(2 * L) / (L / 15.0))


@pytest.fixture(scope="module")
def fs(tmp_path_factory):
    """A set whose 15360- and 4608-bin plans are the A x 256 splits (tune
    table entries with A and no chain)."""
    if torch_cuda() is None:
        pytest.skip("no GPU")
    path = tmp_path_factory.mktemp("rcol") / "tune.txt"
    path.write_text("".join(f"{m} {n} 0 {n // 256} 0 0\n" for n in SIZES for m in (0, 1)))
    s = fftset_amd.FFTSet(tune_file=str(path))
    yield s
    s.destroy()


@pytest.fixture(scope="module")
def fs_lds():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


def _data(L, n, seed):
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal((n, 2 * L)).astype(np.float32)).cuda()
    h = torch.from_numpy((rng.standard_normal((n, 2 * L)) * _decay(L)).astype(np.float32)).cuda()
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


def _fp64(mod, op, xs, hs):
    if op == "fwd":
        return fp64.cx_forward64(xs) if mod == CX else fp64.fo_forward64(xs)
    if op == "inv":
        return fp64.cx_inverse64(xs) if mod == CX else fp64.fo_inverse64(xs)
    if op == "conv":
        return fp64.fo_conv64(xs, hs) if mod == FO else fp64.cx_conv64(xs, hs)
    return None


@pytest.mark.parametrize("L", SIZES)
def test_rcol_plan(fs, L):
    for mod in (CX, FO):
        r = fs.create_fft(mod, L).radices()
        assert r[-3:] == [0, 16, 16], r
        assert int(np.prod(r[:-3])) == L // 256, r


@pytest.mark.parametrize("mod,op", [(CX, "fwd"), (CX, "inv"), (FO, "fwd"), (FO, "inv"), (CX, "kern"),
                                    (FO, "kern"), (CX, "conv"), (FO, "conv"), (CX, "grad"), (FO, "grad")])
@pytest.mark.parametrize("L", SIZES)
def test_rcol_equals_split_launches(fs, mod, op, L):
    # 1500 transforms: more than five workgroups per CU over the grid
    n = 1500
    f = fs.create_fft(mod, L)
    x, h = _data(L, n, 7 + len(op) + len(mod))
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    torch.cuda.synchronize()
    kk = k if op in ("conv", "grad") else None
    fftset_amd.kernel_counts()
    res = _run(f, op, x, kk)
    spec, gen = fftset_amd.kernel_counts()
    assert (spec, gen) == (1, 0), (spec, gen)  # one launch of the column-register kernel
    split = _split(lambda: _run(f, op, x, kk))
    r, s = res.cpu().numpy(), split.cpu().numpy()
    assert np.all(np.isfinite(r))
    assert np.max(fp64.rel_rms(r, s)) <= 1e-6, np.max(fp64.rel_rms(r, s))
    sel = [0, n // 2, n - 1]
    ref = _fp64(mod, op, x.cpu().numpy()[sel], h.cpu().numpy()[sel])
    if ref is not None:
        assert np.max(fp64.rel_rms(r[sel], ref)) <= fp64_tol(L)


@pytest.mark.parametrize("mod", [CX, FO])
@pytest.mark.parametrize("L", SIZES)
def test_rcol_equals_lds_plan(fs, fs_lds, mod, L):
    # layout-free results agree with the LDS engine's plan of the same size
    n = 8
    f, g = fs.create_fft(mod, L), fs_lds.create_fft(mod, L)
    assert g.radices()[-3:] != [0, 16, 16]
    x, h = _data(L, n, 19 + len(mod))
    kf, kg = torch.empty_like(h), torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, kf, h)
    g.execute(fftset_amd.OP_GET_KERNEL, n, kg, h)
    for op in ("fwd", "inv", "conv", "grad"):
        a = _run(f, op, x, kf if op in ("conv", "grad") else None).cpu().numpy()
        b = _run(g, op, x, kg if op in ("conv", "grad") else None).cpu().numpy()
        assert np.max(fp64.rel_rms(a, b)) <= 2e-6, op


@pytest.mark.parametrize("mod", [FO, CX])
@pytest.mark.parametrize("L", SIZES)
def test_rcol_spectra_interchange_with_split_launches(fs, mod, L):
    n = 4
    f = fs.create_fft(mod, L)
    x, h = _data(L, n, 31 + len(mod))
    k_res = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k_res, h)
    k_split = torch.empty_like(h)
    _split(lambda: f.execute(fftset_amd.OP_GET_KERNEL, n, k_split, h))
    torch.cuda.synchronize()
    assert np.max(fp64.rel_rms(k_res.cpu().numpy(), k_split.cpu().numpy())) <= 1e-6
    ref = (fp64.fo_conv64 if mod == FO else fp64.cx_conv64)(x.cpu().numpy(), h.cpu().numpy())
    y1 = _run(f, "conv", x, k_split)
    y2 = _split(lambda: _run(f, "conv", x, k_res))
    for y in (y1, y2):
        assert np.max(fp64.rel_rms(y.cpu().numpy(), ref)) <= fp64_tol(L)


def _at(L, n, off, seed, scale=None):
    # n transforms back to back starting `off` floats into a fresh buffer
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n * 2 * L + 64)
    if scale is not None:
        v[off:off + n * 2 * L] *= np.tile(scale, n)
    return torch.from_numpy(v.astype(np.float32)).cuda()[off:off + n * 2 * L].view(n, 2 * L)


def _run_into(f, op, y, x, k=None):
    y.fill_(float("nan"))
    f.execute(OPS[op], x.shape[0], y, x, kern=k)
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("mod", [CX, FO])
@pytest.mark.parametrize("off_in,off_out", [(2, 2), (16, 0), (0, 26)])
@pytest.mark.parametrize("L", SIZES)
def test_rcol_misaligned_buffers(fs, mod, off_in, off_out, L):
    n = 3
    f = fs.create_fft(mod, L)
    x = _at(L, n, off_in, 70 + off_in)
    h = _at(L, n, off_in, 71 + off_in, _decay(L))
    k = _at(L, n, off_in, 72)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    torch.cuda.synchronize()
    ks = _split(lambda: _run_into(f, "kern", _at(L, n, off_out, 73), h))
    assert np.max(fp64.rel_rms(k.cpu().numpy(), ks.cpu().numpy())) <= 1e-6
    xs, hs = x.cpu().numpy(), h.cpu().numpy()
    for op in ("fwd", "inv", "conv", "grad"):
        kk = k if op in ("conv", "grad") else None
        y = _run_into(f, op, _at(L, n, off_out, 74), x, kk)
        ys = _split(lambda: _run_into(f, op, _at(L, n, off_out, 75), x, kk))
        r = y.cpu().numpy()
        assert np.all(np.isfinite(r)), op
        assert np.max(fp64.rel_rms(r, ys.cpu().numpy())) <= 1e-6, op
        ref = _fp64(mod, op, xs, hs)
        if ref is not None:
            assert np.max(fp64.rel_rms(r, ref)) <= fp64_tol(L), op


@pytest.mark.parametrize("mod,op", [(CX, "fwd"), (FO, "inv"), (FO, "conv"), (CX, "kern")])
@pytest.mark.parametrize("L", SIZES)
def test_rcol_in_place(fs, mod, op, L):
    # out == in: every lane has read its whole column before the first
    # barrier, and outputs are written only after it
    n = 3
    f = fs.create_fft(mod, L)
    x, h = _data(L, n, 41)
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    ref = _run(f, op, x, k if op == "conv" else None).cpu().numpy()
    y = x.clone()
    f.execute(OPS[op], n, y, y, kern=k if op == "conv" else None)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), ref)


@pytest.mark.parametrize("L", SIZES)
def test_rcol_shared_spectrum(fs, L):
    # one spectrum for every transform (kernel stride 0): cached loads
    n = 5
    f = fs.create_fft(FO, L)
    x, h = _data(L, n, 47)
    k = torch.empty_like(h[:1])
    f.execute(fftset_amd.OP_GET_KERNEL, 1, k, h[:1])
    y = torch.full_like(x, float("nan"))
    f.execute(fftset_amd.OP_CONV, n, y, x, kern=k, kern_stride=0)
    torch.cuda.synchronize()
    ref = fp64.fo_conv64(x.cpu().numpy(), np.repeat(h[:1].cpu().numpy(), n, axis=0))
    assert np.max(fp64.rel_rms(y.cpu().numpy(), ref)) <= fp64_tol(L)


@pytest.mark.parametrize("L", SIZES)
def test_rcol_save_spectrum_and_kernel_gradient(fs, L):
    n = 2
    f = fs.create_fft(FO, L)
    x, h = _data(L, n, 53)
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
    gx_ref, gh_ref = fp64.fo_conv_grads64(w.cpu().numpy(), x.cpu().numpy(), h.cpu().numpy())
    assert np.max(fp64.rel_rms(gh.cpu().numpy(), gh_ref)) <= fp64_tol(L)
    gx = _run(f, "grad", w, k)
    assert np.max(fp64.rel_rms(gx.cpu().numpy(), gx_ref)) <= fp64_tol(L)


def test_rcol_noalloc_needs_no_work(fs):
    # no intermediate in HBM: the no-allocation entry point takes counts
    # beyond what its 64 MiB work array could hold for a four-step plan
    L, n = 15360, 600  # 600 x 2L floats = 70 MiB of work rows on the two-launch path
    f = fs.create_fft(CX, L)
    x, h = _data(L, n, 61)
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n, k, h)
    y1 = _run(f, "conv", x, k)
    y2 = torch.full_like(x, float("nan"))
    f.execute_noalloc(fftset_amd.OP_CONV, n, y2, x, kern=k)
    torch.cuda.synchronize()
    assert np.array_equal(y1.cpu().numpy(), y2.cpu().numpy())


def test_rcol_ragged_batch_equals_drop_in(fs):
    L = 15360
    # a ragged batch whose 15360- and 4608-bin classes run the column-register kernel
    # with per-transform pointers, beside LDS classes: their outputs equal
    # the drop-in calls bit for bit (the LDS classes run the ragged kernel,
    # equal to rounding)
    rng = np.random.default_rng(5)
    spec = [(CX, L, fftset_amd.OP_FORWARD, 5), (FO, L, fftset_amd.OP_CONV, 6), (FO, L, fftset_amd.OP_INVERSE, 2),
            (CX, L, fftset_amd.OP_CONV_GRAD_INPUT, 2), (FO, 4608, fftset_amd.OP_CONV, 4),
            (CX, 4608, fftset_amd.OP_FORWARD, 3), (CX, 48, fftset_amd.OP_FORWARD, 9), (FO, 768, fftset_amd.OP_CONV, 5)]
    b = fs.batch()
    jobs = []
    for mod, n, op, count in spec:
        f = fs.create_fft(mod, n)
        for _ in range(count):
            x = torch.from_numpy(rng.standard_normal(2 * n).astype(np.float32)).cuda()
            hh = torch.from_numpy((rng.standard_normal(2 * n) * np.exp(-np.arange(2 * n) / 512.0))
                                  .astype(np.float32)).cuda()
            k = None
            if op in (fftset_amd.OP_CONV, fftset_amd.OP_CONV_GRAD_INPUT):
                k = torch.empty_like(x)
                f.conv_get_kernel(k, hh)
            y = torch.full_like(x, float("nan"))
            b.add(f, op, y, x, k)
            jobs.append((f, op, x, k, y))
    b.commit()
    b.launch()
    torch.cuda.synchronize()
    for f, op, x, k, y in jobs:
        ref = torch.full_like(x, float("nan"))
        f.execute(op, 1, ref, x, kern=k)
        torch.cuda.synchronize()
        if f.L in SIZES:
            assert np.array_equal(y.cpu().numpy(), ref.cpu().numpy()), (f.L, op)
        else:
            assert fp64.rel_rms(y.cpu().numpy(), ref.cpu().numpy()) <= 2e-6, (f.L, op)
