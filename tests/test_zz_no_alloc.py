"""fftset_fft_execute_noalloc through the C ABI against fftset_fft_execute
(work = NULL) on the same plan and inputs: the same kernels, so the outputs
are bit-identical. Two calls in a row per case, the second on a smaller
batch, so the second runs over work space the first left behind."""
import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

FO, CX = fftset_amd.FREQ_OFFSET_REAL, fftset_amd.COMPLEX


@pytest.fixture(scope="module")
def fs():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


CASES = [
    # four-step plans: work space for every operation but get_kernel
    (FO, 61440, fftset_amd.OP_CONV, (9, 4)),
    (CX, 61440, fftset_amd.OP_FORWARD, (7, 2)),
    (FO, 61440, fftset_amd.OP_INVERSE, (5, 3)),
    (FO, 61440, fftset_amd.OP_CONV_GRAD_INPUT, (6, 1)),
    (FO, 61440, fftset_amd.OP_CONV_GRAD_KERNEL, (5, 2)),  # two regions
    (CX, 20480, fftset_amd.OP_CONV, (12, 5)),
    (FO, 61440, fftset_amd.OP_GET_KERNEL, (4, 2)),
    # LDS plans: work space only for the kernel gradient
    (FO, 2048, fftset_amd.OP_CONV_GRAD_KERNEL, (64, 17)),
    (CX, 3840, fftset_amd.OP_FORWARD, (33, 8)),
    (FO, 4096, fftset_amd.OP_CONV, (16, 3)),
]


@pytest.mark.parametrize("mod,L,op,counts", CASES)
def test_noalloc_matches_execute(fs, mod, L, op, counts):
    import torch
    f = fs.create_fft(mod, L)
    rng = np.random.default_rng(L + op)
    n0 = counts[0]
    x = torch.from_numpy(rng.standard_normal((n0, 2 * L)).astype(np.float32)).cuda()
    h = torch.from_numpy(rng.standard_normal((n0, 2 * L)).astype(np.float32)).cuda()
    k = torch.empty_like(h)
    f.execute(fftset_amd.OP_GET_KERNEL, n0, k, h)
    # CONV / GRAD_INPUT take a kernel spectrum; GRAD_KERNEL takes the output
    # gradient (h here) through kern
    kern = None
    if op in (fftset_amd.OP_CONV, fftset_amd.OP_CONV_GRAD_INPUT):
        kern = k
    elif op == fftset_amd.OP_CONV_GRAD_KERNEL:
        kern = h
    for n in counts:
        ref = torch.full((n, 2 * L), float("nan"), device="cuda")
        got = torch.full((n, 2 * L), float("nan"), device="cuda")
        kk = kern[:n] if kern is not None else None
        f.execute(op, n, ref, x[:n], kern=kk)
        f.execute_noalloc(op, n, got, x[:n], kern=kk)
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(r)), (n, "reference")
        assert np.array_equal(r, g), (n, float(np.max(np.abs(r - g))))


def test_noalloc_rejects_work_beyond_the_limit(fs):
    import torch
    L = 61440
    f = fs.create_fft(FO, L)
    per = 2 * L  # conv on a four-step plan: 2L floats per transform
    n_ok = fftset_amd.NOALLOC_WORK_FLOATS // per
    y = torch.zeros(1, device="cuda")
    with pytest.raises(fftset_amd.FFTSetError) as e:
        # rejected before any buffer is touched
        f.execute_noalloc(fftset_amd.OP_CONV, n_ok + 1, y, y, kern=y)
    assert e.value.code == fftset_amd.E_TOO_LARGE
    g = fs.create_fft(FO, 2048)
    with pytest.raises(fftset_amd.FFTSetError) as e:
        g.execute_noalloc(fftset_amd.OP_CONV_GRAD_KERNEL, fftset_amd.NOALLOC_WORK_FLOATS // 4096 + 1, y, y, kern=y)
    assert e.value.code == fftset_amd.E_TOO_LARGE
