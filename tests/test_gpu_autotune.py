"""The autotuner: picks a measured chain / packing / split, later plans use
it, results stay at parity, and tables round-trip through a file."""
import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from oracle import fp64
from oracle.oracle import MOD_COMPLEX as OCX, MOD_FREQ_OFFSET_REAL as OFO
from tests.gpu_util import gpu_conv, gpu_forward, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mod,L", [(fftset_amd.COMPLEX, 3840), (fftset_amd.FREQ_OFFSET_REAL, 4608),
                                   (fftset_amd.FREQ_OFFSET_REAL, 61440)])
def test_autotune_then_plan(oracle, tmp_path, mod, L):
    if torch_cuda() is None:
        pytest.skip("no GPU")
    fs = fftset_amd.FFTSet()
    try:
        before = fs.table_bytes
        ms = fs.autotune(mod, L)
        assert ms > 0
        assert fs.table_bytes == before  # the candidates' tables are freed
        f = fs.create_fft(mod, L)
        assert fs.table_bytes > before
        radices = f.radices()
        with pytest.raises(fftset_amd.FFTSetError):
            fs.autotune(mod, L)  # plan exists: immutable
        rng = np.random.default_rng(L)
        x = rng.standard_normal(2 * L).astype(np.float32)
        h = rng.standard_normal(2 * L).astype(np.float32)
        o = oracle.plan(OFO if mod == fftset_amd.FREQ_OFFSET_REAL else OCX, L)
        assert fp64.rel_rms(gpu_forward(f, x), o.forward(x)) <= 2e-6
        assert fp64.rel_rms(gpu_conv(f, x, h), o.conv(x, o.get_kernel(h))) <= 2e-6
        path = str(tmp_path / "tune.txt")
        fs.tune_save(path)
    finally:
        fs.destroy()
    # a fresh set that loads the table plans the same chain
    fs2 = fftset_amd.FFTSet()
    try:
        fs2.tune_load(path)
        assert fs2.create_fft(mod, L).radices() == radices
    finally:
        fs2.destroy()


def test_autotune_ops_mask(oracle):
    if torch_cuda() is None:
        pytest.skip("no GPU")
    fs = fftset_amd.FFTSet()
    try:
        with pytest.raises(fftset_amd.FFTSetError):
            fs.autotune(fftset_amd.FREQ_OFFSET_REAL, 768, ops=(fftset_amd.OP_CONV_GRAD_KERNEL,))
        assert fs.autotune(fftset_amd.FREQ_OFFSET_REAL, 768, ops=(fftset_amd.OP_CONV,)) > 0
        f = fs.create_fft(fftset_amd.FREQ_OFFSET_REAL, 768)
        rng = np.random.default_rng(768)
        x = rng.standard_normal(2 * 768).astype(np.float32)
        h = rng.standard_normal(2 * 768).astype(np.float32)
        o = oracle.plan(OFO, 768)
        assert fp64.rel_rms(gpu_conv(f, x, h), o.conv(x, o.get_kernel(h))) <= 2e-6
    finally:
        fs.destroy()


def test_autotune_keeps_the_tables_of_existing_plans(oracle):
    # tuning one size must not free tables that a plan made before it uses
    if torch_cuda() is None:
        pytest.skip("no GPU")
    fs = fftset_amd.FFTSet()
    try:
        f = fs.create_fft(fftset_amd.FREQ_OFFSET_REAL, 4096)
        before = fs.table_bytes
        assert before > 0
        fs.autotune(fftset_amd.FREQ_OFFSET_REAL, 4608, ops=(fftset_amd.OP_CONV,))
        assert fs.table_bytes == before
        rng = np.random.default_rng(4096)
        x = rng.standard_normal(2 * 4096).astype(np.float32)
        h = rng.standard_normal(2 * 4096).astype(np.float32)
        o = oracle.plan(OFO, 4096)
        assert fp64.rel_rms(gpu_conv(f, x, h), o.conv(x, o.get_kernel(h))) <= 2e-6
        g = fs.create_fft(fftset_amd.FREQ_OFFSET_REAL, 4608)
        o2 = oracle.plan(OFO, 4608)
        x2 = rng.standard_normal(2 * 4608).astype(np.float32)
        h2 = rng.standard_normal(2 * 4608).astype(np.float32)
        assert fp64.rel_rms(gpu_conv(g, x2, h2), o2.conv(x2, o2.get_kernel(h2))) <= 2e-6
    finally:
        fs.destroy()
