"""Plants one mistake at a time in the oracle's C restatement (built with
-DOR_PLANT=n into a temporary directory) and shows that the checks pinning
the oracle -- the reference's known-answer tests and the FP64 references --
reject every planted variant."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import fp64
from oracle import oracle as orc
from tests import kat

SRC = os.path.join(os.path.dirname(orc.__file__), "fftset_oracle.c")

PLANTS = {
    1: "inner-pass twiddle conjugated",
    2: "FREQOFFSETREAL modulation conjugated",
    3: "odd-bin conjugate dropped in the bin interleave",
    4: "pointwise product not conjugated",
    5: "Stockham output index shifted",
    6: "radix-5 term dropped",
}


def _build(tmp_path, plant):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    so = os.path.join(tmp_path, f"liboracle_plant{plant}.so")
    subprocess.run([cc, "-O1", "-std=gnu11", "-fPIC", "-shared", f"-DOR_PLANT={plant}",
                    "-o", so, SRC, "-lm", "-lpthread"], check=True)
    return so


def _load(so):
    saved = orc._lib, orc.ORACLE_SO
    try:
        orc._lib = None
        orc.ORACLE_SO = so
        return orc.load_oracle()
    finally:
        orc._lib, orc.ORACLE_SO = saved


def _suite_passes(lib):
    """True when every pinning check passes for this build."""
    FO, CX = orc.MOD_FREQ_OFFSET_REAL, orc.MOD_COMPLEX
    for L in (5, 20, 45, 48, 120, 576):
        fo = orc.OraclePlan(lib, FO, L)
        cx = orc.OraclePlan(lib, CX, L)
        x, X = kat.fo_impulse_case(L)
        Y = fo.forward(x)
        if kat.abs_rms_bins(Y, X, L) > kat.FWD_RMS:
            return False
        if kat.abs_rms_bins(fo.inverse(Y) / L, x, L) > kat.INV_RMS:
            return False
        x, X = kat.cx_impulse_case(L)
        Y = cx.forward(x)
        if kat.abs_rms_bins(Y, X, L) > kat.FWD_RMS:
            return False
        x, h, expect = kat.conv_case(L)
        if kat.abs_rms_bins(fo.conv(x, fo.get_kernel(h)), expect, L) > kat.CONV_RMS:
            return False
        rng = np.random.default_rng(L)
        x = rng.standard_normal(2 * L).astype(np.float32)
        h = rng.standard_normal(2 * L).astype(np.float32)
        if fp64.rel_rms(fo.conv(x, fo.get_kernel(h)), fp64.fo_conv64(x, h)) > 1e-5:
            return False
        if fp64.rel_rms(cx.conv(x, cx.get_kernel(h)), fp64.cx_conv64(x, h)) > 1e-5:
            return False
    return True


def test_unplanted_build_passes(tmp_path):
    assert _suite_passes(_load(_build(str(tmp_path), 0)))


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_caught(tmp_path, plant):
    lib = _load(_build(str(tmp_path), plant))
    assert not _suite_passes(lib), PLANTS[plant]
