"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and its host-only helper matches the oracle."""
import ctypes
import re
import subprocess

import pytest

import fftset_amd

pytestmark = pytest.mark.skipif(not fftset_amd.lib_available(), reason="libfftset_amd.so not built")


def _header_symbols():
    text = open(fftset_amd.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    funcs = set(re.findall(r"\b(fftset_\w+)\s*\(", text))
    data = set(re.findall(r"extern\s+const\s+struct\s+\w+\s*\*\s*(\w+)\s*;", text))
    return funcs, data


def test_header_declares_the_reference_entry_points():
    funcs, data = _header_symbols()
    for name in ("fftset_init", "fftset_destroy", "fftset_create_fft", "fftset_fft_forward",
                 "fftset_fft_inverse", "fftset_fft_conv_get_kernel", "fftset_fft_conv",
                 "fftset_recommend_conv_length"):
        assert name in funcs, name
    assert data == {"FFTSET_MODULATION_FREQ_OFFSET_REAL", "FFTSET_MODULATION_COMPLEX"}


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(fftset_amd.LIB_PATH)
    funcs, data = _header_symbols()
    missing = [n for n in sorted(funcs | data) if not hasattr(lib, n)]
    assert not missing, missing


def test_modulation_handles_are_distinct_and_non_null():
    fo = fftset_amd._modulation(fftset_amd.FREQ_OFFSET_REAL)
    cx = fftset_amd._modulation(fftset_amd.COMPLEX)
    assert fo and cx and fo != cx


def test_recommend_conv_length_matches_oracle(oracle):
    for k in list(range(1, 70)) + [100, 127, 128, 129, 333, 512, 1000, 1023, 1024, 4096, 10000]:
        for b in (0, 1, 7, 64, 100, 1000, 4096, 5000, 65536):
            assert fftset_amd.recommend_conv_length(k, b) == oracle.recommend_conv_length(k, b), (k, b)


def test_null_arguments_are_rejected_without_a_device():
    lib = fftset_amd.load_library()
    assert lib.fftset_init(None) != 0
    # execution on a NULL plan records an error instead of crashing
    lib.fftset_fft_forward(None, None, None, None)
    assert lib.fftset_last_error() == fftset_amd.E_INVALID
    assert lib.fftset_last_error() == fftset_amd.OK


def test_header_and_library_link_from_plain_c(tmp_path):
    # the drop-in header is C (not C++) and the library links into a C program
    import __graft_entry__ as ge
    out = tmp_path / "dropin_host"
    r = subprocess.run(ge.c_dropin_command(str(out)), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.exists()
