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


def _is_2a3b(n):
    while n % 2 == 0:
        n //= 2
    while n % 3 == 0:
        n //= 3
    return n == 1


def test_recommend_conv_length_exhaustive(oracle):
    # every kernel length up to 65536 against a sweep of block sizes, through
    # the C ABI: bit-exact with the oracle (src/fftset.c:131-167) and the
    # invariants the formula guarantees -- a multiple of 16 bins of the form
    # 16 * 2^a 3^b whose real length 2r holds k + max(b, 8k) - 1 samples
    lib = fftset_amd.load_library()
    olib = oracle.lib
    for b in (0, 1, 31, 32, 100, 1000, 4096, 8191, 65536, 1 << 20, 1 << 24):
        prev = 0
        for k in range(1, 65537):
            r = lib.fftset_recommend_conv_length(k, b)
            assert r == olib.oracle_recommend_conv_length(k, b), (k, b)
            if r != prev:  # the invariants depend on r only through its value
                assert r % 16 == 0 and _is_2a3b(r // 16), (k, b, r)
                prev = r
            assert 2 * r >= k + max(b, 8 * k) - 1, (k, b, r)


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


# The reference's schedules at the config sizes (its planner at its vector
# widths, src/fftset_vec.c:1004-1062; widths chosen by
# src/fftset_mod_freqoffsetreal.c:1012-1131 and src/fftset_mod_cplx.c:320-347):
# modulation, L -> (first-pass split, inner chain).
_REF_SCHEDULES = [
    ("fo", 512, [4, 2], [4, 16]),
    ("fo", 2048, [4, 2], [4, 4, 16]),
    ("fo", 4096, [4, 2], [4, 4, 4, 8]),
    ("cx", 3840, [4], [3, 4, 4, 4, 5]),
    ("cx", 48, [4], [3, 4]), ("fo", 48, [4], [3, 4]),
    ("cx", 120, [4], [2, 3, 5]), ("fo", 120, [], [2, 3, 4, 5]),
    ("cx", 360, [4], [2, 3, 3, 5]), ("fo", 360, [], [2, 3, 3, 4, 5]),
    ("cx", 768, [4], [3, 4, 16]), ("fo", 768, [4, 2], [4, 4, 6]),
    ("cx", 1440, [4], [2, 3, 3, 4, 5]), ("fo", 1440, [4, 2], [3, 3, 4, 5]),
    ("cx", 4608, [4], [3, 4, 4, 4, 6]), ("fo", 4608, [4, 2], [3, 3, 4, 16]),
    ("cx", 15360, [4], [3, 4, 4, 4, 4, 5]), ("fo", 15360, [4, 2], [2, 3, 4, 4, 4, 5]),
    ("cx", 61440, [4], [3, 4, 4, 4, 4, 4, 5]), ("fo", 61440, [4, 2], [2, 3, 4, 4, 4, 4, 5]),
]


@pytest.mark.parametrize("mod,L,split,inner", _REF_SCHEDULES)
def test_baseline_chain_is_the_reference_schedule(mod, L, split, inner):
    m = fftset_amd.FREQ_OFFSET_REAL if mod == "fo" else fftset_amd.COMPLEX
    assert fftset_amd.baseline_chain(m, L) == (split, inner)


def test_baseline_chain_matches_the_oracle_planner_at_width_one(oracle):
    # at vector width 1 the reference's planner is the oracle's (which
    # restates it for the scalar path): every width-1 structure agrees
    for L in [2, 3, 5, 6, 9, 10, 15, 18, 27, 45, 90, 120, 135, 270, 360, 405, 1215, 3645]:
        split, inner = fftset_amd.baseline_chain(fftset_amd.FREQ_OFFSET_REAL, L)
        assert split == [] and inner == oracle.plan_chain(L), L
