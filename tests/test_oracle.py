"""Pins the oracle (CPU FP32 restatement + FP64 references) before anything is
compared against it: the reference's own known-answer tests, the mathematics
(closed forms, direct O(N^2) sums, identities), and planted mistakes that
each check must catch."""
import numpy as np
import pytest

from oracle import fp64
from oracle.oracle import MOD_COMPLEX as CX, MOD_FREQ_OFFSET_REAL as FO
from tests import kat


# ---------------------------------------------------------------------------
# FP64 references pinned by direct sums (the FFT-based forms are what the
# large-size tests use).

@pytest.mark.parametrize("L", [1, 2, 3, 5, 12, 45, 64])
def test_fp64_forward_forms_agree_with_direct_sums(L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(2 * L)
    assert np.allclose(fp64.cx_forward64(x), fp64.as_interleaved(fp64.dft64_direct(fp64.as_complex(x))),
                       atol=1e-10)
    assert np.allclose(fp64.fo_forward64(x), fp64.fo_forward64_direct(x), atol=1e-10)


@pytest.mark.parametrize("L", [1, 2, 3, 8, 15, 24])
def test_fp64_inverse_forms_are_inverses(L):
    rng = np.random.default_rng(100 + L)
    x = rng.standard_normal(2 * L)
    assert np.allclose(fp64.cx_inverse64(fp64.cx_forward64(x)), L * x, atol=1e-9)
    assert np.allclose(fp64.fo_inverse64(fp64.fo_forward64(x)), L * x, atol=1e-9)


@pytest.mark.parametrize("N", [2, 4, 6, 10, 32, 48])
def test_fp64_convolutions_agree_with_direct_sums(N):
    rng = np.random.default_rng(N)
    x, h = rng.standard_normal((2, N))
    assert np.allclose(fp64.negacyclic_conv64(x, h), fp64.negacyclic_conv64_direct(x, h), atol=1e-10)
    a = x + 1j * h
    b = h[::-1] - 1j * x
    assert np.allclose(fp64.cyclic_conv64(a, b), fp64.cyclic_conv64_direct(a, b), atol=1e-10)


def test_fp64_negacyclic_by_hand():
    # N = 2: y0 = x0 h0 - x1 h1, y1 = x0 h1 + x1 h0
    x = np.array([2.0, 3.0])
    h = np.array([5.0, 7.0])
    assert np.allclose(fp64.negacyclic_conv64_direct(x, h), [2 * 5 - 3 * 7, 2 * 7 + 3 * 5])
    assert np.allclose(fp64.negacyclic_conv64(x, h), [2 * 5 - 3 * 7, 2 * 7 + 3 * 5])


def test_fp64_odd_dft_special_cases():
    # impulse at n=0 gives a flat spectrum of ones; x[n] = cos(pi n (2k0+1)/N)
    # concentrates in bin k0 with weight N/2.
    L = 16
    N = 2 * L
    x = np.zeros(N)
    x[0] = 1
    X = fp64.as_complex(fp64.fo_forward64(x))
    assert np.allclose(X, 1.0)
    k0 = 5
    n = np.arange(N)
    x = np.cos(np.pi * n * (2 * k0 + 1) / N)
    X = fp64.as_complex(fp64.fo_forward64(x))
    expect = np.zeros(L, complex)
    expect[k0] = N / 2
    assert np.allclose(X, expect, atol=1e-9)


@pytest.mark.parametrize("N", [4, 6, 12, 20])
def test_fp64_adjoint_forms_agree_with_definition(N):
    rng = np.random.default_rng(7 * N)
    w, x, h = rng.standard_normal((3, N))
    gx, gh = fp64.fo_conv_grads64(w, x, h)
    gx2, gh2 = fp64.fo_conv_grads64_direct(w, x, h)
    assert np.allclose(gx, gx2, atol=1e-9)
    assert np.allclose(gh, gh2, atol=1e-9)
    # and against central differences of the bilinear loss
    loss = lambda xx, hh: float(np.sum(w * fp64.fo_conv64(xx, hh)))
    eps = 1e-3
    for i in range(0, N, 3):
        e = np.zeros(N)
        e[i] = eps
        assert abs((loss(x + e, h) - loss(x - e, h)) / (2 * eps) - gx[i]) < 1e-6 * (1 + abs(gx[i]))
        assert abs((loss(x, h + e) - loss(x, h - e)) / (2 * eps) - gh[i]) < 1e-6 * (1 + abs(gh[i]))


def test_rel_rms_catches_planted_sign_error():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(64)
    good = fp64.fo_forward64(x)
    bad = good.copy()
    bad[1::2] *= -1  # conjugated spectrum: the classic sign mistake
    assert fp64.rel_rms(good, good) == 0
    assert fp64.rel_rms(bad, good) > 0.1


# ---------------------------------------------------------------------------
# The FP32 restatement against the reference's known-answer tests.

@pytest.mark.parametrize("L", kat.TEST_LENGTHS)
def test_oracle_fo_prime_impulse_kat(oracle, L):
    pl = oracle.plan(FO, L)
    x, X = kat.fo_impulse_case(L)
    Y = pl.forward(x)
    assert kat.abs_rms_bins(Y, X, L) <= kat.FWD_RMS
    back = pl.inverse(Y.astype(np.float32)) / L
    assert kat.abs_rms_bins(back, x, L) <= kat.INV_RMS


@pytest.mark.parametrize("L", kat.TEST_LENGTHS)
def test_oracle_cx_prime_impulse_kat(oracle, L):
    pl = oracle.plan(CX, L)
    x, X = kat.cx_impulse_case(L)
    Y = pl.forward(x)
    assert kat.abs_rms_bins(Y, X, L) <= kat.FWD_RMS
    back = pl.inverse(Y.astype(np.float32)) / L
    # whole buffer, stricter than the reference's first-half check
    assert kat.abs_rms_bins(back, x, L) <= kat.INV_RMS


@pytest.mark.parametrize("L", kat.TEST_LENGTHS)
def test_oracle_conv_kat(oracle, L):
    pl = oracle.plan(FO, L)
    x, h, expect = kat.conv_case(L)
    K = pl.get_kernel(h)
    y = pl.conv(x, K)
    assert kat.abs_rms_bins(y, expect, L) <= kat.CONV_RMS


def test_oracle_in_place_aliasing(oracle):
    # in/out may alias for forward, inverse and conv (fftset.h:93-108)
    L = 48
    rng = np.random.default_rng(11)
    for mod in (FO, CX):
        pl = oracle.plan(mod, L)
        x = rng.standard_normal(2 * L).astype(np.float32)
        ref = pl.forward(x)
        buf = x.copy()
        work = np.empty(2 * L, np.float32)
        from oracle.oracle import _fptr
        pl._lib.oracle_forward(pl._h, _fptr(buf), _fptr(buf), _fptr(work))
        assert np.array_equal(buf, ref)


# ---------------------------------------------------------------------------
# The FP32 restatement against FP64 at random inputs and config sizes.

SIZES = [2, 3, 5, 6, 12, 45, 48, 120, 360, 512, 768, 1440, 2048, 3840, 4096, 4608]


@pytest.mark.parametrize("L", SIZES)
def test_oracle_matches_fp64(oracle, L):
    rng = np.random.default_rng(L)
    lg = max(np.log2(L), 1.0)
    # COMPLEX
    pl = oracle.plan(CX, L)
    x = rng.standard_normal(2 * L).astype(np.float32)
    assert fp64.rel_rms(pl.forward(x), fp64.cx_forward64(x)) < 1e-6 * lg
    assert fp64.rel_rms(pl.inverse(x), fp64.cx_inverse64(x)) < 1e-6 * lg
    h = rng.standard_normal(2 * L).astype(np.float32)
    y = pl.conv(x, pl.get_kernel(h))
    assert fp64.rel_rms(y, fp64.cx_conv64(x, h)) < 1e-6 * lg
    # FREQOFFSETREAL
    pl = oracle.plan(FO, L)
    x = rng.uniform(-1, 1, 2 * L).astype(np.float32)
    X = pl.forward(x)
    assert fp64.rel_rms(X, fp64.fo_forward64(x)) < 1e-6 * lg
    assert fp64.rel_rms(pl.inverse(X), fp64.fo_inverse64(X)) < 1e-6 * lg
    h = (rng.standard_normal(2 * L) * np.exp(-np.arange(2 * L) / (L / 4))).astype(np.float32)
    y = pl.conv(x, pl.get_kernel(h))
    assert fp64.rel_rms(y, fp64.fo_conv64(x, h)) < 1e-6 * lg


def test_oracle_fo_conv_is_negacyclic_not_cyclic(oracle):
    # a full-block signal wraps: the FREQOFFSETREAL conv must match the
    # negacyclic form and must not match the cyclic one
    L = 64
    N = 2 * L
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, N).astype(np.float32)
    h = rng.standard_normal(N).astype(np.float32)
    pl = oracle.plan(FO, L)
    y = pl.conv(x, pl.get_kernel(h))
    neg = L * fp64.negacyclic_conv64_direct(x, h)
    cyc = L * np.real(fp64.cyclic_conv64_direct(x.astype(complex), h.astype(complex)))
    assert fp64.rel_rms(y, neg) < 1e-6
    assert fp64.rel_rms(y, cyc) > 0.1


# ---------------------------------------------------------------------------
# Planner and the conv-length helper.

def test_oracle_planner_scalar_chains(oracle):
    # scalar chains listed for the reference at these sizes (min-cost search,
    # first strictly cheaper candidate wins)
    assert oracle.plan_chain(120) == [2, 3, 4, 5]
    assert oracle.plan_chain(360) == [2, 3, 3, 4, 5]
    # leaves: a length equal to a leaf radix is a single pass
    for r in (2, 3, 4, 5, 6, 8, 16):
        assert oracle.plan_chain(r) == [r]
    # every chain multiplies back to its length, with 5/6/8/16 only last
    for n in (12, 48, 96, 576, 960, 3840, 7680, 61440):
        c = oracle.plan_chain(n)
        assert int(np.prod(c)) == n
        assert all(r in (2, 3, 4) for r in c[:-1])


@pytest.mark.parametrize("n", [7, 25, 49, 11 * 4, 0, 1])
def test_oracle_rejects_unsupported(oracle, n):
    assert oracle.plan_chain(n) == []
    with pytest.raises(ValueError):
        oracle.plan(FO, n)


def test_recommend_conv_length_known_values(oracle):
    # values worked by hand from src/fftset.c:131-167
    assert oracle.recommend_conv_length(1, 0) == 16
    assert oracle.recommend_conv_length(100, 0) == 512
    assert oracle.recommend_conv_length(512, 4096) == 3072
    assert oracle.recommend_conv_length(4096, 0) == 16 * 2048   # ceil(36863/32) = 1152 -> 2048


def test_recommend_conv_length_invariants(oracle):
    for k in (1, 2, 3, 17, 64, 100, 333, 1000, 4096):
        for b in (0, 1, 100, 5000, 65536):
            r = oracle.recommend_conv_length(k, b)
            assert r % 16 == 0
            # real length 2r covers kernel + block - 1
            assert 2 * r >= k + max(b, 8 * k) - 1
            # r/16 is a product of 2s, 3s and 4s
            q = r // 16
            for p in (2, 3):
                while q % p == 0:
                    q //= p
            assert q == 1
            # and the library can plan it
            assert oracle.plan_chain(r) != []
