"""Known-answer tests of the reference, restated (app_fftset_test/fftset_test.c).

The reference's only golden data are closed forms computed inline by its test
program: prime-impulse DFTs for both modulations and an integer-exact linear
convolution of prime indicator sequences, at 28 lengths, with absolute RMS
gates. These helpers build the same inputs and the same FP64 expectations so
the oracle and the GPU library can both be held to them.
"""
import numpy as np

# app_fftset_test/fftset_test.c:327-346
TEST_LENGTHS = [2, 3, 4, 5, 6, 8,
                9, 15,
                27, 45,
                12, 16, 20,
                24, 32, 40,
                48, 64, 80,
                96, 128, 160,
                18, 36, 72, 144, 288, 576]

# Gates (absolute RMS): fftset_test.c:135, 150, 216, 293, 308
FWD_RMS = 4e-6
INV_RMS = 1e-6
CONV_RMS = 1e-5


def _sieve(n):
    s = np.ones(n + 1, bool)
    s[:2] = False
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return np.nonzero(s)[0]


# fftset_test.c:27-68 is 0 followed by every prime up to 3989.
PRIMES = np.concatenate([[0], _sieve(3989)]).astype(np.int64)


def impulse_set(limit):
    """Indices i < limit with i in PRIMES (0 included)."""
    return PRIMES[PRIMES < limit]


def fo_impulse_case(L):
    """FREQOFFSETREAL prime impulse (fftset_test.c:70-156): N = 2L reals with
    ones at prime indices; expected spectrum from the FP64 closed form."""
    N = 2 * L
    x = np.zeros(N, np.float32)
    idx = impulse_set(N)
    x[idx] = 1.0
    k = np.arange(L)
    ang = -(2.0 * np.pi) * np.outer(k + 0.5, idx) / N
    X = np.empty(2 * L, np.float64)
    X[0::2] = np.cos(ang).sum(axis=1)
    X[1::2] = np.sin(ang).sum(axis=1)
    return x, X


def cx_impulse_case(L):
    """COMPLEX prime impulse (fftset_test.c:224-314): the interleaved 2L-float
    buffer has ones at prime float indices, i.e. real parts at even primes
    and imaginary parts at odd ones."""
    x = np.zeros(2 * L, np.float32)
    x[impulse_set(2 * L)] = 1.0
    z = x[0::2].astype(np.float64) + 1j * x[1::2].astype(np.float64)
    n = np.arange(L)
    W = np.exp(-2j * np.pi * (np.outer(n, n) % L) / L)
    Z = W @ z
    X = np.empty(2 * L, np.float64)
    X[0::2] = Z.real
    X[1::2] = Z.imag
    return x, X


def conv_case(L):
    """Integer convolution (fftset_test.c:158-222): kernel = 2/N at primes
    below N/4, signal = 1 at primes below 3N/4, expected = pair counts."""
    N = 2 * L
    h = np.zeros(N, np.float32)
    h[impulse_set(N // 4)] = 2.0 / N
    x = np.zeros(N, np.float32)
    x[impulse_set((N * 3) // 4)] = 1.0
    expect = np.zeros(N, np.float64)
    for a in impulse_set((N * 3) // 4):
        for b in impulse_set(N // 4):
            expect[a + b] += 1.0
    return x, h, expect


def abs_rms_bins(y, ref, nbins):
    """sqrt(sum of squared complex errors / nbins), the reference's metric."""
    d = np.asarray(y, np.float64) - np.asarray(ref, np.float64)
    return float(np.sqrt(np.sum(d * d) / nbins))
