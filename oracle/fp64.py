"""FP64 closed-form references for the fftset transform / convolution path.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Conventions (``L`` = complex bins, every buffer holds ``2L`` floats):

* COMPLEX forward: ``X[k] = sum_{n<L} x[n] e^{-2 pi i nk/L}``, interleaved
  re/im in natural order; inverse is unnormalised so ``inv(fwd(x)) = L x``
  (fftset_test.c:261-311).
* FREQOFFSETREAL forward (``N = 2L`` reals in):
  ``X[k] = sum_{n<N} x[n] e^{-2 pi i n(k+1/2)/N}``, ``k < L``, interleaved
  (fftset.h:62-72); inverse ``y[n] = Re sum_k X[k] e^{+2 pi i n(k+1/2)/N}``
  which returns ``L x`` (fftset_test.c:142-147).
* FREQOFFSETREAL convolution: ``L * negacyclic_N(x, h)``; COMPLEX
  convolution: ``L * cyclic_L(x, h)``.

Every function accepts a batch in the leading dimensions.
"""
import numpy as np

__all__ = [
    "as_complex", "as_interleaved", "cx_forward64", "cx_inverse64",
    "fo_forward64", "fo_inverse64", "negacyclic_conv64", "negacyclic_conv64_direct",
    "cyclic_conv64", "cyclic_conv64_direct", "fo_conv64", "cx_conv64",
    "negacyclic_flip", "fo_conv_grads64", "fo_conv_grads64_direct", "rel_rms",
    "rel_max", "dft64_direct", "fo_forward64_direct",
]


def as_complex(x):
    """(…, 2L) interleaved floats -> (…, L) complex128."""
    x = np.asarray(x, dtype=np.float64)
    return x[..., 0::2] + 1j * x[..., 1::2]


def as_interleaved(z):
    """(…, L) complex -> (…, 2L) float64 interleaved."""
    z = np.asarray(z)
    out = np.empty(z.shape[:-1] + (2 * z.shape[-1],), np.float64)
    out[..., 0::2] = z.real
    out[..., 1::2] = z.imag
    return out


def dft64_direct(z):
    """O(L^2) complex DFT in FP64 (pins the FFT-based forms on small sizes)."""
    z = np.asarray(z, dtype=np.complex128)
    L = z.shape[-1]
    n = np.arange(L)
    W = np.exp(-2j * np.pi * ((np.outer(n, n) % L) / L))
    return z @ W.T


def cx_forward64(x):
    return as_interleaved(np.fft.fft(as_complex(x), axis=-1))


def cx_inverse64(X):
    Z = as_complex(X)
    return as_interleaved(np.fft.ifft(Z, axis=-1) * Z.shape[-1])


def fo_forward64(x):
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[-1]
    L = N // 2
    n = np.arange(N)
    mod = np.exp(-1j * np.pi * n / N)
    return as_interleaved(np.fft.fft(x * mod, axis=-1)[..., :L])


def fo_forward64_direct(x):
    """O(N L) odd-frequency DFT straight from its definition."""
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[-1]
    L = N // 2
    n = np.arange(N)
    k = np.arange(L)
    # exponent n*(k+1/2)/N reduced mod 1 exactly through integers: n*(2k+1) mod 2N
    ph = (np.outer(k, n) * 2 + n[None, :]) % (2 * N)
    W = np.exp(-1j * np.pi * ph / N)
    return as_interleaved(x @ W.T)


def fo_inverse64(X):
    Z = as_complex(X)
    L = Z.shape[-1]
    N = 2 * L
    pad = np.zeros(Z.shape[:-1] + (N,), np.complex128)
    pad[..., :L] = Z
    n = np.arange(N)
    y = np.fft.ifft(pad, axis=-1) * N * np.exp(1j * np.pi * n / N)
    return y.real


def negacyclic_conv64_direct(x, h):
    """y[n] = sum_{m<=n} x[m] h[n-m] - sum_{m>n} x[m] h[N+n-m], O(N^2)."""
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    N = x.shape[-1]
    idx = (np.arange(N)[:, None] - np.arange(N)[None, :])  # n - m
    sign = np.where(idx >= 0, 1.0, -1.0)
    H = h[..., idx % N] * sign  # (…, n, m)
    return np.einsum("...nm,...m->...n", H, x)


def negacyclic_conv64(x, h):
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    N = x.shape[-1]
    th = np.exp(1j * np.pi * np.arange(N) / N)
    c = np.fft.ifft(np.fft.fft(x * th, axis=-1) * np.fft.fft(h * th, axis=-1), axis=-1)
    return (c * np.conj(th)).real


def cyclic_conv64_direct(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    L = a.shape[-1]
    idx = (np.arange(L)[:, None] - np.arange(L)[None, :]) % L
    return np.einsum("...nm,...m->...n", b[..., idx], a)


def cyclic_conv64(a, b):
    return np.fft.ifft(np.fft.fft(a, axis=-1) * np.fft.fft(b, axis=-1), axis=-1)


def fo_conv64(x, h):
    """FREQOFFSETREAL fftset_fft_conv output: L * negacyclic_{2L}(x, h)."""
    x = np.asarray(x, dtype=np.float64)
    L = x.shape[-1] // 2
    return L * negacyclic_conv64(x, h)


def cx_conv64(x, h):
    """COMPLEX fftset_fft_conv output: L * cyclic_L(x, h), interleaved."""
    a = as_complex(x)
    b = as_complex(h)
    return as_interleaved(a.shape[-1] * cyclic_conv64(a, b))


def negacyclic_flip(h):
    """h_flat[0] = h[0], h_flat[k] = -h[N-k] (the negacyclic adjoint of h)."""
    h = np.asarray(h, dtype=np.float64)
    out = np.empty_like(h)
    out[..., 0] = h[..., 0]
    out[..., 1:] = -h[..., :0:-1]
    return out


def fo_conv_grads64(w, x, h):
    """Gradients of sum(w * fo_conv(x, h)) w.r.t. x and h (FP64)."""
    L = np.asarray(x).shape[-1] // 2
    gx = L * negacyclic_conv64(w, negacyclic_flip(h))
    gh = L * negacyclic_conv64(w, negacyclic_flip(x))
    return gx, gh


def fo_conv_grads64_direct(w, x, h):
    """Same gradients from the definition: d/dx[m] sum_n w[n] L sum_m' ..."""
    w = np.asarray(w, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    N = x.shape[-1]
    L = N // 2
    n = np.arange(N)[:, None]
    m = np.arange(N)[None, :]
    d = n - m
    sign = np.where(d >= 0, 1.0, -1.0)
    # y[n] = L sum_m x[m] h[(n-m) mod N] sign(n-m)
    Hm = h[..., d % N] * sign          # dy[n]/dx[m] / L
    gx = L * np.einsum("...n,...nm->...m", w, Hm)
    # dy[n]/dh[k] = L x[(n-k) mod N] sign(n-k)
    Xm = x[..., d % N] * sign          # indexed [n, k]
    gh = L * np.einsum("...n,...nk->...k", w, Xm)
    return gx, gh


def rel_rms(y, ref):
    """sqrt(sum|y-ref|^2 / sum|ref|^2) over the last axis (per transform)."""
    y = np.asarray(y, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    num = np.sum((y - ref) ** 2, axis=-1)
    den = np.sum(ref ** 2, axis=-1)
    return np.sqrt(num / np.maximum(den, 1e-300))


def rel_max(y, ref):
    """max|y-ref| / max|ref| over the last axis (per transform)."""
    y = np.asarray(y, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.max(np.abs(y - ref), axis=-1) / np.maximum(np.max(np.abs(ref), axis=-1), 1e-300)
