"""ctypes binding of the FP32 CPU restatement in ``fftset_oracle.c``.

TEST INFRASTRUCTURE ONLY (see the package docstring).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")

MOD_FREQ_OFFSET_REAL = 0
MOD_COMPLEX = 1

OP_FORWARD = 0
OP_INVERSE = 1
OP_GETKERN = 2
OP_CONV = 3

_lib = None


def oracle_available():
    return os.path.exists(ORACLE_SO)


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(
            "oracle/liboracle.so is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(ORACLE_SO)
    vp, fp, u = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_uint
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [ctypes.c_int, u]
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_destroy.restype = None
    for name in ("oracle_forward", "oracle_inverse"):
        f = getattr(lib, name)
        f.argtypes = [vp, fp, fp, fp]
        f.restype = None
    lib.oracle_get_kernel.argtypes = [vp, fp, fp]
    lib.oracle_get_kernel.restype = None
    lib.oracle_conv.argtypes = [vp, fp, fp, fp, fp]
    lib.oracle_conv.restype = None
    lib.oracle_recommend_conv_length.argtypes = [u, u]
    lib.oracle_recommend_conv_length.restype = u
    lib.oracle_plan_chain.argtypes = [u, ctypes.POINTER(u), u]
    lib.oracle_plan_chain.restype = u
    lib.oracle_plan_passes.argtypes = [vp, ctypes.POINTER(u), u]
    lib.oracle_plan_passes.restype = u
    lib.oracle_run_batch.argtypes = [vp, ctypes.c_int, u, fp, fp, fp, u]
    lib.oracle_run_batch.restype = ctypes.c_int
    _lib = lib
    return lib


def _fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class OraclePlan:
    """One planned transform of ``L`` complex bins for a modulation."""

    def __init__(self, lib, mod, L):
        self._lib = lib
        self.mod = mod
        self.L = int(L)
        self._h = lib.oracle_create(mod, self.L)
        if not self._h:
            raise ValueError(f"unsupported size {L}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.oracle_destroy(h)
            self._h = None

    def radices(self):
        buf = (ctypes.c_uint * 64)()
        n = self._lib.oracle_plan_passes(self._h, buf, 64)
        return [buf[i] for i in range(n)]

    def _buf(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape[-1] == 2 * self.L, (x.shape, self.L)
        return x

    def forward(self, x):
        x = self._buf(x)
        out = np.empty_like(x)
        work = np.empty(2 * self.L, np.float32)
        for o, i in zip(out.reshape(-1, 2 * self.L), x.reshape(-1, 2 * self.L)):
            self._lib.oracle_forward(self._h, _fptr(o), _fptr(i), _fptr(work))
        return out

    def inverse(self, x):
        x = self._buf(x)
        out = np.empty_like(x)
        work = np.empty(2 * self.L, np.float32)
        for o, i in zip(out.reshape(-1, 2 * self.L), x.reshape(-1, 2 * self.L)):
            self._lib.oracle_inverse(self._h, _fptr(o), _fptr(i), _fptr(work))
        return out

    def get_kernel(self, h):
        h = self._buf(h)
        out = np.empty_like(h)
        for o, i in zip(out.reshape(-1, 2 * self.L), h.reshape(-1, 2 * self.L)):
            self._lib.oracle_get_kernel(self._h, _fptr(o), _fptr(i))
        return out

    def conv(self, x, kern):
        x = self._buf(x)
        kern = self._buf(kern)
        out = np.empty_like(x)
        work = np.empty(2 * self.L, np.float32)
        xs = x.reshape(-1, 2 * self.L)
        ks = np.broadcast_to(kern.reshape(-1, 2 * self.L), xs.shape)
        for o, i, k in zip(out.reshape(-1, 2 * self.L), xs, ks):
            k = np.ascontiguousarray(k)
            self._lib.oracle_conv(self._h, _fptr(o), _fptr(i), _fptr(k), _fptr(work))
        return out

    def run_batch(self, op, x, kern=None, nthreads=1):
        """Runs ``op`` over every row of ``x`` (n, 2L) on ``nthreads`` threads,
        one transform per thread at a time."""
        x = self._buf(x)
        n = x.size // (2 * self.L)
        out = np.empty_like(x)
        kp = None
        if op == OP_CONV:
            kern = self._buf(kern)
            assert kern.size == x.size
            kp = _fptr(kern)
        rc = self._lib.oracle_run_batch(self._h, op, n, _fptr(out), _fptr(x), kp, nthreads)
        if rc:
            raise RuntimeError("oracle batch failed")
        return out


class Oracle:
    """Factory for oracle plans; mirrors ``fftset_create_fft``."""

    FREQ_OFFSET_REAL = MOD_FREQ_OFFSET_REAL
    COMPLEX = MOD_COMPLEX

    def __init__(self):
        self.lib = load_oracle()
        self._plans = {}

    def plan(self, mod, L):
        key = (mod, int(L))
        if key not in self._plans:
            self._plans[key] = OraclePlan(self.lib, mod, L)
        return self._plans[key]

    def plan_chain(self, length):
        buf = (ctypes.c_uint * 64)()
        n = self.lib.oracle_plan_chain(int(length), buf, 64)
        return [buf[i] for i in range(n)]

    def recommend_conv_length(self, kernel_length, max_block_size):
        return int(self.lib.oracle_recommend_conv_length(kernel_length, max_block_size))
