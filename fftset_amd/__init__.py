"""fftset_amd -- MI355X (gfx950) implementation of fftset's transform and
fast-convolution path behind fftset's C API.

The product is ``libfftset_amd.so`` (HIP kernels + the C ABI declared in
``include/fftset.h``). This module is a thin ctypes mirror of that API for
Python callers and tests; buffers are device pointers, passed either as
integers or as CUDA/HIP torch tensors (float32, contiguous).

Names follow the reference interface (fftset.h): ``FFTSet.create_fft``,
``FFT.forward``, ``FFT.inverse``, ``FFT.conv_get_kernel``, ``FFT.conv``,
``recommend_conv_length``, the ``FREQ_OFFSET_REAL`` / ``COMPLEX``
modulations, plus the extensions (``execute`` for uniform batches, the
convolution adjoints).
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfftset_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "fftset.h")

OK = 0
E_INVALID = -1
E_NOMEM = -2
E_DEVICE = -3
E_UNSUPPORTED = -4
E_TOO_LARGE = -5
NOALLOC_WORK_FLOATS = 16 << 20

OP_FORWARD = 0
OP_INVERSE = 1
OP_GET_KERNEL = 2
OP_CONV = 3
OP_CONV_GRAD_INPUT = 4
OP_CONV_GRAD_KERNEL = 5
OP_CONV_SAVE_SPECTRUM = 6

_lib = None


class FFTSetError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"fftset error {code} {what}".strip())
        self.code = code


def lib_available():
    return os.path.exists(LIB_PATH)


def load_library():
    """Loads libfftset_amd.so; raises when it has not been built (there is
    no fallback implementation)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m fftset_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, fp, u, sz = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_size_t
    lib.fftset_init.argtypes = [vp]
    lib.fftset_init.restype = ctypes.c_int
    lib.fftset_destroy.argtypes = [vp]
    lib.fftset_destroy.restype = None
    lib.fftset_create_fft.argtypes = [vp, vp, u]
    lib.fftset_create_fft.restype = vp
    for name in ("fftset_fft_forward", "fftset_fft_inverse"):
        f = getattr(lib, name)
        f.argtypes = [vp, fp, fp, fp]
        f.restype = None
    lib.fftset_fft_conv_get_kernel.argtypes = [vp, fp, fp]
    lib.fftset_fft_conv_get_kernel.restype = None
    for name in ("fftset_fft_conv", "fftset_fft_conv_grad_input", "fftset_fft_conv_grad_kernel"):
        f = getattr(lib, name)
        f.argtypes = [vp, fp, fp, fp, fp]
        f.restype = None
    lib.fftset_recommend_conv_length.argtypes = [u, u]
    lib.fftset_recommend_conv_length.restype = u
    lib.fftset_last_error.argtypes = []
    lib.fftset_last_error.restype = ctypes.c_int
    lib.fftset_fft_execute.argtypes = [vp, ctypes.c_int, sz, fp, sz, fp, sz, fp, sz, fp, sz, vp]
    lib.fftset_fft_execute.restype = ctypes.c_int
    lib.fftset_fft_execute_noalloc.argtypes = [vp, ctypes.c_int, sz, fp, sz, fp, sz, fp, sz, vp]
    lib.fftset_fft_execute_noalloc.restype = ctypes.c_int
    lib.fftset_fft_complex_bins.argtypes = [vp]
    lib.fftset_fft_complex_bins.restype = u
    lib.fftset_table_bytes.argtypes = [vp]
    lib.fftset_table_bytes.restype = ctypes.c_ulonglong
    lib.fftset_fft_radices.argtypes = [vp, ctypes.POINTER(u), u]
    lib.fftset_fft_radices.restype = u
    lib.fftset_baseline_chain.argtypes = [vp, u, ctypes.POINTER(u), u]
    lib.fftset_baseline_chain.restype = u
    lib.fftset_kernel_select.argtypes = [u]
    lib.fftset_kernel_select.restype = u
    ull = ctypes.POINTER(ctypes.c_ulonglong)
    lib.fftset_kernel_counts.argtypes = [ull, ull]
    lib.fftset_kernel_counts.restype = None
    lib.fftset_batch_create.argtypes = [vp]
    lib.fftset_batch_create.restype = vp
    lib.fftset_batch_add.argtypes = [vp, vp, ctypes.c_int, fp, fp, fp]
    lib.fftset_batch_add.restype = ctypes.c_int
    lib.fftset_batch_commit.argtypes = [vp, u]
    lib.fftset_batch_commit.restype = ctypes.c_int
    lib.fftset_batch_launch.argtypes = [vp, vp]
    lib.fftset_batch_launch.restype = ctypes.c_int
    lib.fftset_batch_launch_baseline.argtypes = [vp, ctypes.c_int, vp]
    lib.fftset_batch_launch_baseline.restype = ctypes.c_int
    lib.fftset_batch_bytes.argtypes = [vp]
    lib.fftset_batch_bytes.restype = ctypes.c_longlong
    ip = ctypes.POINTER(ctypes.c_int)
    lib.fftset_batch_info.argtypes = [vp, ip, ip, ip, ip]
    lib.fftset_batch_info.restype = ctypes.c_int
    lib.fftset_batch_destroy.argtypes = [vp]
    lib.fftset_batch_destroy.restype = None
    lib.fftset_autotune.argtypes = [vp, vp, u, u, ctypes.POINTER(ctypes.c_float)]
    lib.fftset_autotune.restype = ctypes.c_int
    lib.fftset_tune_save.argtypes = [vp, ctypes.c_char_p]
    lib.fftset_tune_save.restype = ctypes.c_int
    lib.fftset_tune_load.argtypes = [vp, ctypes.c_char_p]
    lib.fftset_tune_load.restype = ctypes.c_int
    lib.fftset_blockconv_create.argtypes = [vp, fp, u, u]
    lib.fftset_blockconv_create.restype = vp
    lib.fftset_blockconv_run.argtypes = [vp, fp, sz, fp, sz, sz, sz, u, vp]
    lib.fftset_blockconv_run.restype = ctypes.c_int
    lib.fftset_blockconv_block_size.argtypes = [vp]
    lib.fftset_blockconv_block_size.restype = u
    lib.fftset_blockconv_complex_bins.argtypes = [vp]
    lib.fftset_blockconv_complex_bins.restype = u
    lib.fftset_blockconv_destroy.argtypes = [vp]
    lib.fftset_blockconv_destroy.restype = None
    _lib = lib
    return lib


def _modulation(name):
    lib = load_library()
    return ctypes.c_void_p.in_dll(lib, name).value


def _ptr(x):
    """Device address of a tensor or an int (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("buffer must be contiguous")
        if hasattr(x, "dtype") and str(x.dtype) != "torch.float32":
            raise ValueError("buffer must be float32")
        return x.data_ptr()
    raise TypeError(f"not a buffer: {type(x)}")


def _clear():
    """Drops a failure recorded by an earlier call on this thread, so the
    check after a drop-in call reports that call only."""
    load_library().fftset_last_error()


def _check(what):
    code = load_library().fftset_last_error()
    if code != OK:
        raise FFTSetError(code, what)


SELECT_GENERIC = 1
SELECT_PERSISTENT = 2
SELECT_FOUR_STEP_FUSED = 4
SELECT_FOUR_STEP_SPLIT = 8


def kernel_select(flags):
    """Sets the process-wide kernel selection (SELECT_* flags); returns the
    previous flags."""
    return int(load_library().fftset_kernel_select(flags))


def kernel_counts():
    """(specialised, generic) kernel launches since the previous call."""
    s, g = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    load_library().fftset_kernel_counts(ctypes.byref(s), ctypes.byref(g))
    return s.value, g.value


def baseline_chain(modulation, complex_bins):
    """The reference's pass structure that BASELINE_REFERENCE_PASSES runs:
    (split radices of its first pass, its planner's inner chain)."""
    buf = (ctypes.c_uint * 64)()
    n = load_library().fftset_baseline_chain(_modulation(modulation), int(complex_bins), buf, 64)
    if n == 0:
        raise FFTSetError(E_UNSUPPORTED, f"baseline_chain({modulation}, {complex_bins})")
    r = [buf[i] for i in range(n)]
    z = r.index(0)
    return r[:z], r[z + 1:]


def recommend_conv_length(kernel_length, max_block_size):
    return int(load_library().fftset_recommend_conv_length(kernel_length, max_block_size))


class FFT:
    """A plan (``const struct fftset_fft *``) owned by an :class:`FFTSet`."""

    def __init__(self, owner, handle, modulation, complex_bins):
        self._owner = owner
        self.handle = handle
        self.modulation = modulation
        self.complex_bins = complex_bins

    @property
    def L(self):
        return self.complex_bins

    def radices(self):
        buf = (ctypes.c_uint * 64)()
        n = load_library().fftset_fft_radices(self.handle, buf, 64)
        return [buf[i] for i in range(n)]

    # drop-in calls (synchronous)
    def forward(self, out, inp, work):
        _clear()
        load_library().fftset_fft_forward(self.handle, _ptr(out), _ptr(inp), _ptr(work))
        _check("forward")

    def inverse(self, out, inp, work):
        _clear()
        load_library().fftset_fft_inverse(self.handle, _ptr(out), _ptr(inp), _ptr(work))
        _check("inverse")

    def conv_get_kernel(self, out, inp):
        _clear()
        load_library().fftset_fft_conv_get_kernel(self.handle, _ptr(out), _ptr(inp))
        _check("conv_get_kernel")

    def conv(self, out, inp, kern, work):
        _clear()
        load_library().fftset_fft_conv(self.handle, _ptr(out), _ptr(inp), _ptr(kern), _ptr(work))
        _check("conv")

    def conv_grad_input(self, grad_in, grad_out, kern, work):
        _clear()
        load_library().fftset_fft_conv_grad_input(self.handle, _ptr(grad_in), _ptr(grad_out), _ptr(kern),
                                                  _ptr(work))
        _check("conv_grad_input")

    def conv_grad_kernel(self, grad_kern, grad_out, inp, work):
        _clear()
        load_library().fftset_fft_conv_grad_kernel(self.handle, _ptr(grad_kern), _ptr(grad_out), _ptr(inp),
                                                   _ptr(work))
        _check("conv_grad_kernel")

    # extension: uniform batch, stream-ordered
    def execute(self, op, count, out, inp, kern=None, work=None, out_stride=None, in_stride=None,
                kern_stride=None, work_stride=None, stream=None):
        n2 = 2 * self.complex_bins
        if out_stride is None:
            out_stride = n2
        if in_stride is None:
            in_stride = n2
        if kern_stride is None:
            kern_stride = n2 if kern is not None else 0
        if work_stride is None:
            work_stride = n2 if work is not None else 0
        if stream is None:
            try:
                import torch
                stream = torch.cuda.current_stream().cuda_stream
            except Exception:
                stream = 0
        rc = load_library().fftset_fft_execute(
            self.handle, op, count, _ptr(out), out_stride, _ptr(inp), in_stride, _ptr(kern), kern_stride,
            _ptr(work), work_stride, stream or None)
        if rc != OK:
            raise FFTSetError(rc, "execute")

    def execute_noalloc(self, op, count, out, inp, kern=None, out_stride=None, in_stride=None, kern_stride=None,
                        stream=None):
        """fftset_fft_execute_noalloc: ``execute`` without a work buffer and
        without any allocation (work space from the library's static device
        array of NOALLOC_WORK_FLOATS floats)."""
        n2 = 2 * self.complex_bins
        out_stride = n2 if out_stride is None else out_stride
        in_stride = n2 if in_stride is None else in_stride
        if kern_stride is None:
            kern_stride = n2 if kern is not None else 0
        rc = load_library().fftset_fft_execute_noalloc(
            self.handle, op, count, _ptr(out), out_stride, _ptr(inp), in_stride, _ptr(kern), kern_stride,
            _stream_handle(stream))
        if rc != OK:
            raise FFTSetError(rc, "execute_noalloc")


BATCH_NO_GRAPH = 1
BATCH_THREADS_512 = 2
BATCH_SCHED_SERIAL = 4
BATCH_SCHED_PHASED = 8
BATCH_PERSISTENT = 16
BATCH_FOUR_STEP_TILES = 32
BASELINE_PER_GROUP = 1
BASELINE_PER_TRANSFORM = 2
BASELINE_PER_PASS = 3
BASELINE_REFERENCE_PASSES = 4
BLOCKCONV_HISTORY = 1


def _stream_handle(stream):
    if stream is None:
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream or None
        except Exception:
            return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream or None
    return stream or None


class Batch:
    """``struct fftset_batch``: a ragged batch run as one graph launch."""

    def __init__(self, fs):
        self._fs = fs
        self._keep = []  # buffers must outlive the batch
        self.handle = load_library().fftset_batch_create(fs.address)
        if not self.handle:
            raise FFTSetError(load_library().fftset_last_error(), "batch_create")

    def add(self, fft, op, out, inp, kern=None):
        self._keep.extend(b for b in (out, inp, kern) if b is not None and not isinstance(b, int))
        rc = load_library().fftset_batch_add(self.handle, fft.handle, op, _ptr(out), _ptr(inp), _ptr(kern))
        if rc < 0:
            raise FFTSetError(rc, "batch_add")
        return rc

    def commit(self, flags=0):
        rc = load_library().fftset_batch_commit(self.handle, flags)
        if rc != OK:
            raise FFTSetError(rc, "batch_commit")

    def launch(self, stream=None):
        rc = load_library().fftset_batch_launch(self.handle, _stream_handle(stream))
        if rc != OK:
            raise FFTSetError(rc, "batch_launch")

    def launch_baseline(self, mode, stream=None):
        rc = load_library().fftset_batch_launch_baseline(self.handle, mode, _stream_handle(stream))
        if rc != OK:
            raise FFTSetError(rc, "batch_launch_baseline")

    @property
    def bytes(self):
        return int(load_library().fftset_batch_bytes(self.handle))

    def info(self):
        vals = [ctypes.c_int() for _ in range(4)]
        rc = load_library().fftset_batch_info(self.handle, *[ctypes.byref(v) for v in vals])
        if rc != OK:
            raise FFTSetError(rc, "batch_info")
        return dict(zip(("tiles", "classes", "lds_bytes", "four_step_classes"), (v.value for v in vals)))

    def destroy(self):
        if getattr(self, "handle", None):
            load_library().fftset_batch_destroy(self.handle)
            self.handle = None
            self._keep = []

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BlockConv:
    """``struct fftset_blockconv``: causal FIR filtering of long real signals
    by overlap-save over FREQOFFSETREAL convolutions (one launch per run)."""

    def __init__(self, fs, taps, max_block_size=0):
        self._fs = fs
        self.taps = int(taps.numel()) if hasattr(taps, "numel") else None
        self.handle = load_library().fftset_blockconv_create(fs.address, _ptr(taps), self.taps, max_block_size)
        if not self.handle:
            raise FFTSetError(load_library().fftset_last_error(), "blockconv_create")

    @property
    def block_size(self):
        return int(load_library().fftset_blockconv_block_size(self.handle))

    @property
    def complex_bins(self):
        return int(load_library().fftset_blockconv_complex_bins(self.handle))

    def run(self, out, inp, nsamples=None, channels=None, in_stride=None, out_stride=None, flags=0, stream=None):
        """Filters every row of ``inp`` (float32 CUDA tensor, rows = channels,
        or a device address with explicit sizes) into ``out``; with
        BLOCKCONV_HISTORY the k-1 samples in front of each row (``in_stride``
        apart) are the signal's past."""
        if nsamples is None:
            nsamples = out.shape[-1]
        if channels is None:
            channels = out.numel() // out.shape[-1] if out.dim() > 1 else 1
        in_stride = nsamples if in_stride is None else in_stride
        out_stride = nsamples if out_stride is None else out_stride
        rc = load_library().fftset_blockconv_run(self.handle, _ptr(out), out_stride, _ptr(inp), in_stride, nsamples,
                                                 channels, flags, _stream_handle(stream))
        if rc != OK:
            raise FFTSetError(rc, "blockconv_run")

    def destroy(self):
        if getattr(self, "handle", None):
            load_library().fftset_blockconv_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class FFTSet:
    """``struct fftset``: owns plans and the HBM twiddle tables."""

    FREQ_OFFSET_REAL = "FFTSET_MODULATION_FREQ_OFFSET_REAL"
    COMPLEX = "FFTSET_MODULATION_COMPLEX"

    def __init__(self, tune_file=None):
        lib = load_library()
        self._mem = (ctypes.c_void_p * 8)()
        if lib.fftset_init(ctypes.addressof(self._mem)) != 0:
            raise FFTSetError(E_DEVICE, "fftset_init failed")
        self._alive = True
        self._plans = {}
        if tune_file:
            self.tune_load(tune_file)

    @property
    def address(self):
        return ctypes.addressof(self._mem)

    def create_fft(self, modulation, complex_bins):
        key = (modulation, int(complex_bins))
        if key in self._plans:
            return self._plans[key]
        lib = load_library()
        h = lib.fftset_create_fft(self.address, _modulation(modulation), int(complex_bins))
        if not h:
            code = lib.fftset_last_error()
            raise FFTSetError(code, f"create_fft({modulation}, {complex_bins})")
        p = FFT(self, h, modulation, int(complex_bins))
        self._plans[key] = p
        return p

    def batch(self):
        return Batch(self)

    def blockconv(self, taps, max_block_size=0):
        """Overlap-save FIR filter with the given taps (float32 CUDA tensor)."""
        return BlockConv(self, taps, max_block_size)

    def autotune(self, modulation, complex_bins, ops=(OP_FORWARD, OP_CONV)):
        """Times candidate chains / packings / splits for a size not yet
        planned on this set on the given operations; returns the winner's
        ms per 64 MiB."""
        ms = ctypes.c_float()
        mask = 0
        for op in ops:
            mask |= 1 << op
        rc = load_library().fftset_autotune(self.address, _modulation(modulation), int(complex_bins), mask,
                                            ctypes.byref(ms))
        if rc != OK:
            raise FFTSetError(rc, f"autotune({modulation}, {complex_bins})")
        return ms.value

    @property
    def table_bytes(self):
        """HBM bytes held by the set's twiddle and modulation tables."""
        return int(load_library().fftset_table_bytes(self.address))

    def tune_save(self, path):
        rc = load_library().fftset_tune_save(self.address, os.fsencode(path))
        if rc != OK:
            raise FFTSetError(rc, "tune_save")

    def tune_load(self, path):
        rc = load_library().fftset_tune_load(self.address, os.fsencode(path))
        if rc != OK:
            raise FFTSetError(rc, "tune_load")

    def destroy(self):
        if getattr(self, "_alive", False):
            load_library().fftset_destroy(self.address)
            self._alive = False
            self._plans.clear()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


FREQ_OFFSET_REAL = FFTSet.FREQ_OFFSET_REAL
COMPLEX = FFTSet.COMPLEX
