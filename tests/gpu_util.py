"""Helpers for GPU tests: device buffers from numpy and one-call wrappers
around the C ABI (through the ctypes mirror in fftset_amd)."""
import numpy as np

import fftset_amd


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        return None
    return torch


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float32)


def empty(n):
    import torch
    return torch.empty(int(n), dtype=torch.float32, device="cuda")


def gpu_forward(fft, x):
    L = fft.complex_bins
    xi, out, work = dev(x), empty(2 * L), empty(2 * L)
    fft.forward(out, xi, work)
    return host(out)


def gpu_inverse(fft, x):
    L = fft.complex_bins
    xi, out, work = dev(x), empty(2 * L), empty(2 * L)
    fft.inverse(out, xi, work)
    return host(out)


def gpu_get_kernel(fft, h):
    L = fft.complex_bins
    hi, out = dev(h), empty(2 * L)
    fft.conv_get_kernel(out, hi)
    return host(out)


def gpu_conv(fft, x, h):
    L = fft.complex_bins
    xi, hi, k, out, work = dev(x), dev(h), empty(2 * L), empty(2 * L), empty(2 * L)
    fft.conv_get_kernel(k, hi)
    fft.conv(out, xi, k, work)
    return host(out)


def gpu_batch(fft, op, x, k=None):
    """Runs `op` over every row of x (n, 2L) with one fftset_fft_execute."""
    import torch
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    xi = dev(x)
    out = torch.empty_like(xi)
    ki = dev(k) if k is not None else None
    work = torch.empty_like(xi) if op == fftset_amd.OP_CONV_GRAD_KERNEL else None
    fft.execute(op, n, out, xi, kern=ki, work=work)
    return host(out)
