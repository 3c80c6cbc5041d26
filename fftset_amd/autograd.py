"""torch.autograd harness for fftset convolutions on the GPU.

``fft_conv(x, h, fft)`` computes ``fftset_fft_conv(x, fftset_fft_conv_get_kernel(h))``
for every row of ``x`` / ``h`` (shape ``(..., 2L)``, float32, on the GPU)
through the C ABI, and its backward calls the library's adjoint kernels:

* dLoss/dx = ``fftset_fft_conv_grad_input``(g, K): the product with the
  conjugate kernel spectrum (negacyclic correlation with h);
* dLoss/dh = the same adjoint with the signal's spectrum X in place of K
  (the spectrum product G conj(X) followed by the inverse: negacyclic
  correlation of g with x). When h needs a gradient the forward pass runs
  FFTSET_OP_CONV_SAVE_SPECTRUM, which writes X beside y, so the backward is
  one pass per gradient.

Everything runs stream-ordered on torch's current stream; there is no
PyTorch fallback -- the library must be built.
"""
import torch

import fftset_amd


def _rows(t, L):
    if t.dtype != torch.float32 or not t.is_cuda:
        raise TypeError("fft_conv expects float32 tensors on the GPU")
    if t.shape[-1] != 2 * L:
        raise ValueError(f"last dimension must be 2L = {2 * L}, got {t.shape[-1]}")
    return t.contiguous(), t.numel() // (2 * L)


class FFTConv(torch.autograd.Function):
    """y = conv(x, h) for an fftset plan; see the module docstring."""

    @staticmethod
    def forward(ctx, x, h, fft):
        L = fft.complex_bins
        x, n = _rows(x, L)
        h, nh = _rows(h, L)
        if nh not in (1, n):
            raise ValueError("h must have one row or as many rows as x")
        K = torch.empty_like(h)
        fft.execute(fftset_amd.OP_GET_KERNEL, nh, K, h)
        y = torch.empty_like(x)
        kst = 0 if nh == 1 else 2 * L
        X = None
        if ctx.needs_input_grad[1]:
            X = torch.empty_like(x)  # the signal's spectrum, for dLoss/dh
            fft.execute(fftset_amd.OP_CONV_SAVE_SPECTRUM, n, y, x, kern=K, kern_stride=kst, work=X)
        else:
            fft.execute(fftset_amd.OP_CONV, n, y, x, kern=K, kern_stride=kst)
        ctx.fft = fft
        ctx.shared_kernel = nh == 1
        ctx.kernel_shape = h.shape
        ctx.save_for_backward(X, K)
        return y

    @staticmethod
    def backward(ctx, gy):
        X, K = ctx.saved_tensors
        fft = ctx.fft
        L = fft.complex_bins
        gy, n = _rows(gy, L)
        gx = gh = None
        if ctx.needs_input_grad[0]:
            gx = torch.empty_like(gy)
            fft.execute(fftset_amd.OP_CONV_GRAD_INPUT, n, gx, gy, kern=K,
                        kern_stride=0 if ctx.shared_kernel else 2 * L)
        if ctx.needs_input_grad[1]:
            gh = torch.empty_like(gy)
            fft.execute(fftset_amd.OP_CONV_GRAD_INPUT, n, gh, gy, kern=X)
            if ctx.shared_kernel:
                gh = gh.reshape(-1, 2 * L).sum(dim=0, keepdim=True).reshape(ctx.kernel_shape)
        return gx, gh, None


def fft_conv(x, h, fft):
    """Differentiable fftset convolution of every row of x with h."""
    return FFTConv.apply(x, h, fft)
