"""The transform object of the domain (``domain.fft``).

Keeps the call surface of the reference's mpi4py-fft ``PFFT`` as it is used
there (src/initialize_domain.py:59-65; calls in src/math_formula.py,
src/RK4_method.py, src/initial.py):

    fft.forward(u, out=None, normalize=False)   -> unscaled R2C (rfftn)
    fft.backward(U, out=None, normalize=True)   -> C2R scaled by 1/N^3 (irfftn)

``out=None`` returns an internal buffer that the next call of the same kind
overwrites, as mpi4py-fft does.  Leading batch dimensions are transformed in
one launch sequence.  Inputs may be torch tensors on the device or numpy
arrays (copied to the device).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib


class Workspace:
    """Named, growable device scratch buffers viewed as real or complex fields."""

    def __init__(self, n, rdtype, cdtype, device):
        self.n = n
        self.nh = n // 2 + 1
        self.rdtype = rdtype
        self.cdtype = cdtype
        self.device = device
        self._bufs = {}

    def _raw(self, name, nbytes):
        b = self._bufs.get(name)
        if b is None or b.numel() < nbytes:
            b = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._bufs[name] = b
        return b

    def cplx(self, name, k):
        n, nh = self.n, self.nh
        esz = torch.empty((), dtype=self.cdtype).element_size()
        nbytes = k * n * n * nh * esz
        return self._raw(name, nbytes)[:nbytes].view(self.cdtype).view(k, n, n, nh)

    def real(self, name, k):
        n = self.n
        esz = torch.empty((), dtype=self.rdtype).element_size()
        nbytes = k * n * n * n * esz
        return self._raw(name, nbytes)[:nbytes].view(self.rdtype).view(k, n, n, n)

    def release(self):
        self._bufs.clear()


def as_device_field(x, dtype, device):
    """Contiguous device tensor of ``dtype`` with the values of ``x``."""
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    elif not torch.is_tensor(x):
        x = torch.as_tensor(x)
    if x.device != device or x.dtype != dtype:
        x = x.to(device=device, dtype=dtype)
    if not x.is_contiguous():
        x = x.contiguous()
    return x


def consecutive(tensors):
    """Base tensor view if the equally-shaped contiguous tensors lie back to
    back in memory (so one batched launch covers them), else None."""
    t0 = tensors[0]
    if not all(t.is_contiguous() and t.dtype == t0.dtype and t.shape == t0.shape for t in tensors):
        return None
    nbytes = t0.numel() * t0.element_size()
    for i, t in enumerate(tensors):
        if t.data_ptr() != t0.data_ptr() + i * nbytes:
            return None
    return True


class SpectralFFT:
    def __init__(self, plan, ws, rdtype, cdtype, device):
        self.plan = plan
        self.ws = ws
        self.rdtype = rdtype
        self.cdtype = cdtype
        self.device = device
        self.N = plan.n
        self.shape = (plan.n,) * 3
        self.spectral_shape = (plan.n, plan.n, plan.nh)

    def _batch(self, t, field_shape):
        nf = int(np.prod(field_shape))
        if t.numel() % nf != 0 or tuple(t.shape[-3:]) != tuple(field_shape):
            raise ValueError(f"expected trailing shape {field_shape}, got {tuple(t.shape)}")
        return t.numel() // nf

    def forward(self, input_array, output_array=None, normalize=False):
        x = as_device_field(input_array, self.rdtype, self.device)
        b = self._batch(x, self.shape)
        direct = (
            output_array is not None
            and torch.is_tensor(output_array)
            and output_array.is_contiguous()
            and output_array.dtype == self.cdtype
            and output_array.device == self.device
            and output_array.numel() == b * self.N * self.N * self.plan.nh
        )
        out = output_array if direct else self.ws.cplx("fft_forward_out", b)
        scale = 1.0 / float(self.N ** 3) if normalize else 1.0
        _lib.call("sles_fft_r2c", self.plan.handle, x.data_ptr(), out.data_ptr(), b, scale,
                  _lib.stream_handle(self.device))
        if output_array is not None:
            if not direct:
                output_array[...] = out.view(output_array.shape)
            return output_array
        return out.view(*x.shape[:-3], *self.spectral_shape)

    def backward(self, input_array, output_array=None, normalize=True):
        X = as_device_field(input_array, self.cdtype, self.device)
        b = self._batch(X, self.spectral_shape)
        direct = (
            output_array is not None
            and torch.is_tensor(output_array)
            and output_array.is_contiguous()
            and output_array.dtype == self.rdtype
            and output_array.device == self.device
            and output_array.numel() == b * self.N ** 3
        )
        out = output_array if direct else self.ws.real("fft_backward_out", b)
        work = self.ws.cplx("fft_backward_work", b)
        scale = 1.0 / float(self.N ** 3) if normalize else 1.0
        _lib.call("sles_fft_c2r", self.plan.handle, X.data_ptr(), out.data_ptr(), work.data_ptr(), b, scale,
                  _lib.stream_handle(self.device))
        if output_array is not None:
            if not direct:
                output_array[...] = out.view(output_array.shape)
            return output_array
        return out.view(*X.shape[:-3], *self.shape)

    # -- in-place variant used internally: destroys ``X`` (work = input) --
    def backward_destroy(self, X, out, batch, scale):
        _lib.call("sles_fft_c2r", self.plan.handle, X.data_ptr(), out.data_ptr(), X.data_ptr(), batch, scale,
                  _lib.stream_handle(self.device))
