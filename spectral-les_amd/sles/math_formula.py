"""Spectral operators of the reference's src/math_formula.py on the GPU.

Same names, positional signatures and in-place-and-return conventions as the
reference (3D branches; the 2D branches are out of scope).  Each function is
a short sequence of libsles launches: a pointwise kernel that builds the
operands, then the batched transform kernels.
"""

from __future__ import annotations

import torch

from . import _lib
from .fft import as_device_field, consecutive

__all__ = ["U_cross_omega", "del_cross_U", "velocity_gradient", "residue_stress"]


def _fields(x, k, dtype, device):
    """List of k contiguous device fields from a list / stacked tensor."""
    if torch.is_tensor(x) and x.dim() == 4:
        items = [x[i] for i in range(k)]
    else:
        items = list(x)
    if len(items) != k:
        raise ValueError(f"expected {k} fields, got {len(items)}")
    return [as_device_field(t, dtype, device) for t in items]


def _store(c, src, k):
    """Copy k result fields ``src`` (k, ...) into ``c`` (stacked tensor or list)."""
    if torch.is_tensor(c) and c.dim() == 4:
        if c.data_ptr() != src.data_ptr():
            c[...] = src
    else:
        for i in range(k):
            if c[i].data_ptr() != src[i].data_ptr():
                c[i][...] = src[i]
    return c


def _out_view(c, k, dtype):
    """c itself if it is a contiguous stacked (k, ...) tensor of dtype, else None."""
    if torch.is_tensor(c) and c.dim() == 4 and c.shape[0] == k and c.is_contiguous() and c.dtype == dtype:
        return c
    if not torch.is_tensor(c) and len(c) == k and consecutive(list(c)) and c[0].dtype == dtype:
        t0 = c[0]
        return t0.as_strided((k, *t0.shape), (t0.numel(), *t0.stride()))
    return None


def U_cross_omega(fft, dim, a, b, c):
    """c_i = F(u x omega)_i, unscaled (math_formula.py:5-13, 3D :10-12)."""
    if dim != 3:
        raise NotImplementedError("2D branch is out of scope")
    dev, rd, cd = fft.device, fft.rdtype, fft.cdtype
    A = _fields(a, 3, rd, dev)
    B = _fields(b, 3, rd, dev)
    prod = fft.ws.real("cross_products", 3)
    s = _lib.stream_handle(dev)
    _lib.call("sles_cross_product", fft.plan.handle, _lib.ptr_array(A), _lib.ptr_array(B),
              _lib.ptr_array([prod[0], prod[1], prod[2]]), s)
    out = _out_view(c, 3, cd)
    target = out if out is not None else fft.ws.cplx("cross_hat", 3)
    _lib.call("sles_fft_r2c", fft.plan.handle, prod.data_ptr(), target.data_ptr(), 3, 1.0, s)
    if out is None:
        _store(c, target, 3)
    return c


def del_cross_U(fft, dim, a, c, K):
    """omega = F^-1(i K x U_hat), not dealiased (math_formula.py:16-25, 3D
    :22-24).  ``K`` must be the domain's wavenumber array: the kernel builds
    the same wavenumbers from indices."""
    if dim != 3:
        raise NotImplementedError("2D branch is out of scope")
    dev, rd, cd = fft.device, fft.rdtype, fft.cdtype
    U = _fields(a, 3, cd, dev)
    w = fft.ws.cplx("curl_hat", 3)
    s = _lib.stream_handle(dev)
    _lib.call("sles_spectral_curl", fft.plan.handle, _lib.ptr_array(U), _lib.ptr_array([w[0], w[1], w[2]]), s)
    out = _out_view(c, 3, rd)
    target = out if out is not None else fft.ws.real("curl_out", 3)
    fft.backward_destroy(w, target, 3, 1.0 / float(fft.N ** 3))
    if out is None:
        _store(c, target, 3)
    return c


_GRAD = ("a11", "a12", "a13", "a21", "a22", "a23", "a31", "a32", "a33")


def _grad_fields(domain):
    return [getattr(domain, n) for n in _GRAD]


def velocity_gradient(domain):
    """a_ij = F^-1(i K_{i-1} U_hat_{j-1}) = d u_j / d x_i into domain.a11..a33
    (math_formula.py:27-62)."""
    fft = domain.fft
    dev, cd = fft.device, fft.cdtype
    U = _fields(domain.U_hat, 3, cd, dev)
    g = fft.ws.cplx("grad_hat", 9)
    s = _lib.stream_handle(dev)
    _lib.call("sles_spectral_gradient", fft.plan.handle, _lib.ptr_array(U), _lib.ptr_array([g[i] for i in range(9)]), s)
    outs = _grad_fields(domain)
    scale = 1.0 / float(fft.N ** 3)
    if consecutive(outs) and outs[0].dtype == fft.rdtype:
        fft.backward_destroy(g, outs[0], 9, scale)
    else:
        tmp = fft.ws.real("grad_out", 9)
        fft.backward_destroy(g, tmp, 9, scale)
        for i, o in enumerate(outs):
            o[...] = tmp[i]


def residue_stress(domain):
    """T_i = 2i dealias sum_j K_j F(nu_R S_ij) into domain.temp_I
    (math_formula.py:64-127, 3D :73-113)."""
    fft = domain.fft
    dev = fft.device
    s = _lib.stream_handle(dev)
    tau = fft.ws.real("stress_products", 6)
    _lib.call("sles_stress_products", fft.plan.handle, _lib.ptr_array(_grad_fields(domain)), domain.nuR.data_ptr(),
              _lib.ptr_array([tau[i] for i in range(6)]), s)
    th = fft.ws.cplx("stress_hat", 6)
    _lib.call("sles_fft_r2c", fft.plan.handle, tau.data_ptr(), th.data_ptr(), 6, 1.0, s)
    T = domain.temp_I
    _lib.call("sles_stress_divergence", fft.plan.handle, _lib.ptr_array([th[i] for i in range(6)]),
              _lib.ptr_array([T[0], T[1], T[2]]), s)
    return domain.temp_I
