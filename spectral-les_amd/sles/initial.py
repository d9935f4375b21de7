"""Initial conditions (reference src/initial.py:7-18) and the seeded
isotropic random field used by the benchmarks."""

from __future__ import annotations

import torch

from .fft import as_device_field
from .math_formula import del_cross_U

__all__ = ["initial_velocity"]


def initial_velocity(domain, u=0, v=0, w=0):
    """U = (u, v, w); U_hat = F(U) unscaled; curl = del_cross_U (initial.py:7-18).
    The field is neither dealiased nor projected, as in the reference."""
    d = domain
    for i, val in enumerate((u, v, w)):
        if torch.is_tensor(val) or hasattr(val, "shape"):
            d.U[i][...] = as_device_field(val, d.rdtype, d.device)
        else:
            d.U[i].fill_(float(val))
    d.fft.forward(d._U, d.U_hat, normalize=False)
    del_cross_U(d.fft, d.dim, d.U_hat, d.curl, None)


def isotropic_spectrum(N, k0=4.0, seed=0, urms=1.0, device="cuda", dtype=None):
    """Seeded solenoidal random velocity for decaying isotropic turbulence.

    Not in the reference (which has no random initial condition): used by the
    256^3 / 512^3 LES workloads.  Returns U_hat (3, N, N, N/2+1) on the device
    in the unscaled forward-transform convention of ``fft.forward``, with

      * E(k) proportional to k^4 exp(-2 (k/k0)^2): amplitude per mode
        |U_hat(k)| ~ k exp(-(k/k0)^2) (shell energy ~ 4 pi k^2 |U_hat|^2),
      * uniform random phases: each component exp(i theta), theta ~ U[0, 2 pi)
        from numpy's default_rng(seed), projected onto the plane normal to k
        and normalised,
      * k = 0, Nyquist rows (|kx| or |ky| or kz = N/2) and dealiased modes
        zeroed,
      * Hermitian symmetry of the kz = 0 plane enforced by averaging
        U(kx, ky, 0) with conj U(-kx, -ky, 0) (so the spectrum is the
        transform of a real field and the zeroed modes stay exactly zero),
      * rescaled to u_rms = sqrt(<|u|^2>/3) = urms (sum of squares of the
        GPU inverse transform, FP64).
    """
    import numpy as np

    from . import _lib
    from .fft import SpectralFFT, Workspace

    dtype = dtype or torch.complex128
    rdtype = torch.float64 if dtype == torch.complex128 else torch.float32
    Nh = N // 2 + 1
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(Nh, dtype=float)
    K0 = kx[:, None, None]
    K1 = kx[None, :, None]
    K2z = kz[None, None, :]
    kk2 = K0 ** 2 + K1 ** 2 + K2z ** 2
    kmag = np.sqrt(kk2)
    xi = np.exp(1j * rng.uniform(0.0, 2 * np.pi, size=(3, N, N, Nh)))
    kdot = K0 * xi[0] + K1 * xi[1] + K2z * xi[2]
    inv = np.where(kk2 > 0, 1.0 / np.where(kk2 > 0, kk2, 1.0), 0.0)
    xi[0] -= K0 * kdot * inv
    xi[1] -= K1 * kdot * inv
    xi[2] -= K2z * kdot * inv
    del kdot
    norm = np.sqrt(np.sum(np.abs(xi) ** 2, axis=0))
    amp = kmag * np.exp(-((kmag / k0) ** 2))
    kmax3 = 2 * (N // 2 + 1)
    keep = (
        (kk2 > 0)
        & (np.abs(K0) < N // 2) & (np.abs(K1) < N // 2) & (K2z < N // 2)
        & (3 * np.abs(K0) < kmax3) & (3 * np.abs(K1) < kmax3) & (3 * K2z < kmax3)
    )
    scale = np.where(keep & (norm > 0), amp / np.where(norm > 0, norm, 1.0), 0.0)
    del amp, norm, keep, kmag, kk2
    xi *= scale[None]
    del scale
    # kz = 0 plane: index -k of fftfreq order is (N - i) % N
    neg = (-np.arange(N)) % N
    plane = xi[:, :, :, 0]
    xi[:, :, :, 0] = 0.5 * (plane + np.conj(plane[:, neg][:, :, neg]))
    del plane
    U_hat = torch.as_tensor(xi, device=device).to(dtype)
    del xi
    plan = _lib.Plan(N, _lib.SLES_FP64 if rdtype == torch.float64 else _lib.SLES_FP32, device)
    fft = SpectralFFT(plan, Workspace(N, rdtype, dtype, torch.device(device)), rdtype, dtype, torch.device(device))
    u = fft.backward(U_hat, normalize=True).clone()
    from .initialize_domain import _plan_for

    rplan, rwork, rout = _plan_for(u[0])
    _lib.call("sles_sum_squares", rplan.handle, _lib.ptr_array([u[0], u[1], u[2]]), 3, rout.data_ptr(),
              rwork.data_ptr(), _lib.stream_handle(torch.device(device)))
    ms = float(rout.item()) / (3.0 * N ** 3)  # <|u|^2>/3
    del u
    _lib.call("sles_scale", plan.handle, U_hat.data_ptr(), 3, urms / np.sqrt(ms),
              _lib.stream_handle(torch.device(device)))
    torch.cuda.synchronize(device)
    return U_hat


def set_spectral_state(domain, U_hat):
    """Install a spectral velocity: U_hat, U_hat1 = U_hat, U and curl from it
    (the state initial_velocity leaves, built from U_hat instead of U)."""
    d = domain
    d.U_hat.copy_(U_hat)
    d.U_hat1.copy_(d.U_hat)
    d.fft.backward(d.U_hat, d._U, normalize=True)
    del_cross_U(d.fft, d.dim, d.U_hat, d.curl, None)
