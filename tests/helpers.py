"""Shared helpers of the GPU parity tests: twin (GPU, oracle) domains fed
identical seeded fields."""

import numpy as np
import torch

from oracle import sles_oracle as O


def rel(a, b):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    b = b.detach().cpu().numpy() if torch.is_tensor(b) else np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def np_(t):
    return t.detach().cpu().numpy()


def random_field(N, seed, k_cut=None):
    """Seeded smooth random velocity (3, N, N, N): low-pass random modes."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, N, N, N))
    if k_cut is not None:
        Uh = np.fft.rfftn(u, axes=(1, 2, 3))
        k = np.fft.fftfreq(N, 1.0 / N)
        kz = np.arange(N // 2 + 1)
        kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
        Uh *= np.exp(-((kk / k_cut) ** 2))
        u = np.fft.irfftn(Uh, s=(N, N, N), axes=(1, 2, 3))
        u /= np.sqrt(np.mean(u * u))
    return u


def twins(N, seed=0, precision="fp64", k_cut=None, fused=False, **kw):
    """(gpu_domain, oracle_domain) with the same initial field."""
    from sles.initialize_domain import InitializeDomain
    from sles.initial import initial_velocity

    u = random_field(N, seed, k_cut=k_cut if k_cut is not None else max(2.0, N / 6))
    g = InitializeDomain(N=N, precision=precision, fused=fused, **kw)
    o = O.Domain(N=N, **kw)
    initial_velocity(g, u[0], u[1], u[2])
    O.initial_velocity(o, u[0], u[1], u[2])
    g.U_hat1.copy_(g.U_hat)
    o.U_hat1[:] = o.U_hat
    return g, o, u
