"""Spectral resampling of the differentiable configuration's target
spectrum (a 512^3 run cut to 256^3): spectral_cut against sampling a
band-limited field on the coarse grid, and the dropped modes (CPU, numpy)."""

import numpy as np
import torch

from sles.studies import spectral_cut, spectral_pad


def _band_limited(N, kmax, seed):
    """Real random field on N^3 with modes |k_i| < kmax only."""
    rng = np.random.default_rng(seed)
    U = np.fft.rfftn(rng.standard_normal((3, N, N, N)), axes=(1, 2, 3))
    k = np.abs(np.fft.fftfreq(N, 1.0 / N))
    kz = np.arange(N // 2 + 1)
    keep = (k[:, None, None] < kmax) & (k[None, :, None] < kmax) & (kz[None, None, :] < kmax)
    return np.fft.irfftn(U * keep, s=(N, N, N), axes=(1, 2, 3))


def test_cut_equals_transform_of_the_sampled_field():
    """For a field band-limited below the coarse Nyquist, cutting the fine
    unscaled spectrum equals the unscaled transform of the field sampled on
    the coarse grid -- which fixes both the kept index set and the
    (N_to/N_from)^3 = 1/8 factor."""
    Nf, Nc = 32, 16
    u = _band_limited(Nf, Nc // 2, 1)
    Uf = torch.as_tensor(np.fft.rfftn(u, axes=(1, 2, 3)))
    got = spectral_cut(Uf, Nf, Nc).numpy()
    ref = np.fft.rfftn(u[:, ::2, ::2, ::2], axes=(1, 2, 3))
    assert got.shape == (3, Nc, Nc, Nc // 2 + 1)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_cut_drops_modes_at_and_above_the_coarse_nyquist():
    """|kx| or |ky| >= N_to/2 is dropped (the -N_to/2 row included), kz <=
    N_to/2 is kept; everything kept is the fine value times 1/8."""
    Nf, Nc = 32, 16
    rng = np.random.default_rng(2)
    U = rng.standard_normal((1, Nf, Nf, Nf // 2 + 1)) + 1j * rng.standard_normal((1, Nf, Nf, Nf // 2 + 1))
    got = spectral_cut(torch.as_tensor(U), Nf, Nc).numpy()
    h = Nc // 2
    kc = np.fft.fftfreq(Nc, 1.0 / Nc).astype(int)
    for i, kx in enumerate(kc):
        for j, ky in enumerate(kc):
            if abs(kx) >= h or abs(ky) >= h:
                assert np.all(got[0, i, j] == 0)
            else:
                assert np.array_equal(got[0, i, j], U[0, kx % Nf, ky % Nf, : h + 1] / 8.0)


def test_pad_then_cut_is_identity_and_pad_interpolates():
    """spectral_pad is the band-limited interpolation: its inverse transform
    sampled every other point is the coarse field, and cutting it back
    returns the coarse spectrum."""
    Nc, Nf = 16, 32
    u = _band_limited(Nc, Nc // 2, 3)
    Uc = torch.as_tensor(np.fft.rfftn(u, axes=(1, 2, 3)))
    Up = spectral_pad(Uc, Nc, Nf)
    uf = np.fft.irfftn(Up.numpy(), s=(Nf,) * 3, axes=(1, 2, 3))
    assert np.max(np.abs(uf[:, ::2, ::2, ::2] - u)) < 1e-12 * np.max(np.abs(u))
    back = spectral_cut(Up, Nf, Nc).numpy()
    assert np.max(np.abs(back - Uc.numpy())) < 1e-12 * np.max(np.abs(Uc.numpy()))
