"""Binned spectra on the GPU: calculate_Ek (src/SM_solver.py:154-192) against
the oracle, the unit-shell spectrum of the spectral loss and its
vector-Jacobian product against torch autograd of the plain restatement
(tests/torch_reference.py), identities that fix the binning independently of
either restatement, and bit-for-bit repeatability (no atomics)."""

import math

import numpy as np
import pytest
import torch

import torch_reference as TR
from helpers import twins
from oracle import sles_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _field(N, seed, dtype=torch.complex128):
    from sles.initial import isotropic_spectrum

    return isotropic_spectrum(N, k0=4.0, seed=seed, urms=1.0, dtype=dtype).contiguous()


@pytest.mark.parametrize("N", [16, 24, 64, 96])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_calculate_ek_matches_oracle_and_repeats(N, precision):
    import sles.SM_solver as S

    g, o, _ = twins(N, seed=N + 3, precision=precision, nu=0.01, method="RK4")
    Eg = S.Ek_plot_variables(g, path=None)
    Eo = O.EkBins(o)
    S.calculate_Ek(g, Eg)
    O.calculate_Ek(o, Eo)
    first = (Eg.temp_k.copy(), Eg.temp_Ek.copy(), Eg._dev.clone())
    for _ in range(3):
        S.calculate_Ek(g, Eg)
        assert torch.equal(Eg._dev, first[2])  # the same bits on every call
    # counts are exact; the bins' mean |K| to round-off
    cnt = Eg._dev[2].cpu().numpy()
    assert np.all(cnt == np.round(cnt))
    np.testing.assert_allclose(Eg.temp_k, Eo.temp_k, rtol=1e-13)
    tol = 1e-11 if precision == "fp64" else 1e-6
    np.testing.assert_allclose(Eg.temp_Ek, Eo.temp_Ek, rtol=tol, atol=1e-14 * Eo.temp_Ek.max())


def test_calculate_ek_counts_every_mode_once():
    """With bins covering the whole half spectrum, the counts add up to
    N^2 (N/2+1) and the Ek sums to sum_k Ek of every mode (a closed form of
    the field, independent of the binning)."""
    import sles.SM_solver as S
    from sles.initialize_domain import InitializeDomain
    from sles.initial import set_spectral_state

    N = 32
    d = InitializeDomain(N=N, nu=0.01, method="RK4")
    U = _field(N, 5)
    set_spectral_state(d, U)
    E = S.Ek_plot_variables(d, path=None)
    S.calculate_Ek(d, E)
    ek, ks, cnt = E._dev.cpu().numpy()
    assert cnt.sum() == N * N * (N // 2 + 1)
    Uc = U.cpu().numpy()
    w = np.full((N, N, N // 2 + 1), 2.0)
    w[..., 0] = 1.0
    total = float((w * (np.abs(Uc) ** 2).sum(0)).sum())
    assert math.isclose(ek.sum(), total, rel_tol=1e-13)
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    kmag = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    assert math.isclose(ks.sum(), float(kmag.sum()), rel_tol=1e-13)


@pytest.mark.parametrize("N,kmax", [(16, 6), (16, 14), (24, 21), (64, 40), (96, 40), (256, 40)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_shell_spectrum_matches_torch_reference(N, kmax, dtype):
    from sles.adjoint import shell_spectrum

    U = _field(N, 7 + N, dtype)
    E = shell_spectrum(U, N, kmax)
    E2 = shell_spectrum(U, N, kmax)
    assert torch.equal(E, E2)
    ref = TR.shell_spectrum(U.cpu().to(torch.complex128), N, kmax)
    rel = float((E.cpu() - ref).abs().max() / ref.abs().max())
    assert rel < 1e-13, rel


def test_shell_spectrum_total_is_parseval():
    """Sum over every shell plus the mean flow = (1/2) sum w |U|^2 / N^6 =
    (1/2) mean |u|^2 of the physical field (Parseval), whatever the shells."""
    from sles.adjoint import shell_spectrum

    N = 48
    g = torch.Generator(device="cuda").manual_seed(3)
    u = torch.randn(3, N, N, N, dtype=torch.float64, device="cuda", generator=g)
    U = torch.fft.rfftn(u, dim=(1, 2, 3)).contiguous()  # the reference's unnormalised forward transform
    kall = int(math.floor(math.sqrt(3.0) * (N // 2) + 0.5))
    E = shell_spectrum(U, N, kall)
    e_phys = 0.5 * float((u ** 2).sum(0).mean())
    e0 = 0.5 * float((U[:, 0, 0, 0].abs() ** 2).sum()) / N ** 6
    assert math.isclose(float(E.sum()) + e0, e_phys, rel_tol=1e-12)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_shell_spectrum_vjp_matches_autograd(dtype):
    from sles.adjoint import shell_spectrum

    N, kmax = 24, 15
    U = _field(N, 11, dtype).requires_grad_(True)
    c = torch.linspace(0.3, 1.7, kmax, dtype=torch.float64, device="cuda")
    (shell_spectrum(U, N, kmax) * c).sum().backward()
    Ur = U.detach().cpu().to(torch.complex128).requires_grad_(True)
    (TR.shell_spectrum(Ur, N, kmax) * c.cpu()).sum().backward()
    g = U.grad.cpu().to(torch.complex128)
    rel = float((g - Ur.grad).abs().max() / Ur.grad.abs().max())
    assert rel < (1e-13 if dtype == torch.complex128 else 1e-6), rel
    # modes outside shells 1..kmax get no gradient
    assert float(g[:, 0, 0, 0].abs().max()) == 0.0
