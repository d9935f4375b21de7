"""The differentiable unit-shell spectrum of the adjoint driver (the
sles_shell_spectrum kernel and its vector-Jacobian product) against the
index_add formulation of the torch reference (CPU) and a brute-force loop,
values and gradients (FP64)."""

import math

import numpy as np
import pytest
import torch

import torch_reference as TR

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _field(N, seed):
    g = torch.Generator().manual_seed(seed)
    shape = (3, N, N, N // 2 + 1)
    return torch.complex(torch.randn(shape, generator=g, dtype=torch.float64),
                         torch.randn(shape, generator=g, dtype=torch.float64))


@pytest.mark.parametrize("N,kmax", [(8, 4), (12, 9), (16, 13), (16, 40)])
def test_shell_spectrum_matches_reference(N, kmax):
    from sles.adjoint import shell_spectrum

    U = _field(N, N + kmax)
    got = shell_spectrum(U.cuda(), N, kmax).cpu()
    ref = TR.shell_spectrum(U, N, kmax=kmax)
    n = min(len(got), len(ref))
    assert torch.allclose(got[:n], ref[:n], rtol=1e-13, atol=0)
    # shells beyond the largest |k| of the grid are empty
    assert torch.all(got[n:] == 0)


def test_shell_spectrum_brute_force():
    from sles.adjoint import shell_spectrum

    N, kmax = 8, 6
    U = _field(N, 3).numpy()
    E = np.zeros(kmax + 1)
    for i in range(N):
        for j in range(N):
            for l in range(N // 2 + 1):
                kx = i if i < N // 2 else i - N
                ky = j if j < N // 2 else j - N
                s = int(math.floor(math.sqrt(kx * kx + ky * ky + l * l) + 0.5))
                w = 1.0 if l in (0, N // 2) else 2.0
                if 1 <= s <= kmax:
                    E[s] += 0.5 * w * float(np.sum(np.abs(U[:, i, j, l]) ** 2)) / N ** 6
    got = shell_spectrum(torch.as_tensor(U).cuda(), N, kmax).cpu().numpy()
    assert np.allclose(got, E[1:], rtol=1e-13, atol=0)


def test_spectrum_loss_gradient_matches_reference():
    from sles.adjoint import spectrum_loss

    N, kmax = 12, 8
    U0 = _field(N, 7)
    target = TR.shell_spectrum(_field(N, 8), N, kmax=kmax)
    U = U0.cuda().requires_grad_(True)
    spectrum_loss(U, target.cuda(), N, kmax=kmax).backward()
    V = U0.clone().requires_grad_(True)
    E = TR.shell_spectrum(V, N, kmax=kmax)
    ((torch.log(E) - torch.log(target)) ** 2).sum().backward()
    assert torch.allclose(U.grad.cpu(), V.grad, rtol=1e-11, atol=1e-14)
