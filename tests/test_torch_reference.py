"""The torch autograd reference used by the adjoint tests agrees with the
oracle (CPU, FP64)."""

import numpy as np
import pytest
import torch

import torch_reference as TR
from helpers import random_field
from oracle import sles_oracle as O


@pytest.mark.parametrize("model", [None, "Smagorinsky", "Vreman"])
def test_torch_rhs_matches_oracle(model):
    N = 16
    u = random_field(N, 3, k_cut=3.0)
    kw = dict(LES_ON=1, LES_model=model, Cs=0.17) if model else {}
    o = O.Domain(N=N, nu=0.01, forcing=0.1, method="RK4", **kw)
    O.initial_velocity(o, *u)
    O.computeRHS(o, o.dU, 1)
    Uh = torch.as_tensor(o.U_hat.copy())
    got = TR.rhs(Uh, 0.17, N, 0.01, model, forcing=0.1).numpy()
    assert np.linalg.norm(got - o.dU) / np.linalg.norm(o.dU) < 1e-12


def test_torch_rk4_matches_oracle():
    N = 12
    u = random_field(N, 4, k_cut=3.0)
    o = O.Domain(N=N, nu=0.01, method="RK4", time_stepSize=0.01, LES_ON=1, LES_model="Vreman", Cs=0.17)
    O.initial_velocity(o, *u)
    o.U_hat1[:] = o.U_hat
    Uh = torch.as_tensor(o.U_hat.copy())
    O.rk4_step(o)
    got = TR.rk4_step(Uh, 0.17, N, 0.01, 0.01, "Vreman").numpy()
    assert np.linalg.norm(got - o.U_hat) / np.linalg.norm(o.U_hat) < 1e-12


@pytest.mark.parametrize("model", [None, "Vreman"])
def test_torch_etd2_matches_oracle(model):
    """Cox-Matthews ETD2 (RK4 bootstrap, then the phi-form update) in the
    autograd reference against the oracle's product ETD2, 4 steps with
    forcing."""
    N = 12
    u = random_field(N, 5, k_cut=3.0)
    kw = dict(LES_ON=1, LES_model=model, Cs=0.25) if model else {}
    o = O.Domain(N=N, nu=0.02, method="ETD2", time_stepSize=0.01, forcing=0.05, **kw)
    O.initial_velocity(o, *u)
    Uh = torch.as_tensor(o.U_hat.copy())
    for _ in range(4):
        O.etd2_step_cox_matthews(o)
    got = TR.etd2_run(Uh, 0.25, N, 0.02, 0.01, 4, model, forcing=0.05).numpy()
    assert np.linalg.norm(got - o.U_hat) / np.linalg.norm(o.U_hat) < 1e-12


def test_torch_phi_functions():
    """phi1, phi2 on both sides of the Taylor switch against their closed
    forms evaluated in extended precision, and their limits at 0.  The direct
    form of phi2 cancels: expm1(z) - z loses about 2 eps / |z| relatively
    (4e-14 just above the switch at |z| = 1e-2), which bounds the tolerance
    there."""
    import mpmath

    z = torch.tensor([0.0, -1e-8, -5e-3, -9.99e-3, -1.001e-2, -0.3, -4.0, -60.0], dtype=torch.float64)
    p1, p2 = TR.phi1(z), TR.phi2(z)
    mpmath.mp.dps = 40
    for zi, a, b in zip(z.tolist(), p1.tolist(), p2.tolist()):
        if zi == 0.0:
            e1, e2 = 1.0, 0.5
        else:
            zm = mpmath.mpf(zi)
            e1 = float((mpmath.exp(zm) - 1) / zm)
            e2 = float((mpmath.exp(zm) - 1 - zm) / zm ** 2)
        assert abs(a - e1) <= 4e-16 * abs(e1), (zi, a, e1)
        tol2 = 4e-16 + (1e-15 / abs(zi) if abs(zi) >= 1e-2 else 0.0)
        assert abs(b - e2) <= tol2 * abs(e2), (zi, b, e2)
