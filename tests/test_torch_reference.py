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
