"""Driver-level behaviour: CFL rule, stop conditions and failure guards
(SM_solver.py:18-83, 245-277), the seeded isotropic initial field, and the
E(k) CSV output."""

import math
import os

import numpy as np
import pytest
import torch

from oracle import sles_oracle as O


def test_compute_dt_rule_and_exit_cpu():
    from sles.SM_solver import DtTooSmall, compute_dt

    dx = 2 * math.pi / 32
    assert compute_dt(None, 2.0, 0.01, dx=dx, CFL=0.5) == min(0.5 * dx / 2.0, 0.1 * dx ** 2 / 0.01)
    assert compute_dt(None, 2.0, 0.0, dx=dx, CFL=0.5) == 0.5 * dx / 2.0
    with pytest.raises(DtTooSmall):
        compute_dt(None, 1e9, 0.01, dx=dx, CFL=0.5)
    assert isinstance(DtTooSmall("x"), SystemExit)
    # NaN max velocity gives a NaN dt (the driver's break condition), not an exit
    assert compute_dt(None, float("nan"), 0.01, dx=dx, CFL=0.5) != compute_dt(None, float("nan"), 0.01, dx=dx,
                                                                              CFL=0.5)


def test_get_index_partition():
    from sles.utilites import get_index

    for N in (7, 64, 129):
        for n in (1, 2, 3, 8):
            parts = [get_index(N, n, r) for r in range(n)]
            assert parts[0][0] == 0 and parts[-1][1] == N
            assert all(parts[i][1] == parts[i + 1][0] for i in range(n - 1))
            sizes = [b - a for a, b in parts]
            assert max(sizes) - min(sizes) <= 1 and sizes == sorted(sizes, reverse=True)


gpu = pytest.mark.gpu


@gpu
def test_driver_loop_matches_oracle_loop(tmp_path):
    """spectralMethodSolver: dt from the CFL rule each step (first CFL 0.3)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from helpers import random_field, rel
    from sles.initial import initial_velocity
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import spectralMethodSolver

    N = 16
    u = random_field(N, 31, k_cut=3.0)
    kw = dict(nu=0.01, final_time=0.05, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17, CFL=0.5)
    g = InitializeDomain(N=N, **kw)
    initial_velocity(g, *u)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        steps = spectralMethodSolver(g, 0.0, No_Ek_plot=2)
        assert os.path.exists("Ek_plot_data.csv")
        header = open("Ek_plot_data.csv").readline().strip().split(",")
        assert header[0].startswith("k-") and header[1].startswith("Ek-")
    finally:
        os.chdir(cwd)
    # oracle loop with the same rule
    o = O.Domain(N=N, **kw)
    O.initial_velocity(o, *u)
    o.dt = O.compute_dt(o.U, O.find_max_velocity(o.U), o.nu, dx=o.dx, CFL=0.3)
    t = o.dt
    o.U_hat1[:] = o.U_hat
    n = 0
    while not (t >= o.T or o.dt < 1e-8):
        t += o.dt
        O.rk4_step(o)
        o.dt = O.compute_dt(o.U, O.find_max_velocity(o.U), o.nu, dx=o.dx, CFL=o.CFL)
        n += 1
    assert steps == n
    assert rel(g.U_hat, o.U_hat) < 1e-10


@gpu
def test_driver_nan_guard_breaks():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import spectralMethodSolver

    g = InitializeDomain(N=16, nu=0.01, final_time=1.0, method="RK4")
    g.U_hat.fill_(float("nan"))
    g._U.fill_(1.0)  # finite first dt, NaN after the first step
    assert spectralMethodSolver(g, 0.0, No_Ek_plot=0) == 0


@gpu
def test_isotropic_initial_field():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sles.initial import isotropic_spectrum

    N = 32
    Uh = isotropic_spectrum(N, k0=4.0, seed=5, urms=1.0)
    Uh2 = isotropic_spectrum(N, k0=4.0, seed=5, urms=1.0)
    assert torch.equal(Uh, Uh2)  # seeded, deterministic
    U = Uh.cpu().numpy()
    u = np.fft.irfftn(U, s=(N, N, N), axes=(1, 2, 3))
    assert math.isclose(np.sqrt(np.mean(u * u)), 1.0, rel_tol=1e-12)  # u_rms over components
    # the stored spectrum is exactly the transform of a real field
    assert np.max(np.abs(np.fft.rfftn(u, axes=(1, 2, 3)) - U)) < 1e-9 * np.max(np.abs(U))
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    K = np.array(np.meshgrid(k, k, kz, indexing="ij"))
    div = np.sum(K * U, axis=0)
    assert np.max(np.abs(div)) < 1e-9 * np.max(np.abs(U))  # solenoidal
    kmag = np.sqrt(np.sum(K * K, 0))
    assert np.all(np.abs(U[:, kmag == 0]) == 0)
    assert np.all(U[:, np.abs(K[0]) == N // 2] == 0)
    keep = np.all(np.abs(K) < 2.0 / 3.0 * (N // 2 + 1), axis=0)
    assert np.all(U[:, ~keep] == 0)
    # |U_hat|^2 per mode ~ k^2 exp(-2 (k/k0)^2) away from the kz = 0 plane
    sel = (kz[None, None, :] * np.ones_like(kmag) > 0) & keep & (kmag > 0)
    ratio = np.sum(np.abs(U) ** 2, 0)[sel] / (kmag[sel] ** 2 * np.exp(-2 * (kmag[sel] / 4.0) ** 2))
    assert np.std(ratio) / np.mean(ratio) < 1e-6
