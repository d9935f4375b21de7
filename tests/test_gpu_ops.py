"""GPU parity of the operator-level call surface (HIP kernels through the C
ABI) against the CPU oracle on identical seeded inputs.

FP64 tolerance: rel-L2 <= 1e-12 for single operators (round-off of two
different FFT algorithms), looser where noted.  FP32: 1e-5."""

import math

import numpy as np
import pytest
import torch

from helpers import np_, rel, twins
from oracle import sles_oracle as O

pytestmark = pytest.mark.gpu

SIZES = [8, 12, 16, 24, 32, 48, 64, 96]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("N", SIZES + [128])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fft_forward_backward_vs_numpy(N, precision):
    from sles.initialize_domain import InitializeDomain

    d = InitializeDomain(N=N, precision=precision, method="RK4")
    rng = np.random.default_rng(N)
    u = rng.standard_normal((2, N, N, N))
    tol = 1e-13 if precision == "fp64" else 3e-6
    X = d.fft.forward(torch.as_tensor(u, device="cuda"))
    ref = np.fft.rfftn(u, axes=(1, 2, 3))
    assert rel(X, ref) < tol
    # non-Hermitian spectral input: c2r semantics must match numpy.irfftn
    Z = rng.standard_normal((2, N, N, N // 2 + 1)) + 1j * rng.standard_normal((2, N, N, N // 2 + 1))
    x = d.fft.backward(torch.as_tensor(Z, device="cuda"))
    refx = np.fft.irfftn(Z, s=(N, N, N), axes=(1, 2, 3))
    assert rel(x, refx) < tol
    x2 = d.fft.backward(torch.as_tensor(Z[0], device="cuda"), normalize=False)
    assert rel(x2, refx[0] * N ** 3) < tol


@pytest.mark.parametrize("N", [192, 384])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fft_12_point_plans_vs_numpy(N, precision):
    """Grid sizes with 12 points per thread (radix 4, 4, 4, 3 [, 2] lines)."""
    from sles.initialize_domain import InitializeDomain

    d = InitializeDomain(N=N, precision=precision, method="RK4")
    rng = np.random.default_rng(N)
    u = rng.standard_normal((N, N, N))
    tol = 1e-13 if precision == "fp64" else 3e-6
    X = d.fft.forward(torch.as_tensor(u, device="cuda"))
    assert rel(X, np.fft.rfftn(u)) < tol
    Z = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    x = d.fft.backward(torch.as_tensor(Z, device="cuda"))
    assert rel(x, np.fft.irfftn(Z, s=(N, N, N))) < tol


@pytest.mark.parametrize(
    "N,precision,wide",
    [(256, "fp64", 3), (256, "fp32", 3), (256, "fp64", 1), (256, "fp32", 2), (512, "fp64", 3), (512, "fp32", 3)],
)
def test_fft_wide_tiles_vs_numpy(N, precision, wide):
    """Wide-tile strided passes (SLES_TUNE_FFT_WIDE; bit 0 x pass, bit 1 y
    pass) against numpy and against the default tiles on the same input.
    The domain is untuned, so the plan holds its built-in default (the
    autotuner's table may choose another width per model)."""
    from sles.initialize_domain import InitializeDomain

    d = InitializeDomain(N=N, precision=precision, method="RK4", tuned=False)
    B = 2 if N <= 256 else 1
    rng = np.random.default_rng(N + wide)
    u = rng.standard_normal((B, N, N, N))
    Z = (rng.standard_normal((B, N, N, N // 2 + 1)) + 1j * rng.standard_normal((B, N, N, N // 2 + 1))) * 1e-3
    ut, Zt = torch.as_tensor(u, device="cuda"), torch.as_tensor(Z, device="cuda")
    tol = 1e-13 if precision == "fp64" else 3e-6
    default = d.plan.get_tuning()["fft_wide"]
    assert default == (3 if N == 512 else 1 if precision == "fp64" else 2)
    d.plan.set_tuning(fft_wide=0)
    X0 = d.fft.forward(ut).clone()
    x0 = d.fft.backward(Zt).clone()
    d.plan.set_tuning(fft_wide=wide)
    try:
        assert d.plan.get_tuning()["fft_wide"] == wide
        X = d.fft.forward(ut)
        ref = np.fft.rfftn(u, axes=(1, 2, 3))
        assert rel(X, ref) < tol
        assert rel(X, X0) < tol
        del ref
        x = d.fft.backward(Zt)
        refx = np.fft.irfftn(Z, s=(N, N, N), axes=(1, 2, 3))
        assert rel(x, refx) < tol
        assert rel(x, x0) < tol
    finally:
        d.plan.set_tuning(fft_wide=default)


@pytest.mark.parametrize("N", [16, 256])
def test_fft_large_roundtrip(N):
    from sles.initialize_domain import InitializeDomain

    d = InitializeDomain(N=N, method="RK4")
    g = torch.Generator(device="cuda").manual_seed(1)
    u = torch.randn(N, N, N, device="cuda", dtype=torch.float64, generator=g)
    X = d.fft.forward(u)
    back = d.fft.backward(X)
    assert rel(back, u) < 1e-14
    # spot check a few bins against a direct sum
    un = np_(u)
    n = np.arange(N)
    Xn = np_(X)
    for (a, b, c) in [(0, 0, 0), (1, 2, 3), (N - 1, N // 2, N // 2), (N // 2, 5, 1)]:
        e0 = np.exp(-2j * np.pi * a * n / N)[:, None, None]
        e1 = np.exp(-2j * np.pi * b * n / N)[None, :, None]
        e2 = np.exp(-2j * np.pi * c * n / N)[None, None, :]
        direct = np.sum(un * e0 * e1 * e2)
        assert abs(Xn[a, b, c] - direct) < 1e-9 * math.sqrt(N ** 3)


@pytest.mark.parametrize("N", [8, 16, 24, 32])
def test_math_formula_ops(N):
    import sles.math_formula as M

    g, o, u = twins(N, seed=N, LES_ON=1, LES_model="Smagorinsky", method="RK4")
    # del_cross_U from initial_velocity
    assert rel(g.curl, o.curl) < 1e-12
    assert rel(g.U_hat, o.U_hat) < 1e-13
    c = torch.zeros_like(g.U_hat)
    M.U_cross_omega(g.fft, 3, g.U, g.curl, c)
    co = np.zeros_like(o.U_hat)
    O.U_cross_omega(o.fft, 3, o.U, o.curl, co)
    assert rel(c, co) < 1e-12
    M.velocity_gradient(g)
    O.velocity_gradient(o)
    for name in ("a11", "a12", "a13", "a21", "a22", "a23", "a31", "a32", "a33"):
        assert rel(getattr(g, name), getattr(o, name)) < 1e-12, name


@pytest.mark.parametrize("model", ["Smagorinsky", "Vreman"])
@pytest.mark.parametrize("N", [12, 32])
def test_sgs_models(model, N):
    import sles.smagorinsky as S
    import sles.vreman as V

    g, o, u = twins(N, seed=3, LES_ON=1, LES_model=model, Cs=0.17, method="RK4")
    mod_g = S.Smagorinsky_model if model[0] == "S" else V.Vreman_model
    mod_o = O.Smagorinsky_model if model[0] == "S" else O.Vreman_model
    Tg = mod_g(g, g.dU)
    To = mod_o(o, o.dU)
    assert rel(g.nuR, o.nuR) < 1e-12
    assert rel(Tg, To) < 1e-12


@pytest.mark.parametrize("les", [None, "Smagorinsky", "Vreman"])
@pytest.mark.parametrize("forcing", [0.0, 0.1])
def test_computeRHS(les, forcing):
    import sles.RK4_method as R

    N = 24
    kw = dict(LES_ON=1, LES_model=les, Cs=0.17) if les else {}
    g, o, u = twins(N, seed=7, nu=0.01, forcing=forcing, method="ETD2", **kw)
    for rk in (0, 1):
        R.computeRHS(g, g.dU, rk)
        O.computeRHS(o, o.dU, rk)
        assert rel(g.dU, o.dU) < 1e-12
        assert rel(g.F1, o.F1) < 1e-12
        assert rel(g.P_hat, o.P_hat) < 1e-12


@pytest.mark.parametrize("method", ["RK4", "RK3"])
@pytest.mark.parametrize("les", [None, "Smagorinsky", "Vreman"])
def test_steps_operator_path(method, les):
    import sles.RK3_method as R3
    import sles.RK4_method as R4

    N = 16
    kw = dict(LES_ON=1, LES_model=les, Cs=0.17) if les else {}
    g, o, u = twins(N, seed=11, nu=0.005, time_stepSize=0.01, method=method, **kw)
    step_g = R4.solveForNextTimestep if method == "RK4" else R3.solveForNextTimestep
    step_o = O.rk4_step if method == "RK4" else O.rk3_step
    for _ in range(5):
        step_g(g)
        step_o(o)
    assert rel(g.U_hat, o.U_hat) < 1e-11
    assert rel(torch.stack(g.U), np.array(o.U)) < 1e-11
    assert rel(g.curl, o.curl) < 1e-11
    if method == "RK4":
        assert rel(g.U_hat1, o.U_hat1) < 1e-11


@pytest.mark.parametrize("form", ["cox_matthews", "reference"])
def test_etd2(form):
    import sles.ETD2_method as E

    N = 16
    from sles.initialize_domain import InitializeDomain
    from sles.initial import initial_velocity
    from helpers import random_field

    u = random_field(N, 12, k_cut=3.0)
    g = InitializeDomain(N=N, nu=0.01, time_stepSize=0.005, method="ETD2", LES_ON=1, LES_model="Vreman",
                         Cs=0.17, etd2_form=form, fused=False)
    o = O.Domain(N=N, nu=0.01, time_stepSize=0.005, method="ETD2", LES_ON=1, LES_model="Vreman", Cs=0.17)
    initial_velocity(g, *u)
    O.initial_velocity(o, *u)
    g.U_hat1.copy_(g.U_hat)
    o.U_hat1[:] = o.U_hat
    step_o = O.etd2_step_cox_matthews if form == "cox_matthews" else O.etd2_step_reference
    nsteps = 4 if form == "cox_matthews" else 2
    for _ in range(nsteps):
        E.solveForNextTimestep(g)
        step_o(o)
    assert rel(g.U_hat, o.U_hat) < 1e-11


def test_reductions_and_spectrum():
    import sles.RK4_method as R
    import sles.SM_solver as S

    N = 32
    g, o, u = twins(N, seed=21, nu=0.01, method="RK4")
    assert math.isclose(R.U_square(g), O.U_square(o), rel_tol=1e-13)
    assert math.isclose(R.dissipation_rate(g), O.dissipation_rate(o), rel_tol=1e-11)
    assert S.find_max_velocity(g.U) == O.find_max_velocity(o.U)
    Eg = S.Ek_plot_variables(g)
    Eo = O.EkBins(o)
    assert Eg.smoothing_para == Eo.smoothing_para
    S.calculate_Ek(g, Eg)
    O.calculate_Ek(o, Eo)
    np.testing.assert_allclose(Eg.temp_k, Eo.temp_k, rtol=1e-13)
    # bins far down the spectrum sit at round-off of the total: compare against the peak
    np.testing.assert_allclose(Eg.temp_Ek, Eo.temp_Ek, rtol=1e-11, atol=1e-14 * Eo.temp_Ek.max())


def test_fp32_ops_close_to_fp64():
    import sles.RK4_method as R

    N = 32
    g, o, u = twins(N, seed=5, precision="fp32", nu=0.01, method="RK4", LES_ON=1, LES_model="Smagorinsky",
                    Cs=0.17)
    R.computeRHS(g, g.dU, 1)
    O.computeRHS(o, o.dU, 1)
    assert rel(g.dU, o.dU) < 2e-5
