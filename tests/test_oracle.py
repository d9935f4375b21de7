"""Pins of the CPU oracle (oracle/sles_oracle.py) against what the mathematics
fixes independently of the oracle's own formulas: brute-force DFTs, closed
forms of the Taylor-Green and ABC flows, matrix identities for the eddy
viscosities, stability polynomials of the steppers and exact mode counts.
The reference ships no tests, golden vectors or fixtures for this path."""

import itertools
import math

import numpy as np
import pytest

from helpers import rank1_shear_flow
from oracle import sles_oracle as O


def brute_rfftn(u):
    N = u.shape[0]
    n = np.arange(N)
    X = np.zeros((N, N, N // 2 + 1), dtype=complex)
    x0, x1, x2 = np.meshgrid(n, n, n, indexing="ij")
    for a, b, c in itertools.product(range(N), range(N), range(N // 2 + 1)):
        X[a, b, c] = np.sum(u * np.exp(-2j * np.pi * (a * x0 + b * x1 + c * x2) / N))
    return X


@pytest.mark.parametrize("N", [4, 6])
def test_fft_forward_is_unscaled_dft(N):
    rng = np.random.default_rng(N)
    u = rng.standard_normal((N, N, N))
    f = O.OracleFFT((N,) * 3)
    np.testing.assert_allclose(f.forward(u), brute_rfftn(u), rtol=0, atol=1e-11)


@pytest.mark.parametrize("N", [4, 6, 8])
def test_fft_backward_inverts_and_scales(N):
    rng = np.random.default_rng(10 + N)
    u = rng.standard_normal((N, N, N))
    f = O.OracleFFT((N,) * 3)
    X = f.forward(u)
    np.testing.assert_allclose(f.backward(X), u, atol=1e-13)
    np.testing.assert_allclose(f.backward(X, normalize=False), u * N ** 3, atol=1e-10)
    # Parseval with rfft doubling (kz = 0 and N/2 planes counted once)
    w = np.full(N // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    assert np.isclose(np.sum(u * u) * N ** 3, np.sum(w * np.abs(X) ** 2), rtol=1e-12)


def test_fft_backward_drops_imag_of_kz_edge_bins():
    N = 8
    f = O.OracleFFT((N,) * 3)
    X = np.zeros((N, N, N // 2 + 1), dtype=complex)
    X[0, 0, 0] = 1j  # imaginary DC bin: ignored by c2r
    X[0, 0, N // 2] = 2j  # imaginary Nyquist bin: ignored
    assert np.max(np.abs(f.backward(X))) < 1e-15


@pytest.mark.parametrize("N", [8, 12, 16, 96, 128, 512])
def test_wavenumbers_and_dealias_count(N):
    d = O.Domain(N=N, method="RK4") if N <= 16 else None
    kmax_times3 = 2 * (N // 2 + 1)  # |k| < 2/3 (N//2+1)  <=>  3|k| < 2 (N//2+1)
    kept_axis = sum(1 for k in range(-N // 2, N // 2) if 3 * abs(k) < kmax_times3)
    kept_z = sum(1 for k in range(0, N // 2 + 1) if 3 * k < kmax_times3)
    if d is not None:
        assert int(d.dealias.sum()) == kept_axis ** 2 * kept_z
        assert d.K[0][N // 2, 0, 0] == -N // 2  # Nyquist negative on axes 0 and 1
        assert d.K[1][0, N // 2, 0] == -N // 2
        assert d.K[2][0, 0, N // 2] == N // 2  # and positive on axis 2
        assert d.K2[1, 2, 3] == 1 + 4 + 9
    # kmax keeps 171 at N = 512, 43 at 128 and 32 at 96
    expect = {512: 171, 128: 43, 96: 32}
    if N in expect:
        assert max(k for k in range(N // 2 + 1) if 3 * k < kmax_times3) == expect[N]


def tgv(N):
    x, y, z = O.grid(N)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    return x, y, z, u, v, w


def test_del_cross_U_taylor_green():
    N = 16
    d = O.Domain(N=N, method="RK4")
    x, y, z, u, v, w = tgv(N)
    O.initial_velocity(d, u, v, w)
    np.testing.assert_allclose(d.curl[0], -np.cos(x) * np.sin(y) * np.sin(z), atol=1e-13)
    np.testing.assert_allclose(d.curl[1], -np.sin(x) * np.cos(y) * np.sin(z), atol=1e-13)
    np.testing.assert_allclose(d.curl[2], 2 * np.sin(x) * np.sin(y) * np.cos(z), atol=1e-13)


def test_abc_flow_is_beltrami():
    N = 16
    d = O.Domain(N=N, method="RK4")
    x, y, z = O.grid(N)
    u, v, w = np.cos(y) + np.sin(z), np.sin(x) + np.cos(z), np.cos(x) + np.sin(y)
    O.initial_velocity(d, u, v, w)
    np.testing.assert_allclose(d.curl[0], u, atol=1e-13)
    np.testing.assert_allclose(d.curl[1], v, atol=1e-13)
    np.testing.assert_allclose(d.curl[2], w, atol=1e-13)
    c = np.zeros_like(d.U_hat)
    O.U_cross_omega(d.fft, 3, d.U, d.curl, c)
    assert np.max(np.abs(c)) < 1e-10


def test_velocity_gradient_taylor_green():
    N = 16
    d = O.Domain(N=N, LES_ON=1, LES_model="Smagorinsky", method="RK4")
    x, y, z, u, v, w = tgv(N)
    O.initial_velocity(d, u, v, w)
    O.velocity_gradient(d)
    s, c = np.sin, np.cos
    expect = {
        "a11": c(x) * c(y) * c(z), "a12": s(x) * s(y) * c(z), "a13": 0 * x,
        "a21": -s(x) * s(y) * c(z), "a22": -c(x) * c(y) * c(z), "a23": 0 * x,
        "a31": -s(x) * c(y) * s(z), "a32": c(x) * s(y) * s(z), "a33": 0 * x,
    }
    for k, e in expect.items():
        np.testing.assert_allclose(getattr(d, k), e, atol=1e-13, err_msg=k)


def _set_gradient(d, A):
    for i in range(3):
        for j in range(3):
            getattr(d, f"a{i + 1}{j + 1}")[:] = A[i][j]


def test_smagorinsky_nu_matrix_identity_and_shear():
    N = 8
    d = O.Domain(N=N, LES_ON=1, LES_model="Smagorinsky", Cs=0.17, method="RK4")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 3, N, N, N))
    _set_gradient(d, A)
    O.smagorinsky_nuR(d)
    S = 0.5 * (A + A.transpose(1, 0, 2, 3, 4))
    expect = (0.17 * d.dx) ** 2 * np.sqrt(2 * np.einsum("ij...,ij...->...", S, S))
    np.testing.assert_allclose(d.nuR, expect, rtol=1e-13)
    # shear u = (sin y, 0, 0): nu = (Cs dx)^2 |cos y|
    x, y, z = O.grid(N)
    O.initial_velocity(d, np.sin(y), 0 * y, 0 * y)
    O.velocity_gradient(d)
    O.smagorinsky_nuR(d)
    np.testing.assert_allclose(d.nuR, (0.17 * d.dx) ** 2 * np.abs(np.cos(y)), atol=1e-15)


def test_vreman_nu_matrix_identity_and_shear():
    N = 8
    Cs = 0.2
    d = O.Domain(N=N, LES_ON=1, LES_model="Vreman", Cs=Cs, method="RK4")
    rng = np.random.default_rng(4)
    A = rng.standard_normal((3, 3, N, N, N))
    _set_gradient(d, A)
    O.vreman_nuR(d)
    beta = np.einsum("mp...,mq...->pq...", A, A)  # beta_pq = sum_m a_mp a_mq
    tr = np.einsum("ii...->...", beta)
    tr2 = np.einsum("ij...,ji...->...", beta, beta)
    B = 0.5 * (tr ** 2 - tr2)  # sum of the principal 2x2 minors
    expect = np.sqrt(B / tr) * 2.5 * Cs ** 2 * d.dx ** 2
    np.testing.assert_allclose(d.nuR, expect, rtol=1e-12)
    np.testing.assert_allclose(d.aTemp, tr, rtol=1e-13)
    # shear: B_beta = 0 exactly -> nu = 0
    x, y, z = O.grid(N)
    _set_gradient(d, [[0 * y, 0 * y, 0 * y], [np.cos(y), 0 * y, 0 * y], [0 * y, 0 * y, 0 * y]])
    O.vreman_nuR(d)
    assert np.all(d.nuR == 0)
    # small gradients below the 1e-4 threshold give 0
    _set_gradient(d, 1e-4 * A / np.sqrt(np.sum(A * A, axis=(0, 1))))
    O.vreman_nuR(d)
    assert np.all(d.nuR == 0)


def random_solenoidal(d, seed):
    """Random solenoidal U_hat with no Nyquist content, Hermitian by construction."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, d.N, d.N, d.N))
    Uh = np.array([np.fft.rfftn(u[i]) for i in range(3)])
    K = d.K.astype(float)
    K2 = np.where(d.K2 == 0, 1, d.K2)
    Uh -= K * np.sum(K * Uh, 0) / K2
    N = d.N
    ny = (np.abs(d.K[0]) == N // 2) | (np.abs(d.K[1]) == N // 2) | (d.K[2] == N // 2)
    Uh[:, ny] = 0
    Uh[:, 0, 0, 0] = 0
    return Uh


def test_residue_stress_constant_nu_is_viscous_term():
    N = 12
    d = O.Domain(N=N, LES_ON=1, LES_model="Smagorinsky", method="RK4")
    d.U_hat[:] = random_solenoidal(d, 5)
    O.velocity_gradient(d)
    d.nuR[:] = 0.37
    T = O.residue_stress(d).copy()
    expect = -0.37 * d.K2 * d.U_hat * d.dealias
    np.testing.assert_allclose(T, expect, atol=1e-9 * np.max(np.abs(expect)))


def test_computeRHS_taylor_green_t0_closed_form():
    N = 16
    nu = 0.01
    d = O.Domain(N=N, nu=nu, method="RK4")
    x, y, z, u, v, w = tgv(N)
    O.initial_velocity(d, u, v, w)
    dU = np.zeros_like(d.U_hat)
    O.computeRHS(d, dU, 0)
    # d u / d t at t = 0 for Taylor-Green: -(u.grad)u - grad p, p = (cos2x + cos2y)(cos2z + 2)/16
    f = d.fft
    expect = np.array([
        f.forward(-np.sin(2 * x) * np.cos(2 * z) / 8),
        f.forward(-np.sin(2 * y) * np.cos(2 * z) / 8),
        f.forward((np.cos(2 * x) + np.cos(2 * y)) * np.sin(2 * z) / 8),
    ]) - nu * 3 * d.U_hat
    np.testing.assert_allclose(dU, expect, atol=1e-9)
    # projection leaves a divergence-free RHS
    assert np.max(np.abs(np.sum(d.K * dU, 0))) < 1e-9


def abc_domain(N, method, nu, forcing, dt):
    d = O.Domain(N=N, nu=nu, method=method, forcing=forcing, time_stepSize=dt)
    x, y, z = O.grid(N)
    O.initial_velocity(d, np.cos(y) + np.sin(z), np.sin(x) + np.cos(z), np.cos(x) + np.sin(y))
    d.U_hat1[:] = d.U_hat
    return d


@pytest.mark.parametrize("method,poly", [
    ("RK4", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24),
    ("RK3", lambda z: 1 + z + z ** 2 + z ** 3 / 2),  # the reference RK3 as written
])
def test_stepper_stability_polynomial_abc(method, poly):
    nu, f, dt = 0.05, 0.3, 0.1
    d = abc_domain(8, method, nu, f, dt)
    U0 = d.U_hat.copy()
    for _ in range(3):
        O.step(d)
    z = dt * (f - nu)  # ABC modes have K^2 = 1 and u x omega = 0
    np.testing.assert_allclose(d.U_hat, U0 * poly(z) ** 3, atol=1e-9 * np.max(np.abs(U0)))


def test_etd2_cox_matthews_exact_without_forcing_and_second_order():
    nu, dt = 0.2, 0.05
    d = abc_domain(8, "ETD2", nu, 0.0, dt)
    U0 = d.U_hat.copy()
    O.step(d)  # RK4 bootstrap
    U1 = d.U_hat.copy()
    for _ in range(4):
        O.step(d)
    # linear part integrated exactly: U_{n+1} = e^{-nu dt} U_n after the bootstrap
    np.testing.assert_allclose(d.U_hat, U1 * np.exp(-nu * dt * 4), atol=1e-12 * np.max(np.abs(U0)))

    # with forcing (a "nonlinear" term F = f U), global error is O(dt^2)
    def err(h, T=1.0):
        dd = abc_domain(8, "ETD2", nu, 0.5, h)
        V0 = dd.U_hat.copy()
        for _ in range(int(round(T / h))):
            O.step(dd)
        return np.max(np.abs(dd.U_hat - V0 * np.exp((0.5 - nu) * T))) / np.max(np.abs(V0))

    e1, e2 = err(0.1), err(0.05)
    assert 3.3 < e1 / e2 < 4.7


def test_etd2_reference_formula_diverges():
    d = abc_domain(8, "ETD2_REFERENCE", 1e-4, 0.0, 2e-3)
    d.method = "ETD2_REFERENCE"
    d.F0 = np.zeros_like(d.U_hat)
    d.F1 = np.zeros_like(d.U_hat)
    x, y, z = O.grid(8)
    # a field with a nonzero projected nonlinear term
    O.initial_velocity(d, np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), 0 * x)
    d.U_hat1[:] = d.U_hat
    d.method = "ETD2"
    O.etd2_step_reference(d)
    O.etd2_step_reference(d)
    # F1 coefficient (1 + h c)/(h c^2) at c = nu K^2 = 3e-4 is ~5.6e9
    assert np.max(np.abs(d.U_hat)) > 1e6 * np.max(np.abs(d.U_hat1))


def test_phi_functions():
    z = np.array([-1e-8, -1e-3, -0.5, -5.0, 0.0])
    p1 = O.phi1(z)
    p2 = O.phi2(z)
    for zi, a, b in zip(z, p1, p2):
        if zi == 0:
            assert a == 1 and b == 0.5
            continue
        ez = math.exp(zi)
        assert math.isclose(a, math.expm1(zi) / zi, rel_tol=1e-14)
        # series for phi2 is the reference near 0, closed form away from it
        if abs(zi) > 1e-2:
            assert math.isclose(b, (ez - 1 - zi) / zi ** 2, rel_tol=1e-12)
        else:
            assert math.isclose(b, 0.5 + zi / 6 + zi ** 2 / 24, rel_tol=1e-10)


def test_reductions_taylor_green():
    N = 16
    nu = 0.02
    d = O.Domain(N=N, nu=nu, method="RK4")
    x, y, z, u, v, w = tgv(N)
    O.initial_velocity(d, u, v, w)
    assert math.isclose(O.U_square(d), 0.25, rel_tol=1e-13)
    assert math.isclose(O.dissipation_rate(d), nu * 3 * 0.25 * N ** 3, rel_tol=1e-12)
    umax = O.find_max_velocity(d.U)
    assert math.isclose(umax, float(np.max(u)), rel_tol=0)
    dt = O.compute_dt(d.U, umax, nu, dx=d.dx, CFL=0.5)
    assert dt == min(0.5 * d.dx / umax, 0.1 * d.dx ** 2 / nu)
    with pytest.raises(SystemExit):
        O.compute_dt(d.U, 1e9, nu, dx=d.dx, CFL=0.5)


def test_calculate_Ek_taylor_green_modes():
    N = 16
    d = O.Domain(N=N, method="RK4")
    x, y, z, u, v, w = tgv(N)
    O.initial_velocity(d, u, v, w)
    E = O.EkBins(d)
    O.calculate_Ek(d, E)
    # brute-force bin bookkeeping on the half grid
    ks = [k if k < N // 2 else k - N for k in range(N)]
    k2s = [a * a + b * b + c * c for a in ks for b in ks for c in range(N // 2 + 1)]
    assert E.smoothing_para == round(math.sqrt(max(k2s))) // 5
    in0 = [k2 for k2 in k2s if k2 < 25]
    amp = N ** 3 / 8.0  # |U_hat| of each of the 8 TGV modes of u and of v
    # stored modes: kz = 1, (kx, ky) = (+-1, +-1); Ek = (|u|^2 + |v|^2) * 2
    expect0 = 4 * (2 * amp ** 2) * 2 / len(in0)
    assert math.isclose(E.temp_Ek[0], expect0, rel_tol=1e-12)
    assert math.isclose(E.temp_k[0], np.mean(np.sqrt(in0)), rel_tol=1e-13)
    assert np.all(E.temp_Ek[1:] < 1e-20 * expect0)


# ----------------------------------------------------------------------------
# The LES branch of computeRHS (RK4_method.py:49-54,69-70): the pressure
# projection acts on dU + T, and T is added once more after the viscous term.
# ----------------------------------------------------------------------------
def _les_domain(N, model, Cs=0.17, nu=0.01, forcing=0.0, seed=7):
    d = O.Domain(N=N, nu=nu, forcing=forcing, LES_ON=1, LES_model=model, Cs=Cs, method="RK4")
    d.U_hat[:] = random_solenoidal(d, seed)
    for i in range(3):
        d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)
    d.curl[:] = O.del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)
    return d


@pytest.mark.parametrize("model", ["Smagorinsky", "Vreman"])
@pytest.mark.parametrize("forcing", [0.0, 0.2])
def test_computeRHS_les_is_solenoidal(model, forcing):
    """K . dU = 0 with LES on a solenoidal field: the projection of dU + T
    (not of dU alone) makes the T added afterwards divergence-free.  T itself
    is not solenoidal, so projecting only dU would fail this."""
    d = _les_domain(12, model, forcing=forcing)
    dU = np.zeros_like(d.U_hat)
    O.computeRHS(d, dU, 1)
    scale = np.max(np.abs(dU)) * np.max(np.abs(d.K))
    assert np.max(np.abs(np.sum(d.K * dU, 0))) < 1e-12 * scale
    assert np.max(np.abs(np.sum(d.K * d.temp_I, 0))) > 1e-3 * scale  # T has a gradient part


@pytest.mark.parametrize("model", ["Smagorinsky", "Vreman"])
def test_computeRHS_les_constant_nuR_closed_form(model, monkeypatch):
    """With nu_R forced to a constant c through a hook, the SGS stress of a
    solenoidal field is -c dealias K^2 U_hat, so computeRHS with LES equals
    the no-LES RHS (pinned by the Taylor-Green closed form) minus
    c dealias K^2 U_hat."""
    c = 0.023
    hook = "smagorinsky_nuR" if model == "Smagorinsky" else "vreman_nuR"

    def const_nu(d):
        d.nuR[:] = c

    monkeypatch.setattr(O, hook, const_nu)
    d = _les_domain(12, model, forcing=0.1)
    dU = np.zeros_like(d.U_hat)
    O.computeRHS(d, dU, 1)
    plain = O.Domain(N=12, nu=d.nu, forcing=0.1, method="RK4")
    plain.U_hat[:] = d.U_hat
    ref = np.zeros_like(d.U_hat)
    O.computeRHS(plain, ref, 1)
    expect = ref - c * d.dealias * d.K2 * d.U_hat
    assert np.max(np.abs(dU - expect)) < 1e-12 * np.max(np.abs(expect))


@pytest.mark.parametrize("model", ["Smagorinsky", "Vreman"])
def test_computeRHS_cs_zero_is_no_les(model):
    """Cs = 0 gives nu_R = 0, T = 0 and the no-LES right-hand side."""
    d = _les_domain(12, model, Cs=0.0)
    dU = np.zeros_like(d.U_hat)
    O.computeRHS(d, dU, 1)
    plain = O.Domain(N=12, nu=d.nu, method="RK4")
    plain.U_hat[:] = d.U_hat
    ref = np.zeros_like(d.U_hat)
    O.computeRHS(plain, ref, 1)
    assert np.all(d.nuR == 0)
    assert np.max(np.abs(dU - ref)) <= 1e-14 * np.max(np.abs(ref))


def _rank1_negative_B_matrix():
    """A rank-1 gradient a = x y^T (B_beta = 0 exactly) for which the
    reference's difference-of-products formula (vreman.py:32-107, evaluated as
    the oracle evaluates it) rounds to B_beta < 0."""
    rng = np.random.default_rng(0)
    for _ in range(1000):
        x, y = rng.standard_normal(3), rng.standard_normal(3)
        a = np.outer(x, y)
        b = a.T @ a
        B = (b[0, 0] * b[1, 1]) - (b[0, 1] * b[0, 1]) + (b[0, 0] * b[2, 2]) - (b[0, 2] * b[0, 2]) + (
            b[1, 1] * b[2, 2]) - (b[1, 2] * b[1, 2])
        if B < 0:
            return a
    raise AssertionError("no rank-1 matrix with B < 0 found")


def test_vreman_nan_for_negative_B():
    """vreman.py:32,119-123: sqrt of a negative B_beta (round-off on a rank-1
    gradient) is NaN, the switch [aa > 1e-4 and sqrt(B) > aa/1e7] is then
    false, and NaN * 0 stays NaN: nu_R is NaN at that point and finite
    elsewhere."""
    N = 4
    d = O.Domain(N=N, LES_ON=1, LES_model="Vreman", Cs=0.17, method="RK4")
    rng = np.random.default_rng(5)
    A = rng.standard_normal((3, 3, N, N, N))
    A[:, :, 1, 2, 3] = _rank1_negative_B_matrix()
    _set_gradient(d, A)
    O.vreman_nuR(d)
    nan = np.isnan(d.nuR)
    assert nan[1, 2, 3] and nan.sum() == 1


def test_vreman_rhs_on_rank1_flow_is_nan_smagorinsky_finite():
    """On the rank-1 flow the reference's Vreman RHS is NaN (some points round
    to B < 0 and the NaN spreads through the transforms), while Smagorinsky's
    (a square root of a sum of squares) stays finite."""
    N = 16
    for model, expect_nan in (("Vreman", True), ("Smagorinsky", False)):
        d = O.Domain(N=N, nu=0.01, LES_ON=1, LES_model=model, Cs=0.17, method="RK4")
        O.initial_velocity(d, *rank1_shear_flow(N))  # helpers.rank1_shear_flow
        dU = np.zeros_like(d.U_hat)
        O.computeRHS(d, dU, 0)
        assert bool(np.isnan(d.nuR).any()) == expect_nan
        assert bool(np.isnan(dU).any()) == expect_nan
