"""CPU FP64 restatement of the reference's RHS and time-step hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline -- never as part of the product path.

The reference (a-k-sangwan/spectral-LES) is pure Python on numpy + mpi4py-fft
(FFTW) + mpi4py.  It cannot be imported here (mpi4py / mpi4py_fft are absent),
so this module restates it on one process:

* ``PFFT`` (mpi4py-fft, version unpinned, default FFTW backend, also unpinned)
  is replaced by ``numpy.fft.rfftn`` / ``irfftn`` over axes (0, 1, 2) -- or
  ``scipy.fft`` with worker threads for the all-cores CPU baseline.  The
  reference always passes ``normalize=False`` to forward and ``normalize=True``
  to backward; round-trip consistency (initial.py:13-15 with
  RK4_method.py:32-34) fixes the convention: forward unscaled, backward scaled
  by 1/N^3.
* MPI collectives on one rank are identities.

Every function cites the reference lines it restates.  Pinning (which check
fixes each function, see tests/test_oracle_*.py):

  FFT forward/backward ........ brute-force DFT at tiny N, Parseval, round trip
  wavenumbers / dealias ........ closed-form mode counts, literal constants
  U_cross_omega / del_cross_U .. Taylor-Green and ABC closed forms
  velocity_gradient ............ Taylor-Green closed-form gradient
  calculate_nuR (Smagorinsky) .. shear-flow closed form, brute-force point check
  calculate_nuR (Vreman) ....... shear flow (B_beta=0), brute-force point check
  residue_stress ............... constant-nu_R reduction to -nu_R K^2 U_hat
  computeRHS ................... K.dU = 0, TGV t=0 closed-form RHS
  RK4 / RK3 / ETD2 steps ....... ABC-flow stability polynomials
  U_square / dissipation_rate .. TGV analytic values
  find_max_velocity/compute_dt . closed form
  calculate_Ek ................. the 8 TGV modes
"""

from __future__ import annotations

import math
import os

import numpy as np

__all__ = [
    "Domain",
    "OracleFFT",
    "U_cross_omega",
    "del_cross_U",
    "velocity_gradient",
    "residue_stress",
    "smagorinsky_nuR",
    "Smagorinsky_model",
    "vreman_nuR",
    "Vreman_model",
    "computeRHS",
    "rk4_step",
    "rk3_step",
    "etd2_step_reference",
    "etd2_step_cox_matthews",
    "U_square",
    "dissipation_rate",
    "find_max_velocity",
    "compute_dt",
    "EkBins",
    "calculate_Ek",
    "initial_velocity",
    "phi1",
    "phi2",
]


# ----------------------------------------------------------------------------
# FFT object: mpi4py-fft PFFT(comm, [N]*3, axes=[0,1,2], dtype=float, grid=(-1,))
# initialize_domain.py:59-65
# ----------------------------------------------------------------------------
class OracleFFT:
    """forward = rfftn (unscaled), backward = irfftn (1/N^3 when normalize)."""

    def __init__(self, shape, workers: int | None = None):
        self.shape = tuple(int(s) for s in shape)
        self.workers = workers
        if workers is not None:
            import scipy.fft as _sfft

            self._lib = _sfft
        else:
            self._lib = None

    def _rfftn(self, x):
        if self._lib is not None:
            return self._lib.rfftn(x, axes=(0, 1, 2), workers=self.workers)
        return np.fft.rfftn(x, axes=(0, 1, 2))

    def _irfftn(self, X):
        if self._lib is not None:
            return self._lib.irfftn(X, s=self.shape, axes=(0, 1, 2), workers=self.workers)
        return np.fft.irfftn(X, s=self.shape, axes=(0, 1, 2))

    def forward(self, x, out=None, normalize=False):
        X = self._rfftn(np.asarray(x, dtype=np.float64))
        if normalize:
            X = X / float(np.prod(self.shape))
        if out is not None:
            out[...] = X
            return out
        return X

    def backward(self, X, out=None, normalize=True):
        x = self._irfftn(np.asarray(X, dtype=np.complex128))
        if not normalize:
            x = x * float(np.prod(self.shape))
        if out is not None:
            out[...] = x
            return out
        return x


# ----------------------------------------------------------------------------
# Domain: InitializeDomain, initialize_domain.py:20-186 (3D branch, one rank)
# ----------------------------------------------------------------------------
class Domain:
    """State bag restating ``InitializeDomain`` for dim=3 on one rank.

    ``No_of_gridpoint`` keeps the reference quirk N = No_of_gridpoint/4 when
    LES is on (initialize_domain.py:38-40); ``N=`` gives the grid directly.
    ``X`` / ``X_full`` coordinate grids and the HDF5 domain ``d`` are left out
    (initial-condition convenience only, initialize_domain.py:83-98,115-119).
    """

    def __init__(
        self,
        nu=0.01,
        final_time=0.0,
        dimension=3,
        time_stepSize=0.01,
        No_of_gridpoint=128,
        forcing=0.0,
        forcing_range=(0, 1),
        method="ETD2",
        LES_ON=0,
        CFL=0.5,
        Cs=0.14,
        LES_model=None,
        N=None,
        fft_workers=None,
    ):
        if dimension != 3:
            raise NotImplementedError("only the 3D branch is restated")
        self.nu = nu
        self.CFL = CFL
        self.T = final_time
        self.dt = time_stepSize
        self.LES_ON = True if LES_ON in (1, True) else False  # :38
        self.NN = 4 if self.LES_ON else 1  # :39
        self.N = int(N) if N is not None else int(No_of_gridpoint / self.NN)  # :40
        self.dx = 2 * np.pi / self.N  # :42
        self.method = method
        self.firstIteration = True  # :45
        self.forcing = forcing
        self.forcing_range = forcing_range
        self.dim = 3
        self.shape = [self.N] * 3
        self.fft = OracleFFT(self.shape, workers=fft_workers)
        N = self.N
        Nh = N // 2 + 1
        self.U = [np.zeros((N, N, N)) for _ in range(3)]  # :66
        self.U_hat = np.zeros((3, N, N, Nh), dtype=np.complex128)  # :67
        self.temp1_Img = np.zeros_like(self.U_hat)  # :68
        self.P = np.zeros((N, N, N))  # :69
        self.P_hat = np.zeros((N, N, Nh), dtype=np.complex128)  # :70
        self.dU = np.zeros_like(self.U_hat)  # :71
        self.U_hat0 = np.zeros_like(self.U_hat)  # :72
        self.U_hat1 = np.zeros_like(self.U_hat)  # :73
        self.F0 = np.zeros_like(self.U_hat) if method == "ETD2" else None  # :74
        self.F1 = np.zeros_like(self.U_hat) if method == "ETD2" else None  # :75
        self.kx = np.fft.fftfreq(N, 1.0 / N)  # :76
        self.kz = self.kx[: (N // 2 + 1)].copy()  # :77
        self.kz[-1] *= -1  # :78
        self.kmax_dealias = 2.0 / 3.0 * (N // 2 + 1)  # :79
        self.w = np.zeros((N, N, N))  # :81
        self.K = np.array(np.meshgrid(self.kx, self.kx, self.kz, indexing="ij"), dtype=int)  # :99-107
        self.curl = np.zeros((3, N, N, N))  # :108
        self.dealias = np.array(  # :109-114
            (abs(self.K[0]) < self.kmax_dealias)
            * (abs(self.K[1]) < self.kmax_dealias)
            * (abs(self.K[2]) < self.kmax_dealias),
            dtype=bool,
        )
        self.K2 = np.sum(self.K * self.K, 0, dtype=int)  # :162
        self.K_over_K2 = self.K.astype(float) / np.where(self.K2 == 0, 1, self.K2).astype(float)  # :163-165
        self.a = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]  # :166
        self.b = [0.5, 0.5, 1.0]  # :167
        if self.LES_ON:  # :168-186
            if LES_model is None:
                raise ValueError("Please choose LES_model (Smagorinsky or Vreman)")
            self.LES_model = LES_model.upper()
            self.Cs = Cs
            self.nuR = np.zeros((N, N, N))
            self.temp_I = np.zeros_like(self.U_hat)
            for name in ("a11", "a12", "a21", "a22", "a13", "a23", "a31", "a32", "a33", "aTemp"):
                setattr(self, name, np.zeros((N, N, N)))
        else:
            self.Cs = Cs
        # Cox-Matthews ETD2 (product form) keeps F(u_{n-1}) here.
        self.F_prev = None


# ----------------------------------------------------------------------------
# math_formula.py
# ----------------------------------------------------------------------------
def U_cross_omega(fft, dim, a, b, c):
    """c_i = F(u x omega)_i, unscaled.  math_formula.py:5-13 (3D :10-12)."""
    c[0] = fft.forward(a[1] * b[2] - a[2] * b[1], c[0], normalize=False)
    c[1] = fft.forward(a[2] * b[0] - a[0] * b[2], c[1], normalize=False)
    c[2] = fft.forward(a[0] * b[1] - a[1] * b[0], c[2], normalize=False)
    return c


def del_cross_U(fft, dim, a, c, K):
    """omega = F^-1(i K x U_hat), not dealiased.  math_formula.py:16-25 (3D :22-24)."""
    c[2] = fft.backward(1j * (K[0] * a[1] - K[1] * a[0]), c[2], normalize=True)
    c[1] = fft.backward(1j * (K[2] * a[0] - K[0] * a[2]), c[1], normalize=True)
    c[0] = fft.backward(1j * (K[1] * a[2] - K[2] * a[1]), c[0], normalize=True)
    return c


# a_ij = F^-1(i K_{i-1} U_hat_{j-1}) = d u_j / d x_i  (math_formula.py:33-61)
GRADIENT_NAMES = ("a11", "a12", "a13", "a21", "a22", "a23", "a31", "a32", "a33")


def velocity_gradient(domain):
    """9 inverse transforms of i K_i U_hat_j.  math_formula.py:27-62."""
    fb = domain.fft.backward
    for name in GRADIENT_NAMES:
        i, j = int(name[1]) - 1, int(name[2]) - 1
        getattr(domain, name)[:] = fb(1j * domain.K[i] * domain.U_hat[j], getattr(domain, name), normalize=True)


def residue_stress(domain):
    """T_i = 2i dealias sum_j K_j F(nu_R S_ij).  math_formula.py:64-127 (3D :73-113)."""
    d = domain
    ff = d.fft.forward
    s11 = ff(d.a11 * d.nuR, normalize=False)
    s12 = ff(0.5 * (d.a12 + d.a21) * d.nuR, normalize=False)
    s13 = ff(0.5 * (d.a13 + d.a31) * d.nuR, normalize=False)
    s22 = ff(d.a22 * d.nuR, normalize=False)
    s23 = ff(0.5 * (d.a23 + d.a32) * d.nuR, normalize=False)
    s33 = ff(d.a33 * d.nuR, normalize=False)
    f = 2.0 * 1j * d.dealias
    d.temp_I[0] = (d.K[0] * s11 + d.K[1] * s12 + d.K[2] * s13) * f
    d.temp_I[1] = (d.K[0] * s12 + d.K[1] * s22 + d.K[2] * s23) * f
    d.temp_I[2] = (d.K[0] * s13 + d.K[1] * s23 + d.K[2] * s33) * f
    return d.temp_I


# ----------------------------------------------------------------------------
# smagorinsky.py / vreman.py
# ----------------------------------------------------------------------------
def smagorinsky_nuR(domain):
    """nu_R = (Cs dx)^2 sqrt(2 S:S).  smagorinsky.py:24-47 (3D :30-39)."""
    d = domain
    d.nuR[:] = (
        d.a11 * d.a11
        + (0.5 * (d.a12 + d.a21) ** 2)
        + d.a22 * d.a22
        + (0.5 * (d.a13 + d.a31) ** 2)
        + (0.5 * (d.a23 + d.a32) ** 2)
        + d.a33 * d.a33
    )
    d.nuR[:] = np.sqrt(2.0 * d.nuR) * (d.Cs * d.dx) ** 2


def Smagorinsky_model(domain, dU):
    """smagorinsky.py:8-21."""
    velocity_gradient(domain)
    smagorinsky_nuR(domain)
    domain.temp_I[:] = residue_stress(domain)
    return domain.temp_I


def vreman_nuR(domain):
    """Vreman eddy viscosity, vreman.py:24-123 (3D :32-123).

    beta_pq = sum_m a_mp a_mq;  B = b11 b22 - b12^2 + b11 b33 - b13^2 + b22 b33 - b23^2;
    nu = sqrt(B) * [aa > 10e-5 and sqrt(B) > aa/10e6] * sqrt(1/aa) * 2.5 Cs^2 dx^2.
    sqrt of a negative B (roundoff) gives NaN, and NaN*0 stays NaN, as in the reference.
    """
    d = domain
    with np.errstate(invalid="ignore", divide="ignore"):
        b11 = d.a11 * d.a11 + d.a21 * d.a21 + d.a31 * d.a31
        b22 = d.a12 * d.a12 + d.a22 * d.a22 + d.a32 * d.a32
        b33 = d.a13 * d.a13 + d.a23 * d.a23 + d.a33 * d.a33
        b12 = d.a11 * d.a12 + d.a21 * d.a22 + d.a31 * d.a32
        b13 = d.a11 * d.a13 + d.a21 * d.a23 + d.a31 * d.a33
        b23 = d.a12 * d.a13 + d.a22 * d.a23 + d.a32 * d.a33
        d.nuR[:] = np.sqrt((b11 * b22) - (b12 * b12) + (b11 * b33) - (b13 * b13) + (b22 * b33) - (b23 * b23))
        d.aTemp[:] = (
            d.a11 * d.a11 + d.a12 * d.a12 + d.a13 * d.a13
            + d.a21 * d.a21 + d.a22 * d.a22 + d.a23 * d.a23
            + d.a31 * d.a31 + d.a32 * d.a32 + d.a33 * d.a33
        )
        d.nuR[:] *= np.where(
            (d.aTemp > 10e-5) * (d.nuR > (d.aTemp / 10e6)),
            np.sqrt(1 / d.aTemp),
            0.0,
        ) * (2.5 * d.Cs * d.Cs * d.dx * d.dx)


def Vreman_model(domain, dU):
    """vreman.py:8-21."""
    velocity_gradient(domain)
    vreman_nuR(domain)
    domain.temp_I[:] = residue_stress(domain)
    return domain.temp_I


# ----------------------------------------------------------------------------
# RK4_method.py / RK3_method.py / ETD2_method.py
# ----------------------------------------------------------------------------
def computeRHS(domain, dU, rk):
    """RK4_method.py:38-71 (RK3_method.py:44-77 is a verbatim copy)."""
    d = domain
    if rk > 0:  # :39-46
        for i in range(3):
            d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)
        d.curl[:] = del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)
    dU[:] = U_cross_omega(d.fft, 3, d.U, d.curl, dU)  # :47
    dU *= d.dealias  # :48
    if d.LES_ON:  # :49-54
        model = Smagorinsky_model if d.LES_model[0] == "S" else Vreman_model
        d.temp_I[:] = model(d, dU)
        d.P_hat[:] = np.sum((dU + d.temp_I) * d.K_over_K2, 0)
    else:  # :55-56
        d.P_hat[:] = np.sum(dU * d.K_over_K2, 0)
    dU -= d.P_hat * d.K  # :57
    if d.forcing > 0.0:  # :58-64
        d.temp1_Img[:] = d.forcing * d.U_hat[:]
        d.temp1_Img[:, 0, 0, 0] = 0.0
        dU += d.temp1_Img
    if d.method == "ETD2":  # :65-66
        d.F1[:] = dU[:]
    if d.nu > 0.0:  # :67-68
        dU -= d.nu * d.K2 * d.U_hat
    if d.LES_ON:  # :69-70
        dU[:] += d.temp_I[:]
    return dU


def rk4_step(domain):
    """Classical RK4, RK4_method.py:20-35.  Requires U_hat1 == U_hat on entry
    (SM_solver.py:33)."""
    d = domain
    d.U_hat0[:] = d.U_hat[:]
    for rk in range(4):
        d.dU[:] = computeRHS(d, d.dU, rk)
        if rk < 3:
            d.U_hat[:] = d.U_hat0 + d.b[rk] * d.dt * d.dU
        d.U_hat1[:] += d.a[rk] * d.dt * d.dU
    d.U_hat[:] = d.U_hat1[:]
    d.curl[:] = del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)
    for i in range(3):
        d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)


def rk3_step(domain):
    """RK3_method.py:20-41, as written (first order: amplification 1+z+z^2+z^3/2)."""
    d = domain
    dt = d.dt
    d.U_hat0[:] = d.U_hat[:]
    d.dU[:] = computeRHS(d, d.dU, rk=0)
    d.U_hat[:] = d.U_hat0 + 0.5 * dt * d.dU
    d.dU[:] = computeRHS(d, d.dU, rk=1)
    d.U_hat[:] = d.U_hat0 - dt * d.dU + 2.0 * dt * d.dU
    d.dU[:] = computeRHS(d, d.dU, rk=2)
    d.U_hat[:] = d.U_hat0 + dt * (1 / 6 * d.dU + 2 / 3 * d.dU + 1 / 6 * d.dU)
    d.curl[:] = del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)
    for i in range(3):
        d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)


def computeFn_for_ETD2(domain):
    """ETD2_method.py:39-46 (the forcing branch :47-55 references an undefined
    attribute and is not restated)."""
    d = domain
    d.F0[:] = d.F1[:]
    d.F1[:] = U_cross_omega(d.fft, 3, d.U, d.curl, d.F1)
    d.F1[:] *= d.dealias
    d.P_hat[:] = np.sum(d.F1 * d.K_over_K2, 0)
    d.F1[:] -= d.P_hat * d.K


def etd2_step_reference(domain):
    """ETD2_method.py:7-37 exactly as written (its F1 coefficient diverges; kept
    only to document the defect)."""
    d = domain
    if d.firstIteration:
        rk4_step(d)
        d.firstIteration = False
        return
    computeFn_for_ETD2(d)
    h, c = d.dt, d.nu * d.K2
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        d.U_hat[:] *= np.exp(-h * c)
        d.U_hat[:] += (d.F1 * ((1 - h * c) + 2 * h * c) + d.F0 * (-np.exp(-h * c) + 1 - h * c)) / np.where(
            d.K2 == 0, 1, (h * c ** 2)
        )
    for i in range(3):
        d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)
    d.curl[:] = del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)


def phi1(z):
    """phi1(z) = (e^z - 1)/z, Taylor series for |z| < 1e-2 (FP64)."""
    z = np.asarray(z, dtype=np.float64)
    out = np.empty_like(z)
    small = np.abs(z) < 1e-2
    zs = z[small]
    out[small] = 1 + zs / 2 * (1 + zs / 3 * (1 + zs / 4 * (1 + zs / 5 * (1 + zs / 6))))
    zl = z[~small]
    out[~small] = np.expm1(zl) / zl
    return out


def phi2(z):
    """phi2(z) = (e^z - 1 - z)/z^2, Taylor series for |z| < 1e-2 (FP64)."""
    z = np.asarray(z, dtype=np.float64)
    out = np.empty_like(z)
    small = np.abs(z) < 1e-2
    zs = z[small]
    out[small] = 0.5 * (1 + zs / 3 * (1 + zs / 4 * (1 + zs / 5 * (1 + zs / 6 * (1 + zs / 7)))))
    zl = z[~small]
    out[~small] = (np.expm1(zl) - zl) / (zl * zl)
    return out


def _full_nonlinear(d):
    """N(U_hat) = projected dealiased (u x omega + SGS) + forcing: the RHS of
    computeRHS (RK4_method.py:38-71) without the viscous term, evaluated at the
    current U_hat (recomputes U and curl like stage rk>0)."""
    dU = np.zeros_like(d.U_hat)
    method = d.method
    d.method = "ETD2"
    computeRHS(d, dU, rk=1)
    d.method = method
    F = d.F1.copy()
    if d.LES_ON:
        F += d.temp_I
    return F


def etd2_step_cox_matthews(domain):
    """Product ETD2: Cox & Matthews (2002) ETD2 in phi form, with the SGS term.

    U+ = e^{hL} U + h (phi1 + phi2)(hL) F_n - h phi2(hL) F_{n-1},  L = -nu K^2.
    The first call is an RK4 step (ETD2_method.py:9); F_{n-1} for the first
    ETD2 step is F evaluated at the state the RK4 bootstrap started from.
    Deviations from ETD2_method.py:16-29: the F1 coefficient is the correct
    one, and F includes the SGS stress divergence.
    """
    d = domain
    if d.firstIteration:
        d.F_prev = _full_nonlinear(d)
        d.U_hat1[:] = d.U_hat
        rk4_step(d)
        d.firstIteration = False
        return
    Fn = _full_nonlinear(d)
    h = d.dt
    z = -h * d.nu * d.K2.astype(np.float64)
    p1 = phi1(z)
    p2 = phi2(z)
    d.U_hat[:] = np.exp(z) * d.U_hat + h * (p1 + p2) * Fn - h * p2 * d.F_prev
    d.F_prev = Fn
    for i in range(3):
        d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)
    d.curl[:] = del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)


def U_square(domain):
    """mean |u|^2 summed over components / N^3.  RK4_method.py:77-84."""
    d = domain
    return float(np.sum(np.array([np.sum(d.U[i] ** 2) for i in range(3)])) / d.N ** 3)


def dissipation_rate(domain):
    """nu * sum_i sum_x u_i F^-1(K^2 U_hat_i), no 1/N^3.  RK4_method.py:87-101."""
    d = domain
    eps = 0.0
    for i in range(3):
        eps += d.nu * np.sum(d.U[i] * d.fft.backward(d.U_hat[i] * d.K2, normalize=True))
    return float(eps)


# ----------------------------------------------------------------------------
# SM_solver.py diagnostics
# ----------------------------------------------------------------------------
def find_max_velocity(U, comm=None):
    """Signed max over all components (numpy.max of the list).  SM_solver.py:245-255."""
    return float(np.max(np.array(U)))


def compute_dt(U, u_max, nu, comm=None, dx=None, CFL=0.5):
    """dt = min(CFL dx/u_max, 0.1 dx^2/nu).  SM_solver.py:258-277 (exit -> SystemExit)."""
    dt1 = CFL * dx / u_max
    dt2 = 0.1 * dx ** 2 / nu if (nu > 0) else 1e6
    dt = min(dt1, dt2)
    if dt < 1e-6:
        raise SystemExit(f"Time step (dt = {dt:g}) is smaller than 1e-6")
    return dt


class EkBins:
    """Ek_plot_variables, SM_solver.py:91-117 (one rank, no plotting)."""

    def __init__(self, domain):
        self.Ek = np.empty(domain.P_hat.shape, dtype=float)
        domain.k2 = np.unique(domain.K2)
        K2 = np.array(np.meshgrid(domain.kx, domain.kx.copy(), domain.kz, indexing="ij"), dtype=int)
        K2 = np.sum(K2 * K2, 0, dtype=int)
        k2 = np.unique(K2)
        self.avgPara = 5
        self.smoothing_para = round(np.sqrt(np.max(k2))) // self.avgPara
        self.temp_k = np.zeros((self.smoothing_para + 1), dtype=float)
        self.temp_Ek = np.zeros((self.smoothing_para + 1), dtype=float)


def calculate_Ek(domain, E):
    """Binned mean spectrum, SM_solver.py:154-192 (rank 0 branch)."""
    E.temp_k[:] = 0
    E.temp_Ek[:] = 0
    E.Ek[:] = np.sum(np.absolute(domain.U_hat[:]) ** 2, 0)
    E.Ek[:, :, 1:] *= 2
    for i in range(E.smoothing_para + 1):
        lo, hi = (i * E.avgPara) ** 2, ((i + 1) * E.avgPara) ** 2
        if np.any(domain.k2[(domain.k2 >= lo) * (domain.k2 < hi)]):
            m = (domain.K2 >= lo) * (domain.K2 < hi)
            E.temp_k[i] = np.mean(np.sqrt(domain.K2[m]))
            E.temp_Ek[i] = np.mean(E.Ek[m])


# ----------------------------------------------------------------------------
# initial.py
# ----------------------------------------------------------------------------
def initial_velocity(domain, u=0, v=0, w=0):
    """initial.py:7-18: copy, U_hat = F(U) unscaled, omega via del_cross_U."""
    d = domain
    d.U[0][:] = u
    d.U[1][:] = v
    d.U[2][:] = w
    for i in range(3):
        d.U_hat[i] = d.fft.forward(d.U[i], d.U_hat[i], normalize=False)
    d.curl[:] = del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)


def grid(N):
    """Physical coordinates X = mgrid * 2 pi / N (initialize_domain.py:83-88)."""
    return np.mgrid[:N, :N, :N].astype(float) * 2 * np.pi / N


def step(domain):
    """Method dispatch of the driver loop, SM_solver.py:52-59."""
    if domain.method == "RK4":
        rk4_step(domain)
    elif domain.method == "ETD2":
        etd2_step_cox_matthews(domain)
    elif domain.method == "ETD2_REFERENCE":
        etd2_step_reference(domain)
    elif domain.method == "RK3":
        rk3_step(domain)
    else:
        raise ValueError("Invalid method name")


def cpu_count():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
