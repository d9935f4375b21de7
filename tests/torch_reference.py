"""Differentiable FP64 restatement of the RHS + RK4 / ETD2 steps in torch (CPU),
used only by the adjoint tests as the autograd reference.

Same algebra as oracle/sles_oracle.py (computeRHS, src/RK4_method.py:38-71
with the Smagorinsky / Vreman models, src/smagorinsky.py:24-39,
src/vreman.py:24-123), written with torch.fft so that torch autograd
differentiates it w.r.t. U_hat and Cs.  Non-differentiable points
(sqrt at S:S = 0, the Vreman mask) take torch's conventions there.
"""

import math

import torch


def wavenumbers(N, dtype=torch.float64):
    k = torch.fft.fftfreq(N, 1.0 / N, dtype=dtype)
    kz = torch.arange(N // 2 + 1, dtype=dtype)
    K0, K1, K2 = torch.meshgrid(k, k, kz, indexing="ij")
    return torch.stack([K0, K1, K2])


def dealias_mask(N, K):
    kmax = 2.0 / 3.0 * (N // 2 + 1)
    return ((K[0].abs() < kmax) & (K[1].abs() < kmax) & (K[2].abs() < kmax)).to(K.dtype)


def irfftn(X, N):
    return torch.fft.irfftn(X, s=(N, N, N), dim=(-3, -2, -1))


def rfftn(x):
    return torch.fft.rfftn(x, dim=(-3, -2, -1))


def nu_smagorinsky(a, Cs, dx):
    s12, s13, s23 = a[0][1] + a[1][0], a[0][2] + a[2][0], a[1][2] + a[2][1]
    ss = a[0][0] ** 2 + 0.5 * s12 ** 2 + a[1][1] ** 2 + 0.5 * s13 ** 2 + 0.5 * s23 ** 2 + a[2][2] ** 2
    return torch.sqrt(2.0 * ss) * (Cs * dx) ** 2


def nu_vreman(a, Cs, dx):
    def beta(p, q):
        return a[0][p] * a[0][q] + a[1][p] * a[1][q] + a[2][p] * a[2][q]

    b11, b22, b33 = beta(0, 0), beta(1, 1), beta(2, 2)
    b12, b13, b23 = beta(0, 1), beta(0, 2), beta(1, 2)
    B = b11 * b22 - b12 ** 2 + b11 * b33 - b13 ** 2 + b22 * b33 - b23 ** 2
    sb = torch.sqrt(B)
    aa = b11 + b22 + b33
    mask = (aa > 1e-4) & (sb > aa / 1e7)
    f = torch.where(mask, torch.sqrt(1.0 / torch.where(mask, aa, torch.ones_like(aa))), torch.zeros_like(aa))
    return sb * f * (2.5 * Cs * Cs * dx * dx)


def rhs(Uh, Cs, N, nu, model=None, forcing=0.0, viscous=True):
    """computeRHS(U_hat) for a (3, N, N, N/2+1) complex128 tensor."""
    K = wavenumbers(N)
    D = dealias_mask(N, K)
    dx = 2 * math.pi / N
    u = irfftn(Uh, N)
    a = [[irfftn(1j * K[i] * Uh[j], N) for j in range(3)] for i in range(3)]  # a[i][j] = d u_j / d x_i
    w0 = a[1][2] - a[2][1]
    w1 = a[2][0] - a[0][2]
    w2 = a[0][1] - a[1][0]
    n0 = u[1] * w2 - u[2] * w1
    n1 = u[2] * w0 - u[0] * w2
    n2 = u[0] * w1 - u[1] * w0
    S = torch.stack([rfftn(n0), rfftn(n1), rfftn(n2)]) * D
    if model is not None:
        nu_r = nu_smagorinsky(a, Cs, dx) if model[0].upper() == "S" else nu_vreman(a, Cs, dx)
        tau = [[None] * 3 for _ in range(3)]
        for i in range(3):
            for j in range(i, 3):
                t = (a[i][i] if i == j else 0.5 * (a[i][j] + a[j][i])) * nu_r
                tau[i][j] = tau[j][i] = rfftn(t)
        T = torch.stack([sum(K[j] * tau[i][j] for j in range(3)) for i in range(3)]) * (2j * D)
        S = S + T
    K2 = (K * K).sum(0)
    den = torch.where(K2 == 0, torch.ones_like(K2), K2)
    P = (S * K).sum(0) / den
    dU = S - K * P
    if forcing > 0:
        f = torch.full_like(K2, forcing)
        f[0, 0, 0] = 0.0
        dU = dU + f * Uh
    if viscous and nu > 0:
        dU = dU - (nu * K2) * Uh
    return dU


def rk4_step(Uh, Cs, N, nu, dt, model=None, forcing=0.0):
    a = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]
    b = [0.5, 0.5, 1.0]
    U0 = Uh
    U1 = Uh
    Y = Uh
    for s in range(4):
        k = rhs(Y, Cs, N, nu, model, forcing)
        if s < 3:
            Y = U0 + b[s] * dt * k
        U1 = U1 + a[s] * dt * k
    return U1


def phi1(z):
    """phi1(z) = (e^z - 1)/z, FP64; Taylor series for |z| < 1e-2 (z = -h nu K^2 <= 0)."""
    small = z.abs() < 1e-2
    zs = torch.where(small, z, torch.zeros_like(z))
    zl = torch.where(small, -torch.ones_like(z), z)
    taylor = 1 + zs / 2 * (1 + zs / 3 * (1 + zs / 4 * (1 + zs / 5 * (1 + zs / 6))))
    return torch.where(small, taylor, torch.expm1(zl) / zl)


def phi2(z):
    """phi2(z) = (e^z - 1 - z)/z^2, FP64; Taylor series for |z| < 1e-2."""
    small = z.abs() < 1e-2
    zs = torch.where(small, z, torch.zeros_like(z))
    zl = torch.where(small, -torch.ones_like(z), z)
    taylor = 0.5 * (1 + zs / 3 * (1 + zs / 4 * (1 + zs / 5 * (1 + zs / 6 * (1 + zs / 7)))))
    return torch.where(small, taylor, (torch.expm1(zl) - zl) / (zl * zl))


def etd2_run(Uh, Cs, N, nu, dt, nsteps, model=None, forcing=0.0):
    """nsteps of the product ETD2 from Uh: the first step is an RK4 step (as
    src/ETD2_method.py:9-11), then Cox-Matthews ETD2 in phi form
        U_{n+1} = e^{hL} U_n + h (phi1 + phi2)(hL) F(U_n) - h phi2(hL) F(U_{n-1}),
    L = -nu K^2, with F the whole RHS but the viscous term (SGS stress and
    forcing included) and F(U_{-1}) := F(U_0) for the first ETD2 step.  The
    phi functions are evaluated in FP64 and hold no dependence on U or Cs."""
    K = wavenumbers(N)
    z = -dt * nu * (K * K).sum(0)
    E = torch.exp(z)
    w1 = dt * (phi1(z) + phi2(z))
    w0 = dt * phi2(z)
    F_prev = rhs(Uh, Cs, N, nu, model, forcing, viscous=False)
    U = rk4_step(Uh, Cs, N, nu, dt, model, forcing)
    for _ in range(1, nsteps):
        Fn = rhs(U, Cs, N, nu, model, forcing, viscous=False)
        U = E * U + w1 * Fn - w0 * F_prev
        F_prev = Fn
    return U


def shell_spectrum(Uh, N, kmax=40):
    """E(k) = sum over unit shells |k| in [k - 1/2, k + 1/2) of (1/2) w |U_hat|^2 / N^6,
    w = 2 for 0 < kz < N/2 (rfft doubling), k = 1..kmax."""
    K = wavenumbers(N)
    kmag = torch.sqrt((K * K).sum(0))
    w = torch.full_like(kmag, 2.0)
    w[..., 0] = 1.0
    w[..., -1] = 1.0
    e = 0.5 * w * (Uh.real ** 2 + Uh.imag ** 2).sum(0) / float(N) ** 6
    idx = torch.floor(kmag + 0.5).to(torch.int64).flatten()
    E = torch.zeros(int(idx.max()) + 1, dtype=e.dtype).index_add(0, idx, e.flatten())
    return E[1 : kmax + 1]
