"""Shared helpers of the GPU parity tests: twin (GPU, oracle) domains fed
identical seeded fields."""

import numpy as np
import torch

from oracle import sles_oracle as O


def rel(a, b):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    b = b.detach().cpu().numpy() if torch.is_tensor(b) else np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def np_(t):
    return t.detach().cpu().numpy()


def random_field(N, seed, k_cut=None):
    """Seeded smooth random velocity (3, N, N, N): low-pass random modes."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, N, N, N))
    if k_cut is not None:
        Uh = np.fft.rfftn(u, axes=(1, 2, 3))
        k = np.fft.fftfreq(N, 1.0 / N)
        kz = np.arange(N // 2 + 1)
        kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
        Uh *= np.exp(-((kk / k_cut) ** 2))
        u = np.fft.irfftn(Uh, s=(N, N, N), axes=(1, 2, 3))
        u /= np.sqrt(np.mean(u * u))
    return u


def twins(N, seed=0, precision="fp64", k_cut=None, fused=False, **kw):
    """(gpu_domain, oracle_domain) with the same initial field."""
    from sles.initialize_domain import InitializeDomain
    from sles.initial import initial_velocity

    u = random_field(N, seed, k_cut=k_cut if k_cut is not None else max(2.0, N / 6))
    g = InitializeDomain(N=N, precision=precision, fused=fused, **kw)
    o = O.Domain(N=N, **kw)
    initial_velocity(g, u[0], u[1], u[2])
    O.initial_velocity(o, u[0], u[1], u[2])
    g.U_hat1.copy_(g.U_hat)
    o.U_hat1[:] = o.U_hat
    return g, o, u


def oracle_loop(o, file_out_time=0.0, No_Ek_plot=5, max_steps=None):
    """The loop of spectralMethodSolver (src/SM_solver.py:18-68) on an oracle
    domain: first dt with CFL 0.3 (:21-23), t = dt (:25), outputs at the loop
    top (:40-45), stop at t >= T or dt < 1e-8 (:46-47), t += dt (:51), step,
    dt from the CFL rule (:60-63), NaN break (:64-67).  Returns (counter1,
    events, dts) with events a list of ("snapshot" | "Ek", t, data) in the
    order the loop writes them (snapshot data: copies of U; Ek data:
    (temp_k, temp_Ek) of calculate_Ek) and dts the dt of every step taken."""
    d = o
    d.dt = O.compute_dt(d.U, O.find_max_velocity(d.U), d.nu, dx=d.dx, CFL=0.3)
    t = d.dt
    file_time = 0.0
    Ek_interval = d.T / No_Ek_plot if No_Ek_plot != 0 else 0
    Ek_counter = 0
    d.U_hat1[:] = d.U_hat
    E = O.EkBins(d) if Ek_interval > 0 else None
    events, dts = [], []
    counter1 = 0
    while True:
        if file_out_time > 0 and t >= file_time:
            events.append(("snapshot", t, np.array(d.U).copy()))
            file_time += file_out_time
        if Ek_interval > 0 and t >= Ek_counter:
            O.calculate_Ek(d, E)
            events.append(("Ek", t, (E.temp_k.copy(), E.temp_Ek.copy())))
            Ek_counter += Ek_interval
        if (t >= d.T) or (d.dt < 1e-8):
            break
        if max_steps is not None and counter1 >= max_steps:
            break
        t += d.dt
        dts.append(d.dt)
        O.step(d)
        d.dt = O.compute_dt(d.U, O.find_max_velocity(d.U), d.nu, dx=d.dx, CFL=d.CFL)
        if d.dt != d.dt:
            break
        counter1 += 1
    return counter1, events, dts


def rank1_shear_flow(N):
    """u = m sin(n . x), n = (1, 2, 3), m = (1, 0.7, -0.8) (m . n = 0):
    solenoidal, with the rank-1 gradient a_ij = n_i m_j cos(n . x) at every
    point, whose B_beta is 0 in exact arithmetic and round-off of either sign
    in floating point (m with integer components would make every product
    exact and B exactly 0)."""
    x, y, z = O.grid(N)
    s = np.sin(x + 2 * y + 3 * z)
    return s, 0.7 * s, -0.8 * s
