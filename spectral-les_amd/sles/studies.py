"""Measurement drivers beyond the single-step benchmark.

precision_study  FP32 form vs the FP64 form on the same seeded field and the
                 same dt sequence: velocity rel-L2, max relative E(k) error,
                 steps/s of each form.
ensemble_run     many members of different grid sizes advanced together, one
                 captured HIP graph per ensemble step (see ensemble.py).

    python -m sles.studies precision --n 256 --steps 100 500 --out result.json
"""

from __future__ import annotations

import argparse
import json
import math
import time

import torch

from . import fused
from .initial import isotropic_spectrum, set_spectral_state
from .initialize_domain import InitializeDomain


def shell_energy(U_hat, N):
    """E(k) = sum over |k| in [k-1/2, k+1/2) of (1/2) w |U_hat|^2 / N^6, FP64."""
    dev = U_hat.device
    k = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=dev)
    kz = torch.arange(N // 2 + 1, dtype=torch.float64, device=dev)
    kmag = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    w = torch.full_like(kmag, 2.0)
    w[..., 0] = 1.0
    w[..., -1] = 1.0
    e = 0.5 * w * (U_hat.real.double() ** 2 + U_hat.imag.double() ** 2).sum(0) / float(N) ** 6
    idx = torch.floor(kmag + 0.5).to(torch.int64).flatten()
    return torch.zeros(int(idx.max()) + 1, dtype=torch.float64, device=dev).index_add(0, idx, e.flatten())


def _rel_l2(a, b):
    return float((a.double() - b.double()).norm() / b.double().norm())


def precision_study(N=256, checkpoints=(100,), seed=1, les="Smagorinsky", Cs=0.17, nu=1e-4, cfl=0.5):
    """Advance FP64 and FP32 forms side by side with the FP64 run's CFL dt."""
    kw = dict(nu=nu, method="RK4", LES_ON=1, LES_model=les, Cs=Cs, CFL=cfl)
    d64 = InitializeDomain(N=N, precision="fp64", **kw)
    d32 = InitializeDomain(N=N, precision="fp32", **kw)
    U0 = isotropic_spectrum(N, k0=4.0, seed=seed, urms=1.0)
    set_spectral_state(d64, U0)
    set_spectral_state(d32, U0.to(torch.complex64))
    del U0
    from .SM_solver import compute_dt, find_max_velocity

    dt = compute_dt(d64.U, find_max_velocity(d64.U), nu, dx=d64.dx, CFL=0.3)
    kmax = int(2.0 / 3.0 * (N // 2 + 1))
    out = {"N": N, "les_model": les, "Cs": Cs, "nu": nu, "seed": seed, "results": []}
    t64 = t32 = 0.0
    steps = max(checkpoints)
    for n in range(1, steps + 1):
        torch.cuda.synchronize()
        a = time.perf_counter()
        fused.rk4_step(d64, dt=dt)
        torch.cuda.synchronize()
        b = time.perf_counter()
        fused.rk4_step(d32, dt=dt)
        torch.cuda.synchronize()
        c = time.perf_counter()
        t64 += b - a
        t32 += c - b
        dt = compute_dt(d64.U, find_max_velocity(d64.U), nu, dx=d64.dx, CFL=cfl)
        if n in checkpoints:
            e64 = shell_energy(d64.U_hat, N)[1 : kmax + 1]
            e32 = shell_energy(d32.U_hat, N)[1 : kmax + 1]
            ek_err = float(((e32 - e64).abs() / e64).max())
            rec = {
                "steps": n,
                "velocity_rel_l2": _rel_l2(d32._U, d64._U),
                "u_hat_rel_l2": _rel_l2(d32.U_hat, d64.U_hat),
                "max_rel_Ek_error_k_le_kmax": ek_err,
                "kmax": kmax,
                "fp64_ms_per_step": 1e3 * t64 / n,
                "fp32_ms_per_step": 1e3 * t32 / n,
                "fp32_speedup": t64 / t32,
            }
            out["results"].append(rec)
            print(json.dumps(rec), flush=True)
    return out


def spectral_cut(U_hat, N_from, N_to):
    """Keep |kx|, |ky| < N_to/2 and kz <= N_to/2 of an N_from^3 spectrum, scaled
    by (N_to/N_from)^3 for the unscaled forward-transform convention."""
    h = N_to // 2
    sel = torch.cat([torch.arange(0, h), torch.arange(N_from - h, N_from)]).to(U_hat.device)
    out = U_hat[:, sel][:, :, sel][..., : h + 1].clone()
    out[:, h] = 0  # kx = -N_to/2 (|kx| < N_to/2 kept)
    out[:, :, h] = 0
    return out * (float(N_to) / float(N_from)) ** 3


def differentiable_study(N=256, N_target=512, steps=50, target_steps=50, dt=2e-3, nu=1e-4, c_vreman=0.08,
                         checkpoint_every=10, seed=2, fd_h=1e-3):
    """Vreman LES, FP32, ETD2, dt = 2e-3 over a `steps` window; loss = squared
    log-mismatch of E(k), k <= 40, against a Smagorinsky N_target^3 run cut to N^3."""
    from .adjoint import DifferentiableLES, spectrum_loss

    # target spectrum
    t = InitializeDomain(N=N_target, nu=nu, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17,
                         precision="fp32")
    set_spectral_state(t, isotropic_spectrum(N_target, seed=seed + 100, dtype=t.cdtype))
    for _ in range(target_steps):
        fused.rk4_step(t, dt=dt, physical=False)
    Ucut = spectral_cut(t.U_hat.to(torch.complex128), N_target, N)
    del t
    torch.cuda.empty_cache()
    target = shell_energy(Ucut, N)[1:41].clone()
    del Ucut
    solver = DifferentiableLES(N=N, nu=nu, dt=dt, nsteps=steps, les_model="Vreman", scheme="ETD2",
                               precision="fp32", checkpoint_every=checkpoint_every)
    # Vreman coefficient c = 2.5 Cs^2  ->  Cs = sqrt(c / 2.5)
    cs0 = math.sqrt(c_vreman / 2.5)
    U0 = isotropic_spectrum(N, seed=seed, dtype=torch.complex64)
    U = U0.clone().requires_grad_(True)
    c = torch.tensor(cs0, dtype=torch.float64, device="cuda", requires_grad=True)
    torch.cuda.synchronize()
    a = time.perf_counter()
    out = solver(U, c)
    torch.cuda.synchronize()
    b = time.perf_counter()
    loss = spectrum_loss(out, target, N, kmax=40)
    loss.backward()
    torch.cuda.synchronize()
    e = time.perf_counter()

    def L(cv):
        with torch.no_grad():
            return float(spectrum_loss(solver(U0, torch.tensor(cv, dtype=torch.float64)), target, N, kmax=40))

    fd = (L(cs0 + fd_h) - L(cs0 - fd_h)) / (2 * fd_h)
    res = {
        "N": N, "steps": steps, "dt": dt, "scheme": "ETD2", "les_model": "Vreman", "precision": "fp32",
        "c_vreman": c_vreman, "Cs": cs0, "checkpoint_every": checkpoint_every,
        "loss": float(loss), "dloss_dCs_adjoint": float(c.grad), "dloss_dCs_central_fd_h1e-3": fd,
        "dloss_dc_vreman_adjoint": float(c.grad) / (5.0 * cs0),
        "grad_U0_norm": float(U.grad.abs().pow(2).sum().sqrt()),
        "forward_s": b - a, "backward_s": e - b,
        "backward_over_forward": (e - b) / (b - a),
    }
    print(json.dumps(res), flush=True)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("precision")
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--steps", type=int, nargs="+", default=[100])
    p.add_argument("--les", default="Smagorinsky")
    p.add_argument("--out", default=None)
    q = sub.add_parser("differentiable")
    q.add_argument("--n", type=int, default=256)
    q.add_argument("--n-target", type=int, default=512)
    q.add_argument("--steps", type=int, default=50)
    q.add_argument("--out", default=None)
    r = sub.add_parser("ensemble")
    r.add_argument("--steps", type=int, default=200)
    r.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "differentiable":
        res = differentiable_study(a.n, a.n_target, a.steps)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        return
    if a.cmd == "ensemble":
        from .ensemble import Ensemble

        ens = Ensemble(((64, 8), (96, 8), (128, 8)), precision="fp32", seed=0)
        ens.capture(steps_per_graph=1)
        rate = ens.run(a.steps)
        res = {"members": [[d.N, d.Cs] for d in ens.members], "steps": a.steps,
               "grid_points_x_steps_per_s": rate, "dt_last": ens.dt_values()}
        print(json.dumps({"grid_points_x_steps_per_s": rate, "steps": a.steps}), flush=True)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        return
    if a.cmd == "precision":
        res = precision_study(a.n, tuple(a.steps), les=a.les)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
