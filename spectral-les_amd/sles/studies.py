"""Measurement drivers beyond the single-step benchmark.

precision_study  FP32 form vs the FP64 form on the same seeded field and the
                 same dt sequence: velocity rel-L2, max relative E(k) error,
                 steps/s of each form.
ensemble         many members of different grid sizes advanced together: one
                 launch per pass per grid size, replays of one captured HIP
                 graph (see ensemble.py); --per-member for one call per member.
ensemble-adjoint d(loss)/dCs of every member through the hand-written adjoint.
differentiable   Vreman ETD2 gradients against finite differences.
fft              generic 3D transforms (batch of 3 fields): ms and GB/s per call.

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
    """E(k), k = 0..kmax_shell, over unit shells |k| in [k-1/2, k+1/2) of
    (1/2) w |U_hat|^2 / N^6 (w the rfft doubling), FP64; entry 0 is the mean
    flow."""
    from .adjoint import shell_spectrum

    kmax = int(math.floor(math.sqrt(3.0) * (N // 2) + 0.5))
    e0 = 0.5 * (U_hat[:, 0, 0, 0].abs().double() ** 2).sum().reshape(1) / float(N) ** 6
    return torch.cat([e0, shell_spectrum(U_hat, N, kmax)])


def _rel_l2(a, b):
    """||a - b||_2 / ||b||_2 in FP64 (complex arrays as (Re, Im) pairs)."""
    if a.is_complex():
        a, b = torch.view_as_real(a), torch.view_as_real(b)
    a, b = a.double(), b.double()
    return float((a - b).norm() / b.norm())


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
            rel = (e32 - e64).abs() / e64
            # shells whose energy is above the FP32 round-off floor of the transforms
            # (E(k) >= 1e-12 of the peak shell; below it FP32 holds only rounding noise)
            resolved = e64 >= 1e-12 * e64.max()
            kres = int(resolved.nonzero().max()) + 1 if bool(resolved.any()) else 0
            rec = {
                "steps": n,
                "velocity_rel_l2": _rel_l2(d32._U, d64._U),
                "u_hat_rel_l2": _rel_l2(d32.U_hat, d64.U_hat),
                "max_rel_Ek_error_k_le_kmax": float(rel.max()),
                "argmax_k": int(rel.argmax()) + 1,
                "Ek_at_argmax_over_peak": float(e64[rel.argmax()] / e64.max()),
                "max_rel_Ek_error_resolved_shells": float(rel[resolved].max()),
                "resolved_kmax": kres,
                "kmax": kmax,
                "fp64_ms_per_step": 1e3 * t64 / n,
                "fp32_ms_per_step": 1e3 * t32 / n,
                "fp32_speedup": t64 / t32,
            }
            out["results"].append(rec)
            print(json.dumps(rec), flush=True)
    return out


def precision_tgv(N=64, steps=200, nu=1 / 1600, dt=0.01):
    """FP32 form against the FP64 form on the Taylor-Green case (no LES, RK4,
    fixed dt): velocity rel-L2 and max relative E(k) error after `steps`."""
    import numpy as np

    from .initial import initial_velocity

    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u, v, w = np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X)
    doms = []
    for prec in ("fp64", "fp32"):
        d = InitializeDomain(N=N, nu=nu, time_stepSize=dt, method="RK4", precision=prec)
        initial_velocity(d, u, v, w)
        d.U_hat1.copy_(d.U_hat)
        doms.append(d)
    for _ in range(steps):
        for d in doms:
            fused.rk4_step(d, dt=dt)
    d64, d32 = doms
    e64 = shell_energy(d64.U_hat, N)[1:]
    e32 = shell_energy(d32.U_hat, N)[1:]
    resolved = e64 >= 1e-12 * e64.max()
    rel = ((e32 - e64).abs() / e64)[resolved]
    res = {"case": "taylor_green", "N": N, "steps": steps, "velocity_rel_l2": _rel_l2(d32._U, d64._U),
           "max_rel_Ek_error_resolved_shells": float(rel.max()), "resolved_kmax": int(resolved.nonzero().max()) + 1}
    print(json.dumps(res), flush=True)
    return res


def spectral_cut(U_hat, N_from, N_to):
    """Keep |kx|, |ky| < N_to/2 and kz <= N_to/2 of an N_from^3 spectrum, scaled
    by (N_to/N_from)^3 for the unscaled forward-transform convention (the
    512^3 -> 256^3 target of the differentiable configuration: x 1/8).  For a
    field whose modes all satisfy |k_i| < N_to/2 this is exactly the transform
    of the field sampled on the N_to^3 grid."""
    h = N_to // 2
    sel = torch.cat([torch.arange(0, h), torch.arange(N_from - h, N_from)]).to(U_hat.device)
    out = U_hat[:, sel][:, :, sel][..., : h + 1].clone()
    out[:, h] = 0  # kx = -N_to/2 (|kx| < N_to/2 kept)
    out[:, :, h] = 0
    return out * (float(N_to) / float(N_from)) ** 3


def spectral_pad(U_hat, N_from, N_to):
    """Zero-padded N_to^3 spectrum of an N_from^3 one (N_to > N_from), scaled
    by (N_to/N_from)^3: the same band-limited field on the finer grid, with
    the N_from/2 Nyquist planes dropped.  spectral_cut(spectral_pad(U)) == U
    for a U without Nyquist modes."""
    h = N_from // 2
    out = torch.zeros(U_hat.shape[:-3] + (N_to, N_to, N_to // 2 + 1), dtype=U_hat.dtype, device=U_hat.device)
    lo = torch.arange(0, h, device=U_hat.device)
    src = torch.cat([lo, torch.arange(N_from - h + 1, N_from, device=U_hat.device)])
    dst = torch.cat([lo, torch.arange(N_to - h + 1, N_to, device=U_hat.device)])
    sub = U_hat[..., src, :, :][..., :, src, :][..., :h]
    out[..., dst[:, None], dst[None, :], :h] = sub
    return out * (float(N_to) / float(N_from)) ** 3


def differentiable_study(N=256, N_target=512, steps=50, target_steps=50, dt=2e-3, nu=1e-4, c_vreman=0.08,
                         checkpoint_every=10, seed=2, fd_h=1e-3, urms=0.25):
    """Vreman LES, FP32, ETD2, dt = 2e-3 over a `steps` window; loss = squared
    log-mismatch of E(k), k <= 40, against a Smagorinsky N_target^3 run cut to N^3.

    ETD2 advances the nonlinear term with the two-step (Adams-Bashforth-like)
    weights h(phi1 + phi2), -h phi2, whose stability interval on the imaginary
    axis ends near h |k| |u| = 0.5 (amplification 1.13 per step at 0.68, 1.5
    at 1.0).  At N = 256 (|k| up to 147 after dealiasing) and h = 2e-3 that
    needs max |u| below ~1.5, so the initial field has u_rms = 0.25 (max |u|
    about 1.1); with u_rms = 1 the ETD2 run diverges within 50 steps in both
    precisions while RK4 at the same h stays stable."""
    from .adjoint import DifferentiableLES, spectrum_loss

    # target spectrum
    t = InitializeDomain(N=N_target, nu=nu, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17,
                         precision="fp32")
    set_spectral_state(t, isotropic_spectrum(N_target, seed=seed + 100, urms=urms, dtype=t.cdtype))
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
    U0 = isotropic_spectrum(N, seed=seed, urms=urms, dtype=torch.complex64)
    from .SM_solver import find_max_velocity

    set_spectral_state(solver.d, U0)
    umax0 = float(find_max_velocity(solver.d.U))

    def forward_backward():
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
        return U, c, loss, b - a, e - b

    forward_backward()  # warm-up (workspaces, plans)
    U, c, loss, t_fwd, t_bwd = forward_backward()

    def L(cv):
        with torch.no_grad():
            return float(spectrum_loss(solver(U0, torch.tensor(cv, dtype=torch.float64)), target, N, kmax=40))

    fd = (L(cs0 + fd_h) - L(cs0 - fd_h)) / (2 * fd_h)

    # initial-field gradient along the radial direction U0: d/da L((1 + a) U0) at a = 0
    def LU(a_):
        with torch.no_grad():
            return float(spectrum_loss(solver((1.0 + a_) * U0, torch.tensor(cs0, dtype=torch.float64)), target, N,
                                       kmax=40))

    fdU = (LU(fd_h) - LU(-fd_h)) / (2 * fd_h)
    adjU = float((U.grad.real.double() * U0.real.double() + U.grad.imag.double() * U0.imag.double()).sum())
    res = {
        "N": N, "steps": steps, "dt": dt, "scheme": "ETD2", "les_model": "Vreman", "precision": "fp32", "urms": urms,
        "c_vreman": c_vreman, "Cs": cs0, "checkpoint_every": checkpoint_every,
        "loss": float(loss.detach()), "dloss_dCs_adjoint": float(c.grad), "dloss_dCs_central_fd_h1e-3": fd,
        "adjoint_vs_fd_rel_diff": abs(float(c.grad) - fd) / max(abs(fd), 1e-300),
        "dloss_dc_vreman_adjoint": float(c.grad) / (5.0 * cs0),
        "grad_U0_norm": float(U.grad.abs().pow(2).sum().sqrt()),
        "dloss_dalpha_U0_adjoint": adjU, "dloss_dalpha_U0_central_fd_h1e-3": fdU,
        "U0_direction_rel_diff": abs(adjU - fdU) / max(abs(fdU), 1e-300),
        "max_u_initial": umax0,
        "forward_s": t_fwd, "backward_s": t_bwd,
        "backward_over_forward": t_bwd / t_fwd,
    }
    print(json.dumps(res), flush=True)
    return res


def adjoint_cost(N=256, scheme="RK4", les="Smagorinsky", precision="fp32", steps=20, checkpoint_every=10,
                 dt=1e-3, Cs=0.17, urms=0.25, seed=4, keep_last_segment=True):
    """Backward / forward time of a DifferentiableLES window (loss = sum of
    E(k), k <= N/3, so the adjoint seed is cheap): the cost of the adjoint step
    relative to the forward step, with checkpointing (keep_last_segment=False:
    every segment recomputed in the backward sweep)."""
    from .adjoint import DifferentiableLES, shell_spectrum

    s = DifferentiableLES(N=N, nu=1e-4, dt=dt, nsteps=steps, les_model=les, scheme=scheme, precision=precision,
                          checkpoint_every=checkpoint_every, keep_last_segment=keep_last_segment)
    cd = torch.complex64 if precision == "fp32" else torch.complex128
    U0 = isotropic_spectrum(N, seed=seed, urms=urms, dtype=cd)

    def once():
        U = U0.clone().requires_grad_(True)
        c = torch.tensor(Cs, dtype=torch.float64, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        a = time.perf_counter()
        out = s(U, c)
        torch.cuda.synchronize()
        b = time.perf_counter()
        shell_spectrum(out, N, N // 3).sum().backward()
        torch.cuda.synchronize()
        e = time.perf_counter()
        return b - a, e - b

    once()
    f, bwd = once()
    res = {"N": N, "scheme": scheme, "les_model": les, "precision": precision, "steps": steps,
           "checkpoint_every": checkpoint_every, "keep_last_segment": keep_last_segment, "forward_s": f,
           "backward_s": bwd, "backward_over_forward": bwd / f}
    print(json.dumps(res), flush=True)
    return res


def ensemble_rate(groups=((64, 8), (96, 8), (128, 8)), steps=200, steps_per_graph=10, batched=True,
                  precision="fp32", seed=0):
    """Aggregate grid points x steps / s of an ensemble (Smagorinsky RK4,
    Cs ~ U[0.10, 0.20], own seeds, CFL dt on the device) advanced by replays
    of one captured graph of steps_per_graph ensemble steps; batched: one
    launch per pass per grid size, else one sles_step_drive call per member.
    One untimed replay first."""
    from .ensemble import S_STATUS, S_STEPS, Ensemble

    ens = Ensemble(groups, precision=precision, seed=seed, batched=batched)
    ens.capture(steps_per_graph=steps_per_graph)
    ens.graph.replay()  # untimed: first replay
    torch.cuda.synchronize()
    rate = ens.run(steps)
    st = ens.states()
    ok = bool((st[:, S_STATUS] == 0).all()) and bool((st[:, S_STEPS] == steps + steps_per_graph).all())
    return {"groups": [list(g) for g in groups], "members": [[d.N, d.Cs] for d in ens.members], "steps": steps,
            "steps_per_graph": steps_per_graph, "batched": batched, "precision": precision,
            "grid_points_x_steps_per_s": rate, "all_members_stepped": ok, "dt_last": ens.dt_values()}


def ensemble_adjoint(groups=((64, 8), (96, 8), (128, 8)), steps=200, checkpoint_every=10, nu=1e-4, seed=0,
                     cs_range=(0.10, 0.20), fd_check=True):
    """d(loss)/dCs for every member of the ensemble after `steps` RK4 steps
    (Smagorinsky, FP32 form, hand-written adjoint with checkpointing).  Member
    i: its own seed, Cs ~ U[cs_range], fixed dt from the CFL rule (0.3) on its
    initial field.  Loss_i = sum_{k=1}^{kmax} (log E_T(k) - log E_0(k))^2, the
    squared log-change of the member's spectrum over the window, kmax =
    min(N/3, 24): beyond k ~ 25 the k0 = 4 initial spectrum (amplitude
    k exp(-(k/4)^2)) is below the FP32 range and E_0(k) would be 0.  Returns the
    per-member gradients, a central-difference check (h = 1e-3) on the first
    member, and the forward + adjoint throughput over all members."""
    import numpy as np

    from .adjoint import DifferentiableLES, shell_spectrum
    from .SM_solver import compute_dt, find_max_velocity

    rng = np.random.default_rng(seed)
    members = []
    i = 0
    for N, count in groups:
        for _ in range(count):
            Cs = float(rng.uniform(*cs_range))
            s = DifferentiableLES(N=N, nu=nu, dt=1e-3, nsteps=steps, les_model="Smagorinsky", scheme="RK4",
                                  precision="fp32", checkpoint_every=checkpoint_every)
            U0 = isotropic_spectrum(N, k0=4.0, seed=seed * 1000 + i, urms=1.0, dtype=torch.complex64)
            set_spectral_state(s.d, U0)
            s.dt = compute_dt(s.d.U, find_max_velocity(s.d.U), nu, dx=s.d.dx, CFL=0.3)
            s.d.dt = s.dt
            kmax = min(N // 3, 24)
            E0 = shell_spectrum(U0, N, kmax).detach()
            members.append((s, U0, Cs, kmax, E0))
            i += 1

    def loss_of(s, U0, cs, kmax, E0):
        E = shell_spectrum(s(U0, cs), s.N, kmax)
        return ((torch.log(E) - torch.log(E0)) ** 2).sum()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grads = []
    for s, U0, Cs, kmax, E0 in members:
        c = torch.tensor(Cs, dtype=torch.float64, device="cuda", requires_grad=True)
        L_ = loss_of(s, U0, c, kmax, E0)
        L_.backward()
        grads.append({"N": s.N, "Cs": Cs, "dt": s.dt, "loss": float(L_.detach()), "dloss_dCs": float(c.grad)})
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    points = sum(s.N ** 3 for s, *_ in members)
    out = {"members": grads, "steps": steps, "checkpoint_every": checkpoint_every,
           "forward_plus_adjoint_s": elapsed,
           "grid_points_x_steps_per_s_forward_plus_adjoint": points * steps / elapsed}
    if fd_check:
        s, U0, Cs, kmax, E0 = members[0]
        h = 1e-3

        def L(cv):
            with torch.no_grad():
                return float(loss_of(s, U0, torch.tensor(cv, dtype=torch.float64), kmax, E0))

        fd = (L(Cs + h) - L(Cs - h)) / (2 * h)
        out["fd_check_member0"] = {"adjoint": grads[0]["dloss_dCs"], "central_fd_h1e-3": fd,
                                   "rel_diff": abs(grads[0]["dloss_dCs"] - fd) / max(abs(fd), 1e-300)}
    print(json.dumps({k: v for k, v in out.items() if k != "members"}), flush=True)
    return out


def stream_ceiling(N=256, precision="fp64", mixes=((15, 6),), tiles=(0, 1, 2), xcd=1, reps=10, verbose=True,
                   pad=None):
    """Bandwidth of the strided x-pass access pattern with no arithmetic
    (sles_stream_probe): for each tile shape and (fields read, fields
    written) mix, the algorithmic bytes over the mean launch time.  The
    ceiling of a pass with that tile shape and traffic mix on this device.
    The fields have the fused step's workspace layout: kx planes padded by
    ``pad`` 128-byte lines (None: the plan's default)."""
    from . import _lib

    prec = _lib.SLES_FP64 if precision == "fp64" else _lib.SLES_FP32
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    plan = _lib.Plan(N, prec, torch.device("cuda"))
    plan.set_tuning(xcd=xcd)
    if pad is not None:
        plan.set_tuning(pad=pad)
    pad = plan.get_tuning()["pad"]
    esz = 16 if precision == "fp64" else 8
    plane = N * (N // 2 + 1) + pad * 128 // esz
    nin = max(r for r, _ in mixes)
    nout = max(w for _, w in mixes)
    a = torch.zeros((nin, N, plane), dtype=cdt, device="cuda")
    b = torch.zeros((max(nout, 1), N, plane), dtype=cdt, device="cuda")
    C = N * N * (N // 2 + 1) * esz  # algorithmic bytes of one field (padding excluded)
    s = torch.cuda.current_stream().cuda_stream
    out = {}
    for tile in tiles:
        for r, w in mixes:
            def once():
                _lib.call("sles_stream_probe", plan.handle, a.data_ptr(), r, b.data_ptr(), w, tile, s)

            once()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                once()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1) / reps
            gbs = (r + w) * C / (ms * 1e-3) / 1e9
            out[f"tile{tile}_{r}r{w}w"] = {"ms": round(ms, 4), "GBps": round(gbs, 1)}
            if verbose:
                print(f"[stream] N={N} {precision} pad {pad} tile {tile} {r} read + {w} written: {ms:.4f} ms "
                      f"{gbs:.0f} GB/s", flush=True)
    del a, b
    torch.cuda.empty_cache()
    return out


def fft_throughput(sizes=(256, 512), precisions=("fp64", "fp32"), batch=3, reps=10):
    """Generic 3D transforms (sles_fft_r2c / sles_fft_c2r) of `batch` fields
    with each plan's default tiles: ms per call and algorithmic GB/s,
    batch (R + 5 C) bytes per call (z pass R in / C out, y and x passes C in /
    C out, each field once)."""
    from . import _lib

    out = {}
    for N in sizes:
        for pr in precisions:
            fp64 = pr == "fp64"
            plan = _lib.Plan(N, _lib.SLES_FP64 if fp64 else _lib.SLES_FP32, torch.device("cuda"))
            s = 8 if fp64 else 4
            R, C = N ** 3 * s, N * N * (N // 2 + 1) * 2 * s
            u = torch.randn(batch, N, N, N, dtype=torch.float64 if fp64 else torch.float32, device="cuda")
            U = torch.empty(batch, N, N, N // 2 + 1, dtype=torch.complex128 if fp64 else torch.complex64,
                            device="cuda")
            st = torch.cuda.current_stream().cuda_stream
            calls = {
                "r2c": lambda: _lib.call("sles_fft_r2c", plan.handle, u.data_ptr(), U.data_ptr(), batch, 1.0, st),
                "c2r": lambda: _lib.call("sles_fft_c2r", plan.handle, U.data_ptr(), u.data_ptr(), U.data_ptr(),
                                         batch, 1.0 / N ** 3, st),
            }
            res = {"tuning": plan.get_tuning()}
            for name, fn in calls.items():
                fn()
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    fn()
                e1.record()
                e1.synchronize()
                ms = e0.elapsed_time(e1) / reps
                gbs = batch * (R + 5 * C) / (ms * 1e-3) / 1e9
                res[name] = {"ms": round(ms, 4), "algorithmic_GBps": round(gbs, 1),
                             "frac_of_8TBps": round(gbs / 8000.0, 4)}
            out[f"{N}_{pr}"] = res
            print(f"{N}_{pr}", json.dumps(res), flush=True)
            del u, U, plan
            torch.cuda.empty_cache()
    return out


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
    r.add_argument("--steps-per-graph", type=int, default=10)
    r.add_argument("--per-member", action="store_true", help="one sles_step_drive call per member")
    r.add_argument("--out", default=None)
    pt = sub.add_parser("precision-tgv")
    pt.add_argument("--n", type=int, default=64)
    pt.add_argument("--steps", type=int, default=200)
    rc = sub.add_parser("adjoint-cost")
    rc.add_argument("--n", type=int, default=256)
    rc.add_argument("--scheme", default="RK4")
    rc.add_argument("--les", default="Smagorinsky")
    rc.add_argument("--precision", default="fp32")
    rc.add_argument("--steps", type=int, default=20)
    rc.add_argument("--recompute-all", action="store_true", help="recompute the last segment too")
    ff = sub.add_parser("fft")
    ff.add_argument("--n", type=int, nargs="+", default=[256, 512])
    ff.add_argument("--reps", type=int, default=10)
    st = sub.add_parser("stream")
    st.add_argument("--n", type=int, nargs="+", default=[256, 512])
    st.add_argument("--precision", nargs="+", default=["fp64", "fp32"])
    st.add_argument("--pad", type=int, nargs="+", default=[None], help="plane padding, 128-byte lines")
    st.add_argument("--mixes", nargs="+", default=["9:9", "15:6", "12:6", "3:6", "1:1"],
                    help="fields read:written per launch")
    st.add_argument("--tiles", type=int, nargs="+", default=[0, 1, 2])
    ra = sub.add_parser("ensemble-adjoint")
    ra.add_argument("--steps", type=int, default=200)
    ra.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "differentiable":
        res = differentiable_study(a.n, a.n_target, a.steps)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        return
    if a.cmd == "stream":
        # default mixes of the x passes: the x-forward + update of RK4 stages
        # 0 / 1-2 / 3 (9+9, 15+6, 12+6 fields), the x-inverse (3+6) and a
        # plain 1+1; with the next x-inverse fused the x-forward writes 6 more
        # (9:15 15:12 12:12)
        mixes = tuple(tuple(int(v) for v in m.split(":")) for m in a.mixes)
        for n in a.n:
            for prec in a.precision:
                for pad in a.pad:
                    stream_ceiling(n, prec, mixes=mixes, tiles=tuple(a.tiles), pad=pad)
        return
    if a.cmd == "fft":
        fft_throughput(tuple(a.n), reps=a.reps)
        return
    if a.cmd == "precision-tgv":
        precision_tgv(a.n, a.steps)
        return
    if a.cmd == "adjoint-cost":
        adjoint_cost(a.n, a.scheme, a.les, a.precision, a.steps, keep_last_segment=not a.recompute_all)
        return
    if a.cmd == "ensemble-adjoint":
        res = ensemble_adjoint(steps=a.steps)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        return
    if a.cmd == "ensemble":
        res = ensemble_rate(steps=a.steps, steps_per_graph=a.steps_per_graph, batched=not a.per_member)
        print(json.dumps({k: res[k] for k in ("grid_points_x_steps_per_s", "steps", "steps_per_graph", "batched")}),
              flush=True)
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
