"""Parity at the benchmark configurations' own sizes and settings.

* Decaying isotropic LES 256^3 FP64 (Smagorinsky Cs = 0.17, nu = 1e-4, RK4,
  CFL rule: first dt 0.3, then 0.5): the device driver loop against the
  oracle's loop for 10 steps -- velocity rel-L2 <= 1e-10, the dt after the
  last step and the calculate_Ek bins (src/SM_solver.py:154-192).
* The same at 512^3 in the precision form: FP32 vs FP64, 3 RK4 steps on the
  FP64 run's dt sequence, velocity rel-L2 <= 1e-5.
* Differentiable Vreman LES 256^3 FP32, ETD2, dt = 2e-3, checkpoints every 5
  over 10 steps, loss = squared log-mismatch of E(k), k <= 40, against a
  512^3 Smagorinsky run cut to 256^3: dL/dCs and a derivative along U0
  against central finite differences (h = 1e-3).
* The ensemble (64^3, 96^3, 128^3 members, FP32, Smagorinsky, own Cs and
  seed, graph-captured steps with device dt): each member against the FP64
  oracle after 3 steps, a member's d(loss)/dCs against central FD, and the
  full 8 x 64^3 + 8 x 96^3 + 8 x 128^3 ensemble replayed: every member
  finite, solenoidal and losing energy.
"""

import math

import numpy as np
import pytest
import torch

from helpers import oracle_loop, rel
from oracle import sles_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _isotropic_physical(N, seed, precision="fp64", urms=1.0):
    """Seeded k^4 exp(-2 (k/4)^2) solenoidal field (u_rms = 1) as physical
    components, for initial_velocity on both sides."""
    from sles.initial import isotropic_spectrum
    from sles.fft import SpectralFFT, Workspace
    from sles import _lib

    Uh = isotropic_spectrum(N, k0=4.0, seed=seed, urms=urms)
    plan = _lib.Plan(N, _lib.SLES_FP64, torch.device("cuda"))
    fft = SpectralFFT(plan, Workspace(N, torch.float64, torch.complex128, torch.device("cuda")), torch.float64,
                      torch.complex128, torch.device("cuda"))
    u = fft.backward(Uh, normalize=True).cpu().numpy().copy()
    del Uh
    torch.cuda.empty_cache()
    return u


def test_256_fp64_driver_loop_matches_oracle():
    from sles.initial import initial_velocity
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import Ek_plot_variables, calculate_Ek, spectralMethodSolver

    N, nsteps = 256, 10
    u = _isotropic_physical(N, 3)
    kw = dict(nu=1e-4, final_time=10.0, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17, CFL=0.5)
    g = InitializeDomain(N=N, **kw)
    initial_velocity(g, *u)
    steps = spectralMethodSolver(g, 0.0, No_Ek_plot=0, max_steps=nsteps, chunk=nsteps)
    torch.cuda.synchronize()
    o = O.Domain(N=N, fft_workers=min(O.cpu_count(), 16), **kw)
    O.initial_velocity(o, *u)
    del u
    n, _, dts = oracle_loop(o, 0.0, No_Ek_plot=0, max_steps=nsteps)
    assert steps == n == nsteps
    ru, rU = rel(torch.stack(g.U), np.array(o.U)), rel(g.U_hat, o.U_hat)
    print(f"[256 fp64] {nsteps} CFL steps: velocity rel-L2 {ru:.3e}, U_hat rel-L2 {rU:.3e}, "
          f"dt {dts[0]:.6g} .. {dts[-1]:.6g}", flush=True)
    assert ru < 1e-10 and rU < 1e-10
    assert math.isclose(g.dt, o.dt, rel_tol=1e-10)  # the dt after the last step
    E = Ek_plot_variables(g, path=None)
    calculate_Ek(g, E)
    Eo = O.EkBins(o)
    O.calculate_Ek(o, Eo)
    assert np.allclose(E.temp_k, Eo.temp_k, rtol=1e-13, atol=0)
    # every bin to 1e-13 of the peak bin; bins above 1e-8 of the peak (the
    # rest hold round-off of modes ~1e-4 below the largest and smaller: the
    # k^4 exp(-2 (k/4)^2) field has ~e^-200 of the peak at k = 40) to 1e-10
    peak = float(Eo.temp_Ek.max())
    absd = float(np.max(np.abs(E.temp_Ek - Eo.temp_Ek))) / peak
    res = Eo.temp_Ek >= 1e-8 * peak
    err = float(np.max(np.abs(E.temp_Ek[res] - Eo.temp_Ek[res]) / Eo.temp_Ek[res]))
    print(f"[256 fp64] calculate_Ek: {len(Eo.temp_Ek)} bins, max |diff| / peak {absd:.3e}; {int(res.sum())} bins "
          f"above 1e-8 of the peak, max rel diff {err:.3e}", flush=True)
    assert absd < 1e-13
    assert err < 1e-10


def test_512_fp32_form_vs_fp64():
    from sles import fused
    from sles.initial import isotropic_spectrum, set_spectral_state
    from sles.initialize_domain import InitializeDomain

    N = 512
    kw = dict(nu=1e-4, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17)
    d64 = InitializeDomain(N=N, precision="fp64", **kw)
    d32 = InitializeDomain(N=N, precision="fp32", **kw)
    Uh = isotropic_spectrum(N // 2, k0=4.0, seed=9, urms=1.0)  # the 256^3 field on the 512^3 grid
    from sles.studies import spectral_pad

    Uh = spectral_pad(Uh, N // 2, N)
    set_spectral_state(d64, Uh)
    set_spectral_state(d32, Uh.to(torch.complex64))
    del Uh
    for k in range(3):
        fused.rk4_step(d64, dt_device=True, cfl=0.3 if k == 0 else 0.5)
        dt = float(d64._scalars[1])
        assert dt >= 1e-6
        fused.rk4_step(d32, dt=dt)
    torch.cuda.synchronize()
    rv = sum(float((d32._U[i].double() - d64._U[i]).pow(2).sum()) for i in range(3)) / sum(
        float(d64._U[i].pow(2).sum()) for i in range(3))
    rv = math.sqrt(rv)
    print(f"[512] FP32 vs FP64 after 3 RK4 steps: velocity rel-L2 {rv:.3e}", flush=True)
    assert rv < 1e-5


def test_256_fp32_etd2_vreman_adjoint_vs_fd():
    from sles import fused
    from sles.adjoint import DifferentiableLES, spectrum_loss
    from sles.initial import isotropic_spectrum, set_spectral_state
    from sles.initialize_domain import InitializeDomain
    from sles.studies import shell_energy, spectral_cut, spectral_pad

    N, NT, dt = 256, 512, 2e-3
    # target: a 512^3 Smagorinsky run (FP32) cut to 256^3
    t = InitializeDomain(N=NT, nu=1e-4, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17,
                         precision="fp32")
    set_spectral_state(t, spectral_pad(isotropic_spectrum(N, seed=102, urms=0.25), N, NT).to(torch.complex64))
    for _ in range(3):
        fused.rk4_step(t, dt=dt, physical=False)
    target = shell_energy(spectral_cut(t.U_hat.to(torch.complex128), NT, N), N)[1:41].clone()
    del t
    torch.cuda.empty_cache()
    assert bool((target > 0).all())
    s = DifferentiableLES(N=N, nu=1e-4, dt=dt, nsteps=10, les_model="Vreman", scheme="ETD2", precision="fp32",
                          checkpoint_every=5)
    cs0 = math.sqrt(0.08 / 2.5)
    U0 = isotropic_spectrum(N, seed=2, urms=0.25, dtype=torch.complex64)
    U = U0.clone().requires_grad_(True)
    c = torch.tensor(cs0, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = spectrum_loss(s(U, c), target, N, kmax=40)
    loss.backward()
    h = 1e-3

    def L(U_, cv):
        with torch.no_grad():
            return float(spectrum_loss(s(U_, torch.tensor(cv, dtype=torch.float64)), target, N, kmax=40))

    fd = (L(U0, cs0 + h) - L(U0, cs0 - h)) / (2 * h)
    gcs = float(c.grad)
    adj_a = float((torch.conj(U.grad) * U0).real.sum().double())  # d/dalpha L((1 + alpha) U0)
    fd_a = (L(U0 * (1 + h), cs0) - L(U0 * (1 - h), cs0)) / (2 * h)
    print(f"[256 fp32 etd2 vreman] loss {float(loss):.6g}: dL/dCs adjoint {gcs:.6g} fd {fd:.6g}; "
          f"d/dalpha adjoint {adj_a:.6g} fd {fd_a:.6g}", flush=True)
    # FD of an FP32 loss: rounding noise ~1e-7 L over 2h; stated tolerance 2e-3
    assert abs(gcs - fd) < 2e-3 * abs(fd)
    assert abs(adj_a - fd_a) < 2e-3 * abs(fd_a)


def _oracle_member(d, U_hat0):
    """Oracle twin of an ensemble member from its initial spectrum."""
    o = O.Domain(N=d.N, nu=d.nu, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=d.Cs, CFL=0.5,
                 fft_workers=min(O.cpu_count(), 16))
    o.U_hat[:] = U_hat0
    for i in range(3):
        o.U[i][:] = o.fft.backward(o.U_hat[i], o.U[i], normalize=True)
    o.curl[:] = O.del_cross_U(o.fft, 3, o.U_hat, o.curl, o.K)
    return o


def test_ensemble_members_match_oracle():
    """One member each of 64^3, 96^3, 128^3 (FP32 form, own seed and Cs),
    advanced by three replays of the captured graph of batched launches: the
    first step at the CFL dt 0.3, the next ones at 0.5, all on the device; each
    member against the FP64 oracle run with its own CFL dt."""
    from sles.ensemble import S_DT_APPLY, S_STEPS, Ensemble

    ens = Ensemble(((64, 1), (96, 1), (128, 1)), precision="fp32", seed=7)
    init = [d.U_hat.to(torch.complex128).cpu().numpy() for d in ens.members]
    ens.capture(steps_per_graph=1)
    for _ in range(3):
        ens.graph.replay()
    torch.cuda.synchronize()
    st = ens.states()
    for i, (d, U0) in enumerate(zip(ens.members, init)):
        o = _oracle_member(d, U0)
        o.U_hat1[:] = o.U_hat
        for k in range(3):
            o.dt = O.compute_dt(o.U, O.find_max_velocity(o.U), o.nu, dx=o.dx, CFL=0.3 if k == 0 else 0.5)
            O.rk4_step(o)
        r = rel(d.U_hat.to(torch.complex128), o.U_hat)
        print(f"[ensemble] member {d.N}^3 Cs={d.Cs:.4f}: FP32 vs FP64 oracle after 3 steps rel-L2 {r:.3e}",
              flush=True)
        assert r < 1e-5
        assert st[i, S_STEPS] == 3
        assert math.isclose(float(st[i, S_DT_APPLY]), o.dt, rel_tol=1e-5)  # the device dt of step 3


def test_ensemble_member_adjoint_dcs_vs_fd():
    """d(loss)/dCs of an ensemble member (64^3 FP32 Smagorinsky RK4, 20 steps,
    checkpoints every 10, loss = squared log-change of its spectrum) through
    the hand-written adjoint against central FD, h = 1e-3 (FD of an FP32
    loss: stated tolerance 2e-3)."""
    from sles.studies import ensemble_adjoint

    out = ensemble_adjoint(groups=((64, 1),), steps=20, checkpoint_every=10, seed=3)
    chk = out["fd_check_member0"]
    print(f"[ensemble adjoint] {chk}", flush=True)
    assert chk["rel_diff"] < 2e-3


def _div_energy(U_hat, N):
    dev = U_hat.device
    k = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=dev)
    kz = torch.arange(N // 2 + 1, dtype=torch.float64, device=dev)
    U = U_hat.to(torch.complex128)
    div = k[:, None, None] * U[0] + k[None, :, None] * U[1] + kz[None, None, :] * U[2]
    kmag = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    scale = float((kmag * U.abs().pow(2).sum(0).sqrt()).max())
    w = torch.full_like(kmag, 2.0)
    w[..., 0] = 1.0
    w[..., -1] = 1.0
    return float(div.abs().max()) / scale, float((w * U.abs().pow(2).sum(0)).sum())


def test_full_ensemble_graph_replays():
    """The full ensemble of the batched configuration (8 x 64^3 + 8 x 96^3 +
    8 x 128^3, FP32, Cs ~ U[0.10, 0.20], own seeds) under one captured graph:
    3 replays; every member stays finite and solenoidal and loses kinetic
    energy (nu > 0, eddy viscosity >= 0, no forcing)."""
    from sles.ensemble import Ensemble

    from sles.ensemble import S_DT_APPLY, S_STATUS, S_STEPS

    ens = Ensemble(((64, 8), (96, 8), (128, 8)), precision="fp32", seed=0)
    assert len(ens.members) == 24 and len({d.Cs for d in ens.members}) == 24
    ens.capture(steps_per_graph=1)
    torch.cuda.synchronize()
    before = [_div_energy(d.U_hat, d.N)[1] for d in ens.members]
    for _ in range(3):
        ens.graph.replay()
    torch.cuda.synchronize()
    st = ens.states()
    for i, (d, e0) in enumerate(zip(ens.members, before)):
        assert bool(torch.isfinite(torch.view_as_real(d.U_hat)).all()), d.N
        div, e = _div_energy(d.U_hat, d.N)
        assert div < 1e-5, (d.N, div)
        assert e < e0, (d.N, e, e0)
        # every member took its 3 steps (the device dt guard skipped none)
        assert st[i, S_STEPS] == 3 and st[i, S_STATUS] == 0 and st[i, S_DT_APPLY] >= 1e-6
