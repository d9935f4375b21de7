"""Hand-written adjoint: the VJP kernel against finite differences of the
forward RHS kernel (dot-product test), d/dCs against finite differences, and
whole RK4 / ETD2 trajectories against an FP64 torch-autograd reference
(tests/torch_reference.py, CPU) and central finite differences (h = 1e-3
for Cs, as in the differentiable-LES configuration)."""

import math

import numpy as np
import pytest
import torch

import torch_reference as TR
from helpers import random_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rdot(a, b):
    """Real inner product of complex arrays viewed as (Re, Im) vectors."""
    return float((torch.conj(a) * b).real.sum())


def _solver(N, les, scheme="RK4", nsteps=3, dt=0.01, nu=0.01, precision="fp64", K=2, forcing=0.0):
    from sles.adjoint import DifferentiableLES

    return DifferentiableLES(N=N, nu=nu, dt=dt, nsteps=nsteps, les_model=les, scheme=scheme, precision=precision,
                             checkpoint_every=K, forcing=forcing)


def _state(N, seed, k_cut=3.0):
    u = random_field(N, seed, k_cut=k_cut)
    U = np.fft.rfftn(u, axes=(1, 2, 3))
    return torch.as_tensor(U, device="cuda")


@pytest.mark.parametrize("les", [None, "Smagorinsky", "Vreman"])
@pytest.mark.parametrize("N", [16, 24])
@pytest.mark.parametrize("forcing", [0.0, 0.1])
def test_vjp_dot_product(les, N, forcing):
    """<J(Y) v, mu> = <v, J(Y)^T mu> with J v by central differences."""
    s = _solver(N, les, forcing=forcing)
    s._set_cs(0.17)
    Y = _state(N, 1)
    g = torch.Generator(device="cuda").manual_seed(2)
    v = torch.complex(torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64),
                      torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64))
    mu = torch.complex(torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64),
                       torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64))
    eps = 1e-6 * float(Y.abs().max()) / float(v.abs().max())
    fp = torch.empty_like(Y)
    fm = torch.empty_like(Y)
    s.rhs(Y + eps * v, fp)
    s.rhs(Y - eps * v, fm)
    Jv = (fp - fm) / (2 * eps)
    JTmu = torch.empty_like(Y)
    s.vjp(Y, mu, JTmu)
    lhs, rhs = rdot(mu, Jv), rdot(JTmu, v)
    assert abs(lhs - rhs) < 1e-6 * max(abs(lhs), abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("les", ["Smagorinsky", "Vreman"])
def test_vjp_cs_derivative(les):
    N = 16
    s = _solver(N, les)
    Y = _state(N, 3)
    g = torch.Generator(device="cuda").manual_seed(4)
    mu = torch.complex(torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64),
                       torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64))
    Cs, h = 0.15, 1e-5
    out = torch.empty_like(Y)

    def f(c):
        s._set_cs(c)
        s.rhs(Y, out)
        return rdot(mu, out)

    fd = (f(Cs + h) - f(Cs - h)) / (2 * h)
    s._set_cs(Cs)
    s.cs_grad.zero_()
    s.vjp(Y, mu, out)
    got = float(s.cs_grad.item())
    assert abs(got - fd) < 1e-6 * abs(fd), (got, fd)


def _torch_reference_grads(U0, Cs, N, nu, dt, nsteps, les, target):
    """FP64 autograd on CPU through torch_reference.rk4_step."""
    U = U0.detach().cpu().clone().requires_grad_(True)
    c = torch.tensor(Cs, dtype=torch.float64, requires_grad=True)
    X = U
    for _ in range(nsteps):
        X = TR.rk4_step(X, c, N, nu, dt, les)
    E = TR.shell_spectrum(X, N, kmax=target.numel())
    loss = ((torch.log(E) - torch.log(target)) ** 2).sum()
    loss.backward()
    return float(loss), U.grad, float(c.grad)


@pytest.mark.parametrize("les", ["Smagorinsky", "Vreman"])
def test_rk4_adjoint_vs_autograd_and_fd(les):
    from sles.adjoint import spectrum_loss

    N, nu, dt, nsteps = 16, 0.01, 0.01, 4
    U0 = _state(N, 7)
    target = TR.shell_spectrum(torch.as_tensor(_state(N, 8).cpu()), N, kmax=6) * 1.3
    Cs = 0.17
    ref_loss, ref_gU, ref_gCs = _torch_reference_grads(U0, Cs, N, nu, dt, nsteps, les, target)

    s = _solver(N, les, nsteps=nsteps, dt=dt, nu=nu, K=3)
    U = U0.clone().requires_grad_(True)
    c = torch.tensor(Cs, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = spectrum_loss(s(U, c), target.cuda(), N, kmax=6)
    loss.backward()
    assert math.isclose(float(loss), ref_loss, rel_tol=1e-10)
    gU = U.grad.cpu()
    assert float((gU - ref_gU).abs().max()) < 1e-8 * float(ref_gU.abs().max())
    assert math.isclose(float(c.grad), ref_gCs, rel_tol=1e-8)
    # central finite differences in Cs, h = 1e-3
    h = 1e-3

    def L(cv):
        with torch.no_grad():
            return float(spectrum_loss(s(U0, torch.tensor(cv, dtype=torch.float64)), target.cuda(), N, kmax=6))

    fd = (L(Cs + h) - L(Cs - h)) / (2 * h)
    assert abs(float(c.grad) - fd) < 1e-4 * abs(fd) + 1e-12
    # directional finite difference in U0
    g = torch.Generator(device="cuda").manual_seed(9)
    v = torch.complex(torch.randn(U0.shape, generator=g, device="cuda", dtype=torch.float64),
                      torch.randn(U0.shape, generator=g, device="cuda", dtype=torch.float64))
    v *= float(U0.abs().max()) * 1e-6 / float(v.abs().max())
    with torch.no_grad():
        Lp = float(spectrum_loss(s(U0 + v, torch.tensor(Cs)), target.cuda(), N, kmax=6))
        Lm = float(spectrum_loss(s(U0 - v, torch.tensor(Cs)), target.cuda(), N, kmax=6))
    dd = (Lp - Lm) / 2
    assert abs(rdot(U.grad, v) - dd) < 1e-5 * abs(dd)


@pytest.mark.parametrize("les", ["Vreman", "Smagorinsky"])
def test_etd2_adjoint_vs_fd(les):
    from sles.adjoint import spectrum_loss

    N, nu, dt, nsteps = 16, 0.01, 0.005, 6
    U0 = _state(N, 11)
    target = TR.shell_spectrum(torch.as_tensor(_state(N, 12).cpu()), N, kmax=6)
    s = _solver(N, les, scheme="ETD2", nsteps=nsteps, dt=dt, nu=nu, K=4)
    Cs = 0.08 if les == "Vreman" else 0.17
    U = U0.clone().requires_grad_(True)
    c = torch.tensor(Cs, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = spectrum_loss(s(U, c), target.cuda(), N, kmax=6)
    loss.backward()
    h = 1e-3

    def L(U_, cv):
        with torch.no_grad():
            return float(spectrum_loss(s(U_, torch.tensor(cv, dtype=torch.float64)), target.cuda(), N, kmax=6))

    fd = (L(U0, Cs + h) - L(U0, Cs - h)) / (2 * h)
    assert abs(float(c.grad) - fd) < 1e-4 * abs(fd) + 1e-12
    g = torch.Generator(device="cuda").manual_seed(13)
    v = torch.complex(torch.randn(U0.shape, generator=g, device="cuda", dtype=torch.float64),
                      torch.randn(U0.shape, generator=g, device="cuda", dtype=torch.float64))
    v *= float(U0.abs().max()) * 1e-6 / float(v.abs().max())
    dd = (L(U0 + v, Cs) - L(U0 - v, Cs)) / 2
    assert abs(rdot(U.grad, v) - dd) < 1e-5 * abs(dd)


@pytest.mark.parametrize("les", [None, "Smagorinsky", "Vreman"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("with_lam2", [True, False])
def test_fused_etd2_adjoint_step_matches_combine_plus_vjp(les, precision, with_lam2):
    """sles_etd2_adjoint_rhs == sles_etd2_adjoint_combine followed by
    sles_adjoint_rhs(accumulate, linear = 2), values and d/dCs."""
    import ctypes

    from sles import _lib, fused

    N = 16
    s = _solver(N, les, scheme="ETD2", dt=0.004, nu=0.02, precision=precision, forcing=0.1)
    s._set_cs(0.15)
    d = s.d
    cd = d.cdtype
    Y = _state(N, 21).to(cd)
    g = torch.Generator(device="cuda").manual_seed(22)

    def rnd():
        return torch.complex(torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64),
                             torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64)).to(cd)

    lam1, lam2 = rnd(), (rnd() if with_lam2 else None)
    mu = torch.empty_like(lam1)
    ref = torch.empty_like(lam1)
    _lib.call("sles_etd2_adjoint_combine", d.plan.handle, s.dt, s.nu, lam1.data_ptr(),
              lam2.data_ptr() if with_lam2 else None, mu.data_ptr(), ref.data_ptr(), d.stream)
    s.cs_grad.zero_()
    s.vjp(Y, mu, ref, accumulate=True, linear=2)
    cs_ref = float(s.cs_grad.item())
    out = torch.empty_like(lam1)
    s.cs_grad.zero_()
    prm = fused.params(d, dt=s.dt)
    _lib.call("sles_etd2_adjoint_rhs", d.plan.handle, ctypes.byref(prm), Y.data_ptr(), lam1.data_ptr(),
              lam2.data_ptr() if with_lam2 else None, out.data_ptr(), s.dcoeff(),
              s.cs_grad.data_ptr() if d.LES_ON else None, s.adj_work.data_ptr(), d.stream)
    torch.cuda.synchronize()
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert float((out - ref).abs().max()) <= tol * float(ref.abs().max())
    if les:
        assert abs(float(s.cs_grad.item()) - cs_ref) <= tol * abs(cs_ref)


@pytest.mark.parametrize("les", [None, "Vreman"])
def test_rk4_stage_states(les):
    """sles_rk4_stage_states: the stage states equal U + dt b_s k_s built from
    the RHS entry point, and the final state equals a plain fused RK4 step
    bit for bit."""
    from sles import fused

    N = 16
    s = _solver(N, les, dt=0.01, nu=0.02)
    s._set_cs(0.15)
    U = _state(N, 31)
    Ys, new = s._rk4_stages(U.clone())
    assert torch.equal(Ys[0], U)
    # reference stages from the RHS
    k = torch.empty_like(U)
    ref = [U]
    for st in range(3):
        s.rhs(ref[st], k)
        ref.append(U + s.dt * (0.5, 0.5, 1.0)[st] * k)
    for a, b in zip(Ys[1:], ref[1:]):
        assert float((a - b).abs().max()) < 1e-12 * float(b.abs().max())
    d = s.d
    d.U_hat.copy_(U)
    d.U_hat1.copy_(U)
    fused.rk4_step(d, dt=s.dt, physical=False)
    assert torch.equal(d.U_hat, new)


@pytest.mark.parametrize("les", [None, "Smagorinsky", "Vreman"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vjp_combined(les, precision):
    """sles_adjoint_rhs_combined against its composition from separate calls
    (sles_lincomb then sles_adjoint_rhs), with and without z, the d/dCs
    accumulation, and a plain sles_adjoint_rhs call afterwards (the combined
    call stores its mu in the workspace; the plain call must not read it)."""
    N = 24
    s = _solver(N, les, precision=precision, forcing=0.05)
    s._set_cs(0.16)
    cd = s.d.cdtype
    Y = _state(N, 5).to(cd)
    g = torch.Generator(device="cuda").manual_seed(6)

    def rnd():
        return torch.complex(torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64),
                             torch.randn(Y.shape, generator=g, device="cuda", dtype=torch.float64)).to(cd)

    x, z = rnd(), rnd()
    a, b = 0.37, -1.25
    tol = 1e-12 if precision == "fp64" else 2e-5
    mu = torch.empty_like(Y)
    ref = torch.empty_like(Y)
    out = torch.empty_like(Y)
    for zz, bb in ((z, b), (None, 0.0)):
        s.lincomb(a, x, bb, zz if zz is not None else x, mu)
        s.cs_grad.zero_()
        s.vjp(Y, mu, ref)
        cs_ref = float(s.cs_grad.item())
        s.cs_grad.zero_()
        s.vjp_combined(Y, a, x, bb, zz, out)
        assert float((out - ref).abs().max()) <= tol * float(ref.abs().max())
        if les:
            assert abs(float(s.cs_grad.item()) - cs_ref) <= tol * abs(cs_ref)
    # plain call after a combined one reads its own mu
    s.vjp_combined(Y, a, x, b, z, out)
    s.vjp(Y, x, out)
    s.vjp(Y, x, ref)
    assert torch.equal(out, ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sum_fields(precision):
    N = 16
    s = _solver(N, None, precision=precision)
    cd = s.d.cdtype
    g = torch.Generator(device="cuda").manual_seed(8)
    ts = [torch.complex(torch.randn((3, N, N, N // 2 + 1), generator=g, device="cuda", dtype=torch.float64),
                        torch.randn((3, N, N, N // 2 + 1), generator=g, device="cuda", dtype=torch.float64)).to(cd)
          for _ in range(5)]
    out = torch.empty_like(ts[0])
    s.sum_fields(ts, out)
    ref = ts[0].clone()
    for t in ts[1:]:
        ref = ref + t
    assert torch.equal(out, ref)
    s.sum_fields(ts[:1], out)
    assert torch.equal(out, ts[0])


@pytest.mark.parametrize("scheme", ["RK4", "ETD2"])
def test_keep_last_segment_is_bitwise_neutral(scheme):
    """Keeping the last segment's states in the forward pass (no recomputation
    for it) gives the same final state and gradients, bit for bit, as
    recomputing every segment from its checkpoint."""
    from sles.adjoint import DifferentiableLES, spectrum_loss

    N, nsteps, K = 16, 7, 3
    U0 = _state(N, 41)
    target = TR.shell_spectrum(torch.as_tensor(_state(N, 42).cpu()), N, kmax=6).cuda()
    res = []
    for keep in (True, False):
        s = DifferentiableLES(N=N, nu=0.01, dt=0.005, nsteps=nsteps, les_model="Vreman", scheme=scheme,
                              precision="fp64", checkpoint_every=K, keep_last_segment=keep)
        U = U0.clone().requires_grad_(True)
        c = torch.tensor(0.1, dtype=torch.float64, device="cuda", requires_grad=True)
        out = s(U, c)
        loss = spectrum_loss(out, target, N, kmax=6)
        loss.backward()
        res.append((out.detach().clone(), U.grad.clone(), float(c.grad)))
    assert torch.equal(res[0][0], res[1][0])
    assert torch.equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]


def _torch_reference_etd2_grads(U0, Cs, N, nu, dt, nsteps, les, target):
    """FP64 autograd on CPU through torch_reference.etd2_run (RK4 bootstrap,
    then Cox-Matthews ETD2 with the SGS term in F)."""
    U = U0.detach().cpu().to(torch.complex128).clone().requires_grad_(True)
    c = torch.tensor(Cs, dtype=torch.float64, requires_grad=True)
    X = TR.etd2_run(U, c, N, nu, dt, nsteps, les)
    E = TR.shell_spectrum(X, N, kmax=target.numel())
    loss = ((torch.log(E) - torch.log(target)) ** 2).sum()
    loss.backward()
    return float(loss.detach()), U.grad, float(c.grad)


# Tolerances of the ETD2 + Vreman adjoint (the differentiable configuration:
# Vreman c = 2.5 Cs^2 = 0.08, ETD2, checkpoints with recomputation) against
# the FP64 autograd reference and central finite differences (h = 1e-3 in Cs;
# in U0 a 1e-6 relative perturbation along a random direction in FP64, 1e-3
# along U0 itself in FP32).  FP64: the adjoint is the exact transpose of the FP64 step, so it
# meets autograd to 1e-8 and FD to its truncation error (1e-4).  FP32: the
# forward and adjoint run in FP32 with FP64 sums; both gradients are compared
# with the FP64 autograd reference to 2e-4 (FP32 round-off accumulated over
# the window) and with FD of the FP32 loss to 2e-3 (FD of an FP32 loss:
# rounding noise of ~1e-7 L over 2h).
ETD2_TOL = {
    "fp64": dict(loss=1e-10, grad_ref=1e-8, cs_ref=1e-8, cs_fd=1e-4, dir_fd=1e-5, eps=1e-6),
    "fp32": dict(loss=1e-5, grad_ref=2e-4, cs_ref=2e-4, cs_fd=2e-3, dir_fd=2e-3, eps=1e-3),
}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("N", [16, 24])
def test_etd2_vreman_adjoint_vs_autograd_and_fd(precision, N):
    """ETD2 + Vreman (the differentiable LES configuration at small N): 7
    steps with a checkpoint every 3 (three segments: two recomputed, the last
    kept), loss = squared log-mismatch of E(k) against a target spectrum;
    loss, dL/dU0 and dL/dCs against the FP64 torch-autograd reference and
    central finite differences."""
    from sles.adjoint import spectrum_loss

    tol = ETD2_TOL[precision]
    nu, dt, nsteps, kmax = 0.01, 0.005, 7, 6
    Cs = math.sqrt(0.08 / 2.5)
    U0 = _state(N, 51)
    target = TR.shell_spectrum(torch.as_tensor(_state(N, 52).cpu()), N, kmax=kmax)
    ref_loss, ref_gU, ref_gCs = _torch_reference_etd2_grads(U0, Cs, N, nu, dt, nsteps, "Vreman", target)

    s = _solver(N, "Vreman", scheme="ETD2", nsteps=nsteps, dt=dt, nu=nu, precision=precision, K=3)
    cd = s.d.cdtype
    U0c = U0.to(cd)
    U = U0c.clone().requires_grad_(True)
    c = torch.tensor(Cs, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = spectrum_loss(s(U, c), target.cuda(), N, kmax=kmax)
    loss.backward()
    gU = U.grad.detach().cpu().to(torch.complex128)
    gCs = float(c.grad)
    assert math.isclose(float(loss.detach()), ref_loss, rel_tol=tol["loss"])
    assert float((gU - ref_gU).abs().max()) < tol["grad_ref"] * float(ref_gU.abs().max())
    assert math.isclose(gCs, ref_gCs, rel_tol=tol["cs_ref"]), (gCs, ref_gCs)

    def L(U_, cv):
        with torch.no_grad():
            return float(spectrum_loss(s(U_, torch.tensor(cv, dtype=torch.float64)), target.cuda(), N, kmax=kmax))

    h = 1e-3
    fd = (L(U0c, Cs + h) - L(U0c, Cs - h)) / (2 * h)
    assert abs(gCs - fd) < tol["cs_fd"] * abs(fd), (gCs, fd)
    if precision == "fp64":
        g = torch.Generator(device="cuda").manual_seed(53)
        v = torch.complex(torch.randn(U0.shape, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(U0.shape, generator=g, device="cuda", dtype=torch.float64))
        v *= float(U0.abs().max()) * tol["eps"] / float(v.abs().max())
    else:
        # along U0 itself (d/dalpha L((1 + alpha) U0)): a random direction's
        # derivative is below the FP32 loss's rounding noise
        v = tol["eps"] * U0
    dd = (L(U0c + v.to(cd), Cs) - L(U0c - v.to(cd), Cs)) / 2
    assert abs(rdot(U.grad.to(torch.complex128), v) - dd) < tol["dir_fd"] * abs(dd), (rdot(U.grad, v), dd)


@pytest.mark.parametrize("scheme", ["RK4", "ETD2"])
def test_backward_uses_its_own_forward_cs(scheme):
    """Two forwards through one solver at different Cs, then the backward of
    the first: its gradients equal those of a forward + backward at that Cs
    alone (the backward recomputes and differentiates at the Cs its own
    forward ran with), and the solver's Cs is left as it was."""
    from sles.adjoint import DifferentiableLES, spectrum_loss

    N, nsteps, K = 16, 5, 2
    U0 = _state(N, 61)
    target = TR.shell_spectrum(torch.as_tensor(_state(N, 62).cpu()), N, kmax=6).cuda()

    def make():
        return DifferentiableLES(N=N, nu=0.01, dt=0.005, nsteps=nsteps, les_model="Vreman", scheme=scheme,
                                 precision="fp64", checkpoint_every=K)

    def graph(s, cs):
        U = U0.clone().requires_grad_(True)
        c = torch.tensor(cs, dtype=torch.float64, device="cuda", requires_grad=True)
        return U, c, spectrum_loss(s(U, c), target, N, kmax=6)

    s = make()
    U1, c1, loss1 = graph(s, 0.10)
    U2, c2, loss2 = graph(s, 0.25)  # a second forward at another Cs before the first backward
    loss1.backward()
    ref = make()
    Ur, cr, lossr = graph(ref, 0.10)
    lossr.backward()
    assert torch.equal(U1.grad, Ur.grad)
    assert float(c1.grad) == float(cr.grad)
    assert s.d.Cs == 0.25
    loss2.backward()
    ref2 = make()
    Ur2, cr2, lossr2 = graph(ref2, 0.25)
    lossr2.backward()
    assert torch.equal(U2.grad, Ur2.grad)
    assert float(c2.grad) == float(cr2.grad)
