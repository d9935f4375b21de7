"""Differentiable LES time integration: hand-written adjoint behind
``torch.autograd.Function``.

    solver = DifferentiableLES(N=256, nu=1e-4, dt=2e-3, nsteps=50, les_model="Vreman",
                               scheme="ETD2", precision="fp32", checkpoint_every=10)
    U_N = solver(U0, Cs)            # U0: (3, N, N, N/2+1) complex, Cs: 0-dim tensor
    loss(U_N).backward()            # -> U0.grad, Cs.grad

The forward pass runs the fused time steps (``sles_step``) and keeps the
state every ``checkpoint_every`` steps.  The backward pass walks the
segments in reverse: it recomputes the segment from its checkpoint (RK4:
the four stage states of every step; ETD2: the state and nonlinear term of
every step) and applies the transposed step, whose core is the hand-written
vector-Jacobian product of the RHS (``sles_adjoint_rhs``, one fused
five-pass launch sequence).  Gradients follow torch's convention for complex
tensors (d/dRe + i d/dIm); d/dCs is accumulated in FP64 on the device.

RK4 (src/RK4_method.py:20-35), for Y_s the stage states, k_s = f(Y_s):
    mu_3 = dt a_3 lam;  lY_3 = J_3^T mu_3
    mu_s = dt a_s lam + dt b_s lY_{s+1};  lY_s = J_s^T mu_s   (s = 2, 1, 0)
    lam_n = lam + sum_s lY_s
ETD2 (Cox-Matthews, U_{n+1} = E U_n + w1 F(U_n) - w0 F(U_{n-1})):
    lam_n = E lam_{n+1} + J_F(U_n)^T (w1 lam_{n+1} - w0 lam_{n+2})
with the first step an RK4 step whose F(U_0) also feeds the second step.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _lib, fused

RK4_A = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)
RK4_B = (0.5, 0.5, 1.0)


class DifferentiableLES:
    def __init__(self, N, nu, dt, nsteps, les_model="Smagorinsky", scheme="RK4", precision="fp32",
                 checkpoint_every=10, forcing=0.0, device="cuda", keep_last_segment=True):
        from .initialize_domain import InitializeDomain

        if scheme not in ("RK4", "ETD2"):
            raise ValueError("scheme must be 'RK4' or 'ETD2'")
        kw = dict(LES_ON=1, LES_model=les_model, Cs=0.1) if les_model else {}
        self.d = InitializeDomain(N=N, nu=nu, time_stepSize=dt, method=scheme, forcing=forcing,
                                  precision=precision, device=device, **kw)
        self.N, self.nu, self.dt, self.nsteps = N, nu, dt, int(nsteps)
        self.scheme = scheme
        self.K = max(1, int(checkpoint_every))
        # the forward pass keeps what the backward sweep needs for the last
        # segment (RK4: 12 fields per step, ETD2: 3), so it is not recomputed
        self.keep_last_segment = bool(keep_last_segment)
        self.les_model = les_model
        self.adj_work = torch.empty(int(_lib.lib.sles_adjoint_work_bytes(self.d.plan.handle, self.d.sgs_model)),
                                    dtype=torch.uint8, device=self.d.device)
        self.cs_grad = torch.zeros(1, dtype=torch.float64, device=self.d.device)

    # -- coefficient and its Cs derivative -------------------------------------
    def _set_cs(self, Cs):
        self.d.Cs = float(Cs)

    def dcoeff(self):
        d = self.d
        if not d.LES_ON:
            return 0.0
        if d.LES_model[0] == "S":
            return 2.0 * d.Cs * d.dx * d.dx
        return 5.0 * d.Cs * d.dx * d.dx

    # -- primitives -------------------------------------------------------------
    def rhs(self, Y, out):
        d = self.d
        d.U_hat.copy_(Y)
        fused.rhs(d, out)

    def nonlinear(self, Y, out):
        d = self.d
        d.U_hat.copy_(Y)
        fused.nonlinear(d, out)

    def vjp(self, Y, mu, out, accumulate=False, linear=1):
        """out (+)= J(Y)^T mu; accumulates d/dCs into self.cs_grad."""
        d = self.d
        prm = fused.params(d, dt=self.dt)
        _lib.call("sles_adjoint_rhs", d.plan.handle, ctypes.byref(prm), Y.data_ptr(), mu.data_ptr(),
                  out.data_ptr(), 1 if accumulate else 0, int(linear), self.dcoeff(),
                  self.cs_grad.data_ptr() if d.LES_ON else None, self.adj_work.data_ptr(), d.stream)

    def vjp_combined(self, Y, a, x, b, z, out):
        """out = J(Y)^T (a x + b z) (z may be None): one fused call
        (sles_adjoint_rhs_combined); accumulates d/dCs like vjp."""
        d = self.d
        prm = fused.params(d, dt=self.dt)
        _lib.call("sles_adjoint_rhs_combined", d.plan.handle, ctypes.byref(prm), Y.data_ptr(), float(a), x.data_ptr(),
                  float(b), z.data_ptr() if z is not None else None, out.data_ptr(), self.dcoeff(),
                  self.cs_grad.data_ptr() if d.LES_ON else None, self.adj_work.data_ptr(), d.stream)

    def sum_fields(self, terms, out):
        """out = terms[0] + terms[1] + ... (3 spectral fields each)."""
        ptrs = (ctypes.c_void_p * len(terms))(*[t.data_ptr() for t in terms])
        _lib.call("sles_sum_fields", self.d.plan.handle, 3, len(terms), ptrs, out.data_ptr(), self.d.stream)

    def axpby(self, a, x, b, y):
        _lib.call("sles_axpby", self.d.plan.handle, 3, float(a), x.data_ptr(), float(b), y.data_ptr(), self.d.stream)

    # -- forward ------------------------------------------------------------------
    def _rk4_stages(self, U):
        """Stage states Y_0..Y_3 of one RK4 step from U, and the new state: one
        fused RK4 step (the forward pass's own kernels, so the recomputed
        states equal the forward ones bit for bit) that also writes Y_1..Y_3.
        U is overwritten with the new state (it is returned); Y_0 is the
        step's own copy of U (written by the step as U_hat0), so no field is
        copied."""
        d = self.d
        y0 = torch.empty_like(U)
        st = torch.empty((9,) + tuple(U.shape[1:]), dtype=U.dtype, device=U.device)
        prm = fused.params(d, dt=self.dt)
        _lib.call("sles_rk4_stage_states", d.plan.handle, ctypes.byref(prm), U.data_ptr(),
                  y0.data_ptr(), d.U_hat1.data_ptr(), st.data_ptr(), fused.work(d).data_ptr(), d.stream)
        return [y0, st[0:3], st[3:6], st[6:9]], U

    def run_forward(self, U0, keep=None):
        """Advance nsteps; returns the final state and the checkpoints
        {step: (U, F_prev)}.  With keep_last_segment, what the backward sweep
        needs for the last segment is kept as well (entry "tail": ETD2 the
        states [U_n, ...], RK4 the stage states [Y_0..Y_3] of every step,
        written by the same fused step kernels): the sweep starts there and
        needs no recomputation for it."""
        d = self.d
        ckpt = {}
        d.U_hat.copy_(U0)
        d.U_hat1.copy_(U0)
        d.firstIteration = True
        last = ((self.nsteps - 1) // self.K) * self.K
        tail = []
        for n in range(self.nsteps):
            if n % self.K == 0:
                ckpt[n] = (d.U_hat.clone(), d.F0.clone() if (self.scheme == "ETD2" and n > 0) else None)
            keep = self.keep_last_segment and n >= last
            if self.scheme == "RK4":
                if keep:
                    Ys, _ = self._rk4_stages(d.U_hat)  # advances d.U_hat in place
                    tail.append(Ys)
                else:
                    fused.rk4_step(d, physical=False)
            else:
                if keep:
                    tail.append(ckpt[n][0] if n == last else d.U_hat.clone())
                fused.etd2_step(d, physical=False)
        if tail:
            ckpt["tail"] = (last, tail)
        return d.U_hat.clone(), ckpt

    # -- backward -----------------------------------------------------------------
    def run_backward(self, ckpt, gU):
        lam = gU.clone()
        self.cs_grad.zero_()
        if self.scheme == "RK4":
            return self._backward_rk4(ckpt, lam)
        return self._backward_etd2(ckpt, lam)

    def _backward_rk4(self, ckpt, lam):
        tail_start, tail_stages = ckpt.get("tail", (None, None))
        for start in sorted((k for k in ckpt if k != "tail"), reverse=True):
            end = min(start + self.K, self.nsteps)
            if start == tail_start:
                stages = tail_stages  # kept by the forward pass
            else:
                U = ckpt[start][0].clone()
                stages = []
                for n in range(start, end):
                    Ys, U = self._rk4_stages(U)
                    stages.append(Ys)
            for Ys in reversed(stages):
                lam = self._rk4_step_adjoint(Ys, lam)
        return lam

    def _backward_etd2(self, ckpt, lam_final):
        """lam_n = E lam_{n+1} + J_F(U_n)^T (w1 lam_{n+1} - w0 lam_{n+2})."""
        d = self.d
        N = self.nsteps
        mu = torch.empty_like(lam_final)
        # adjoints of the states U_{n+1} and U_{n+2} during the sweep
        lam_next = lam_final.clone()  # lam_{n+1}
        lam_next2 = None  # lam_{n+2}
        tail_start, tail_states = ckpt.get("tail", (None, None))
        for start in sorted((k for k in ckpt if k != "tail"), reverse=True):
            end = min(start + self.K, N)
            if start == tail_start:
                states = tail_states  # kept by the forward pass
            else:
                # recompute U_start .. U_{end-1}
                U, Fp = ckpt[start]
                states = []
                d.U_hat.copy_(U)
                d.U_hat1.copy_(U)
                if start == 0:
                    d.firstIteration = True
                else:
                    d.F0.copy_(Fp)
                    d.firstIteration = False
                for n in range(start, end):
                    states.append(d.U_hat.clone())
                    fused.etd2_step(d, physical=False)
            for n in reversed(range(start, end)):
                Un = states[n - start]
                if n >= 1:
                    # lam_n = E lam_{n+1} + J_F(U_n)^T (w1 lam_{n+1} - w0 lam_{n+2}), one fused call
                    lam_out = torch.empty_like(lam_next)
                    prm = fused.params(d, dt=self.dt)
                    _lib.call("sles_etd2_adjoint_rhs", d.plan.handle, ctypes.byref(prm), Un.data_ptr(),
                              lam_next.data_ptr(), lam_next2.data_ptr() if lam_next2 is not None else None,
                              lam_out.data_ptr(), self.dcoeff(), self.cs_grad.data_ptr() if d.LES_ON else None,
                              self.adj_work.data_ptr(), d.stream)
                    lam_next2, lam_next = lam_next, lam_out
                else:
                    # step 0: RK4 bootstrap U_1 = RK4(U_0), plus F(U_0) in the step to U_2
                    Ys, _ = self._rk4_stages(Un.clone())
                    lam0 = self._rk4_step_adjoint(Ys, lam_next)
                    if lam_next2 is not None:
                        # U_2 = ... - w0 F(U_0): contribution J_F(U_0)^T (-w0 lam_2)
                        zero = torch.zeros_like(lam_next2)
                        tmp = torch.empty_like(lam_next2)
                        _lib.call("sles_etd2_adjoint_combine", d.plan.handle, float(self.dt), float(self.nu),
                                  zero.data_ptr(), lam_next2.data_ptr(), mu.data_ptr(), tmp.data_ptr(), d.stream)
                        self.vjp(Un, mu, lam0, accumulate=True, linear=2)
                    lam_next2, lam_next = lam_next, lam0
        return lam_next

    def lincomb(self, a, x, b, y, out):
        _lib.call("sles_lincomb", self.d.plan.handle, 3, float(a), x.data_ptr(), float(b), y.data_ptr(),
                  out.data_ptr(), self.d.stream)

    def _rk4_step_adjoint(self, Ys, lam):
        """lam_n from lam_{n+1} for one RK4 step with stage states Ys: four
        fused calls that form mu_s on the fly, then one summation pass."""
        lY = [None] * 4
        for s in (3, 2, 1, 0):
            lY[s] = torch.empty_like(lam)
            if s == 3:
                # lY_3 = J_3^T (dt a_3 lam)
                self.vjp_combined(Ys[3], self.dt * RK4_A[3], lam, 0.0, None, lY[3])
            else:
                # lY_s = J_s^T (dt b_s lY_{s+1} + dt a_s lam)
                self.vjp_combined(Ys[s], self.dt * RK4_B[s], lY[s + 1], self.dt * RK4_A[s], lam, lY[s])
        acc = torch.empty_like(lam)
        self.sum_fields([lam, lY[3], lY[2], lY[1], lY[0]], acc)  # lam + sum_s lY_s
        return acc

    def __call__(self, U0, Cs):
        return _Advance.apply(U0, Cs, self)


class _Advance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, Cs, solver):
        cs = float(Cs.item() if torch.is_tensor(Cs) else Cs)
        solver._set_cs(cs)
        out, ckpt = solver.run_forward(U0.detach().to(solver.d.cdtype))
        ctx.solver = solver
        ctx.ckpt = ckpt
        ctx.cs = cs  # this forward's coefficient: the solver may run other forwards before this backward
        ctx.cs_dtype = Cs.dtype if torch.is_tensor(Cs) else torch.float64
        ctx.cs_device = Cs.device if torch.is_tensor(Cs) else solver.d.device
        ctx.in_dtype = U0.dtype
        return out

    @staticmethod
    def backward(ctx, gU):
        solver = ctx.solver
        # recompute and differentiate at the Cs of this graph's forward pass
        # (finite-difference probes, line searches or a second loss may have
        # run the solver at another Cs since), then restore the solver's Cs
        prev = solver.d.Cs
        solver._set_cs(ctx.cs)
        try:
            lam = solver.run_backward(ctx.ckpt, gU.detach().to(solver.d.cdtype).contiguous())
            gCs = solver.cs_grad.clone().reshape(()).to(dtype=ctx.cs_dtype, device=ctx.cs_device)
        finally:
            solver._set_cs(prev)
        return lam.to(ctx.in_dtype), gCs, None


_SPEC = {}


def _spectrum_plan(N, dtype, device, kmax):
    """(plan, workspace) of the shell spectrum for (N, precision, device),
    cached; the workspace grows to the largest kmax asked for."""
    key = (N, dtype, str(device))
    plan, work = _SPEC.get(key, (None, None))
    if plan is None:
        prec = _lib.SLES_FP64 if dtype == torch.complex128 else _lib.SLES_FP32
        plan = _lib.Plan(N, prec, device)
    nbytes = int(_lib.lib.sles_spectrum_work_bytes(plan.handle, kmax))
    if work is None or work.numel() < nbytes:
        work = torch.empty(nbytes, dtype=torch.uint8, device=device)
    _SPEC[key] = (plan, work)
    return plan, work


class _ShellSpectrum(torch.autograd.Function):
    """E(k) over unit shells (sles_shell_spectrum) with its vector-Jacobian
    product (sles_shell_spectrum_vjp)."""

    @staticmethod
    def forward(ctx, U_hat, N, kmax):
        U = U_hat.detach().contiguous()
        plan, work = _spectrum_plan(N, U.dtype, U.device, kmax)
        E = torch.empty(kmax, dtype=torch.float64, device=U.device)
        _lib.call("sles_shell_spectrum", plan.handle, U.data_ptr(), kmax, E.data_ptr(), work.data_ptr(),
                  _lib.stream_handle(U.device))
        ctx.save_for_backward(U)
        ctx.N, ctx.kmax = N, kmax
        return E

    @staticmethod
    def backward(ctx, gE):
        (U,) = ctx.saved_tensors
        plan, _ = _spectrum_plan(ctx.N, U.dtype, U.device, ctx.kmax)
        g = gE.detach().to(torch.float64).contiguous()
        gU = torch.empty_like(U)
        _lib.call("sles_shell_spectrum_vjp", plan.handle, U.data_ptr(), ctx.kmax, g.data_ptr(), gU.data_ptr(),
                  _lib.stream_handle(U.device))
        return gU, None, None


def shell_spectrum(U_hat, N, kmax=40):
    """E(k) over unit shells |k| in [k - 1/2, k + 1/2), k = 1..kmax:
    sum (1/2) w |U_hat|^2 / N^6 with the rfft doubling w (1 on the kz = 0 and
    N/2 planes, 2 elsewhere), FP64, differentiable.  One deterministic
    binning kernel (segmented wavefront sums, no atomics) and its hand-written
    vector-Jacobian product; U_hat (3, N, N, N/2+1) complex on the GPU."""
    if not U_hat.is_cuda or U_hat.dtype not in (torch.complex64, torch.complex128):
        raise ValueError("shell_spectrum: expected a complex64/complex128 CUDA tensor (3, N, N, N/2+1)")
    if tuple(U_hat.shape) != (3, N, N, N // 2 + 1):
        raise ValueError(f"shell_spectrum: expected shape (3, {N}, {N}, {N // 2 + 1}), got {tuple(U_hat.shape)}")
    return _ShellSpectrum.apply(U_hat, int(N), int(kmax))


def spectrum_loss(U_hat, target_E, N, kmax=40):
    """Squared log-mismatch sum_{k <= kmax} (log E(k) - log E_target(k))^2."""
    E = shell_spectrum(U_hat, N, kmax)
    return ((torch.log(E) - torch.log(target_E[:kmax])) ** 2).sum()
