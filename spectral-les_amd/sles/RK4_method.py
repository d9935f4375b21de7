"""Classical RK4 time step and the right-hand side, on the GPU.

Call surface of the reference's src/RK4_method.py: ``solveForNextTimestep``,
``computeRHS``, ``U_square``, ``dissipation_rate``.

``solveForNextTimestep`` runs the fused kernels (one RK stage = five
transform passes with all pointwise work folded in) when ``domain.fused`` is
set, otherwise the operator-by-operator sequence of the reference, which also
leaves every intermediate array (dU, P_hat, temp_I, nuR, a11..a33, ...) as
the reference does.  Both require U_hat1 == U_hat on entry
(src/SM_solver.py:33), and both leave U_hat, U_hat1, U and curl as the
reference does.
"""

from __future__ import annotations

import torch

from . import _lib
from .math_formula import U_cross_omega, del_cross_U
from .smagorinsky import Smagorinsky_model
from .vreman import Vreman_model

__all__ = ["solveForNextTimestep", "computeRHS", "U_square", "dissipation_rate"]


def _fused_available(domain):
    from . import fused

    return fused.supported(domain)


def solveForNextTimestep(domain):
    """One RK4 step (RK4_method.py:20-35)."""
    if getattr(domain, "fused", False) and _fused_available(domain):
        from . import fused

        fused.rk4_step(domain)
        return
    _rk4_step_ops(domain)


def _rk4_step_ops(domain):
    d = domain
    s = d.stream
    d.U_hat0.copy_(d.U_hat)
    for rk in range(4):
        computeRHS(d, d.dU, rk)
        c0 = d.b[rk] * d.dt if rk < 3 else 0.0
        _lib.call("sles_stage_update", d.plan.handle, 0, d.U_hat.data_ptr(), d.U_hat0.data_ptr(),
                  d.U_hat1.data_ptr(), d.dU.data_ptr(), c0, d.a[rk] * d.dt, s)
    d.U_hat.copy_(d.U_hat1)
    del_cross_U(d.fft, d.dim, d.U_hat, d.curl, None)
    d.fft.backward(d.U_hat, d._U, normalize=True)


def computeRHS(domain, dU, rk):
    """RK4_method.py:38-71: u x omega, dealias, SGS stress, projection,
    forcing, ETD2 capture of F1, viscous term."""
    d = domain
    if rk > 0:
        d.fft.backward(d.U_hat, d._U, normalize=True)
        del_cross_U(d.fft, d.dim, d.U_hat, d.curl, None)
    U_cross_omega(d.fft, d.dim, d.U, d.curl, dU)
    T = None
    if d.LES_ON:
        model = Smagorinsky_model if d.LES_model[0] == "S" else Vreman_model
        model(d, dU)
        T = d.temp_I
    f1 = d.F1 if (d.method == "ETD2" and d.F1 is not None) else None
    _lib.call(
        "sles_rhs_finish",
        d.plan.handle,
        _lib.ptr_array([dU[0], dU[1], dU[2]]),
        _lib.ptr_array([T[0], T[1], T[2]]) if T is not None else None,
        _lib.ptr_array([d.U_hat[0], d.U_hat[1], d.U_hat[2]]),
        d.P_hat.data_ptr(),
        _lib.ptr_array([f1[0], f1[1], f1[2]]) if f1 is not None else None,
        float(d.nu),
        float(d.forcing),
        d.stream,
    )
    return dU


def _reduce(domain, name, *args):
    out = domain._scalars[0:1]
    _lib.call(name, domain.plan.handle, *args, out.data_ptr(), domain._red_work.data_ptr(), domain.stream)
    return float(out.item())


def U_square(domain):
    """sum_i sum_x u_i^2 / N^3 in FP64 (RK4_method.py:77-84)."""
    d = domain
    return _reduce(d, "sles_sum_squares", _lib.ptr_array(d.U), 3) / d.N ** d.dim


def dissipation_rate(domain):
    """nu * sum_i sum_x u_i F^-1(K^2 U_hat_i), FP64 sum (RK4_method.py:87-101)."""
    d = domain
    tmp = d.ws.cplx("k2_uhat", 3)
    _lib.call("sles_mul_k2", d.plan.handle, d.U_hat.data_ptr(), tmp.data_ptr(), 3, d.stream)
    lap = d.ws.real("k2_u", 3)
    d.fft.backward_destroy(tmp, lap, 3, 1.0 / float(d.N ** 3))
    # reference: eps += nu * sum(U_i * ...) per component
    eps = 0.0
    for i in range(3):
        eps += d.nu * _reduce(d, "sles_dot", _lib.ptr_array([d.U[i]]), _lib.ptr_array([lap[i]]), 1)
    return eps
