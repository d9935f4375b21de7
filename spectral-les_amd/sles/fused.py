"""Fused time steps: one libsles ``sles_step`` call per step.

Every RHS evaluation of computeRHS (src/RK4_method.py:38-71) runs as five
transform passes with the pointwise work fused in; the stage updates of
RK4 (src/RK4_method.py:20-35), the reference RK3 (src/RK3_method.py:20-41)
and Cox-Matthews ETD2 are fused into the last pass.  The step leaves
U_hat, U_hat1 (RK4), U_hat0, U and curl as the reference leaves them;
intermediate arrays (dU, P_hat, temp_I, nuR, a11..a33) are not written --
use ``domain.fused = False`` for the operator-by-operator path that keeps
them.
"""

from __future__ import annotations

import ctypes

from . import _lib

SCHEME_RK4 = 0
SCHEME_RK3 = 1
SCHEME_ETD2 = 2
SCHEME_RHS = 3
SCHEME_NONLINEAR = 4


class StepParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("forcing", ctypes.c_double),
        ("sgs_coeff", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("les_model", ctypes.c_int),
        ("dt_device", ctypes.c_int),
    ]


def supported(domain) -> bool:
    return True


def work(domain):
    nbytes = int(_lib.lib.sles_step_work_bytes(domain.plan.handle, domain.sgs_model))
    return domain.ws._raw("fused_work", nbytes)


def params(domain, dt=None, dt_device=False, cfl=None) -> StepParams:
    p = StepParams()
    p.dt = float(domain.dt if dt is None else dt)
    p.nu = float(domain.nu)
    p.forcing = float(domain.forcing)
    p.sgs_coeff = float(domain.sgs_coeff())
    p.cfl = float(domain.CFL if cfl is None else cfl)
    p.dx = float(domain.dx)
    p.les_model = int(domain.sgs_model)
    p.dt_device = 1 if dt_device else 0
    return p


def step(domain, scheme, aux=None, dt_device=False, cfl=None, physical=True, dt=None):
    d = domain
    prm = params(d, dt=dt, dt_device=dt_device, cfl=cfl)
    w = work(d)
    _lib.call(
        "sles_step",
        d.plan.handle,
        scheme,
        ctypes.byref(prm),
        d.U_hat.data_ptr(),
        d.U_hat0.data_ptr(),
        d.U_hat1.data_ptr(),
        aux.data_ptr() if aux is not None else None,
        d._U.data_ptr() if physical else None,
        d.curl.data_ptr() if physical else None,
        d._scalars.data_ptr(),
        w.data_ptr(),
        d.stream,
    )


def rk4_step(domain, **kw):
    step(domain, SCHEME_RK4, **kw)


def rk3_step(domain, **kw):
    step(domain, SCHEME_RK3, **kw)


def etd2_step(domain, **kw):
    """Cox-Matthews ETD2; the first call is an RK4 step that also stores
    F(u_0) in F0 (the F_{n-1} of the next call)."""
    d = domain
    if d.firstIteration:
        d.U_hat1.copy_(d.U_hat)
        step(d, SCHEME_RK4, aux=d.F0, **kw)
        d.firstIteration = False
        return
    step(d, SCHEME_ETD2, aux=d.F0, **kw)


def rhs(domain, out):
    """out = computeRHS(U_hat) (full right-hand side), no side effects."""
    step(domain, SCHEME_RHS, aux=out, physical=False)


def nonlinear(domain, out):
    """out = computeRHS(U_hat) + nu K^2 U_hat (everything but the viscous term)."""
    step(domain, SCHEME_NONLINEAR, aux=out, physical=False)
