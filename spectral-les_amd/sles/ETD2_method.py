"""ETD2 time step on the GPU.

Two forms, chosen by ``domain.etd2_form``:

* ``"cox_matthews"`` (default, the product form): Cox & Matthews (2002) ETD2
  with L = -nu K^2,
      U_hat <- e^{hL} U_hat + h (phi1 + phi2)(hL) F_n - h phi2(hL) F_{n-1},
  phi functions in FP64 (Taylor series near 0), where F is the full
  nonlinear term of computeRHS *including* the SGS stress divergence.
  F_n lives in ``domain.F1`` and F_{n-1} in ``domain.F0``.
* ``"reference"``: src/ETD2_method.py:7-57 literally.  Its F1 coefficient
  (1 + h nu K^2)/(h (nu K^2)^2) is not the ETD2 one and grows without bound
  for small nu K^2, and its nonlinear term has no SGS part; it is kept so the
  defect can be reproduced and documented.

Both start with one RK4 step (src/ETD2_method.py:8-12).
"""

from __future__ import annotations

from . import _lib
from . import RK4_method as RK4
from .math_formula import U_cross_omega, del_cross_U
from .smagorinsky import Smagorinsky_model
from .vreman import Vreman_model

__all__ = ["solveForNextTimestep"]


def _refresh_physical(d):
    d.fft.backward(d.U_hat, d._U, normalize=True)
    del_cross_U(d.fft, d.dim, d.U_hat, d.curl, None)


def full_nonlinear(d, out):
    """out = computeRHS(...) without the viscous term, at the current U_hat,
    from the current U and curl: P(dealias(u x omega) + T) + forcing + T."""
    U_cross_omega(d.fft, d.dim, d.U, d.curl, out)
    T = None
    if d.LES_ON:
        model = Smagorinsky_model if d.LES_model[0] == "S" else Vreman_model
        model(d, out)
        T = d.temp_I
    _lib.call(
        "sles_rhs_finish", d.plan.handle, _lib.ptr_array([out[0], out[1], out[2]]),
        _lib.ptr_array([T[0], T[1], T[2]]) if T is not None else None,
        _lib.ptr_array([d.U_hat[0], d.U_hat[1], d.U_hat[2]]), d.P_hat.data_ptr(), None,
        0.0, float(d.forcing), d.stream)
    return out


def computeFn_for_ETD2(domain):
    """src/ETD2_method.py:39-46: F0 <- F1; F1 = P(dealias F(u x omega))."""
    d = domain
    d.F0.copy_(d.F1)
    U_cross_omega(d.fft, d.dim, d.U, d.curl, d.F1)
    _lib.call("sles_rhs_finish", d.plan.handle, _lib.ptr_array([d.F1[0], d.F1[1], d.F1[2]]), None,
              _lib.ptr_array([d.U_hat[0], d.U_hat[1], d.U_hat[2]]), d.P_hat.data_ptr(), None, 0.0, 0.0,
              d.stream)


def solveForNextTimestep(domain):
    d = domain
    form = getattr(d, "etd2_form", "cox_matthews")
    if form == "reference":
        if d.firstIteration:
            RK4.solveForNextTimestep(d)
            d.firstIteration = False
            return
        computeFn_for_ETD2(d)
        _lib.call("sles_etd2_update", d.plan.handle, 0, d.U_hat.data_ptr(), d.F1.data_ptr(), d.F0.data_ptr(),
                  float(d.dt), float(d.nu), d.stream)
        _refresh_physical(d)
        return
    if form != "cox_matthews":
        raise ValueError("etd2_form must be 'cox_matthews' or 'reference'")
    if getattr(d, "fused", False):
        from . import fused

        fused.etd2_step(d)
        return
    if d.firstIteration:
        full_nonlinear(d, d.F0)  # F(u_0), the F_{n-1} of the first ETD2 step
        d.U_hat1.copy_(d.U_hat)
        RK4.solveForNextTimestep(d)
        d.firstIteration = False
        return
    full_nonlinear(d, d.F1)
    _lib.call("sles_etd2_update", d.plan.handle, 1, d.U_hat.data_ptr(), d.F1.data_ptr(), d.F0.data_ptr(),
              float(d.dt), float(d.nu), d.stream)
    d.F0, d.F1 = d.F1, d.F0
    _refresh_physical(d)
