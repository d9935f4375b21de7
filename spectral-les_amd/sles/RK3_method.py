"""The reference's RK3 step on the GPU (src/RK3_method.py:20-41), as written.

As written the scheme overwrites k1 and its linear amplification factor is
1 + z + z^2 + z^3/2 (first order); it is reproduced, not corrected.
``computeRHS``, ``U_square`` and ``dissipation_rate`` are the RK4 module's
(src/RK3_method.py:44-107 are verbatim copies of them).
"""

from __future__ import annotations

from . import _lib
from .RK4_method import computeRHS, dissipation_rate, U_square  # noqa: F401
from .math_formula import del_cross_U

__all__ = ["solveForNextTimestep", "computeRHS", "U_square", "dissipation_rate"]


def solveForNextTimestep(domain):
    d = domain
    if getattr(d, "fused", False):
        from . import fused

        fused.rk3_step(d)
        return
    dt = d.dt
    s = d.stream
    d.U_hat0.copy_(d.U_hat)
    computeRHS(d, d.dU, rk=0)
    _lib.call("sles_stage_update", d.plan.handle, 1, d.U_hat.data_ptr(), d.U_hat0.data_ptr(), None,
              d.dU.data_ptr(), 0.5 * dt, 0.0, s)
    computeRHS(d, d.dU, rk=1)
    _lib.call("sles_stage_update", d.plan.handle, 2, d.U_hat.data_ptr(), d.U_hat0.data_ptr(), None,
              d.dU.data_ptr(), dt, 0.0, s)
    computeRHS(d, d.dU, rk=2)
    _lib.call("sles_stage_update", d.plan.handle, 3, d.U_hat.data_ptr(), d.U_hat0.data_ptr(), None,
              d.dU.data_ptr(), dt, 0.0, s)
    del_cross_U(d.fft, d.dim, d.U_hat, d.curl, None)
    d.fft.backward(d.U_hat, d._U, normalize=True)
