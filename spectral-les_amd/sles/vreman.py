"""Vreman eddy viscosity on the GPU (reference src/vreman.py:8-123)."""

from __future__ import annotations

from . import _lib
from .math_formula import _grad_fields, residue_stress, velocity_gradient

__all__ = ["Vreman_model"]


def Vreman_model(domain, dU):
    """velocity_gradient -> calculate_nuR -> residue_stress; returns temp_I
    (vreman.py:8-21)."""
    velocity_gradient(domain=domain)
    calculate_nuR(domain)
    domain.temp_I[:] = residue_stress(domain=domain)
    return domain.temp_I


def calculate_nuR(domain):
    """Vreman nu_R into domain.nuR and alpha:alpha into domain.aTemp
    (vreman.py:24-123), coefficient 2.5 Cs^2 dx^2."""
    _lib.call("sles_vreman_nu", domain.plan.handle, _lib.ptr_array(_grad_fields(domain)),
              domain.nuR.data_ptr(), domain.aTemp.data_ptr(),
              2.5 * domain.Cs * domain.Cs * domain.dx * domain.dx, domain.stream)
