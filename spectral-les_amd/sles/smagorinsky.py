"""Smagorinsky eddy viscosity on the GPU (reference src/smagorinsky.py:8-47)."""

from __future__ import annotations

from . import _lib
from .math_formula import _grad_fields, residue_stress, velocity_gradient

__all__ = ["Smagorinsky_model"]


def Smagorinsky_model(domain, dU):
    """velocity_gradient -> calculate_nuR -> residue_stress; returns temp_I
    (smagorinsky.py:8-21; ``dU`` is unused there too)."""
    velocity_gradient(domain=domain)
    calculate_nuR(domain)
    domain.temp_I[:] = residue_stress(domain=domain)
    return domain.temp_I


def calculate_nuR(domain):
    """nu_R = sqrt(2 S:S) (Cs dx)^2 into domain.nuR (smagorinsky.py:24-39)."""
    _lib.call("sles_smagorinsky_nu", domain.plan.handle, _lib.ptr_array(_grad_fields(domain)),
              domain.nuR.data_ptr(), (domain.Cs * domain.dx) ** 2, domain.stream)
