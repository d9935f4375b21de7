"""sles: the RHS and time-step hot path of the pseudo-spectral LES solver
a-k-sangwan/spectral-LES, on MI355X (gfx950) HIP kernels.

Modules mirror the reference package ``src``:

    initialize_domain  InitializeDomain
    math_formula       U_cross_omega, del_cross_U, velocity_gradient, residue_stress
    smagorinsky        Smagorinsky_model (calculate_nuR)
    vreman             Vreman_model (calculate_nuR)
    RK4_method         solveForNextTimestep, computeRHS, U_square, dissipation_rate
    RK3_method         solveForNextTimestep (as written in the reference)
    ETD2_method        solveForNextTimestep (Cox-Matthews form; reference form on request)
    initial            initial_velocity
    SM_solver          spectralMethodSolver, find_max_velocity, compute_dt, calculate_Ek
    utilites           get_index

All arithmetic runs in libsles.so; importing the package fails if it is missing.
"""

from . import _lib  # noqa: F401  (loads libsles.so, raises if absent)

__version__ = "0.1.0"
