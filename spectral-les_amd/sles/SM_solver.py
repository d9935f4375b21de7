"""Driver-side reductions and the time loop (reference src/SM_solver.py).

* ``find_max_velocity`` / ``compute_dt`` (SM_solver.py:245-277): the CFL rule
  on a device FP64 max reduction (signed max over all components, as
  ``numpy.max`` of the list U).
* ``Ek_plot_variables`` / ``calculate_Ek`` (SM_solver.py:91-117,154-192): the
  binned mean spectrum from one device pass with FP64 bin sums.
* ``spectralMethodSolver`` (SM_solver.py:18-83): the time loop with the same
  dt rule, stop conditions and NaN guard.  HDF5 dumps and plotting are not
  part of this package; E(k) is appended to ``Ek_plot_data.csv`` with the
  reference's ``k-{t}`` / ``Ek-{t}`` columns when ``No_Ek_plot > 0``.
"""

from __future__ import annotations

import math
import os
import time

import numpy as np
import torch

from . import _lib
from . import ETD2_method as ETD2
from . import RK3_method as RK3
from . import RK4_method as RK4

__all__ = [
    "spectralMethodSolver",
    "find_max_velocity",
    "compute_dt",
    "Ek_plot_variables",
    "calculate_Ek",
    "DtTooSmall",
]


class DtTooSmall(SystemExit):
    """Raised where the reference calls exit() for dt < 1e-6 (SM_solver.py:273-276)."""


def _field_list(U):
    if torch.is_tensor(U) and U.dim() == 4:
        return [U[i] for i in range(U.shape[0])]
    return list(U)


def find_max_velocity(U, comm=None):
    """Signed max over every component of U (SM_solver.py:245-255), FP64."""
    fields = _field_list(U)
    dev = fields[0].device
    from .initialize_domain import _plan_for

    plan, work, out = _plan_for(fields[0])
    _lib.call("sles_max", plan.handle, _lib.ptr_array(fields), len(fields), out.data_ptr(), work.data_ptr(),
              _lib.stream_handle(dev))
    return float(out.item())


def compute_dt(U, u_max, nu, comm=None, dx=None, CFL=0.5):
    """dt = min(CFL dx / u_max, 0.1 dx^2 / nu) (SM_solver.py:258-277)."""
    dt1 = CFL * dx / u_max if u_max != 0 else math.copysign(math.inf, CFL * dx)
    dt2 = 0.1 * dx ** 2 / nu if (nu > 0) else 1e6
    dt = min(dt1, dt2)
    if dt < 1e-6:
        raise DtTooSmall(f"Time step (dt = {dt:g}) is smaller than 1e-6")
    return dt


class Ek_plot_variables:
    """Bin bookkeeping of the E(k) spectrum (SM_solver.py:91-117)."""

    def __init__(self, domain):
        N = domain.N
        kx = domain.kx
        kz = domain.kz
        k2max = int(np.max(kx ** 2)) * 2 + int(np.max(kz ** 2))
        self.avgPara = 5
        self.smoothing_para = round(math.sqrt(k2max)) // self.avgPara
        nb = self.smoothing_para + 1
        self.temp_k = np.zeros(nb, dtype=float)
        self.temp_Ek = np.zeros(nb, dtype=float)
        self._dev = torch.zeros(3, nb, dtype=torch.float64, device=domain.device)


def calculate_Ek(domain, E):
    """temp_k[i] = mean |k| and temp_Ek[i] = mean Ek over modes with K^2 in
    [(5i)^2, (5i+5)^2), Ek = sum_i |U_hat_i|^2 doubled for kz >= 1
    (SM_solver.py:154-192)."""
    nb = E.smoothing_para + 1
    dev = E._dev
    uh = domain.U_hat
    if not uh.is_contiguous():
        uh = uh.contiguous()
    _lib.call("sles_spectrum", domain.plan.handle, uh.data_ptr(), nb, dev[0].data_ptr(), dev[1].data_ptr(),
              dev[2].data_ptr(), domain.stream)
    ek, ks, cnt = dev.cpu().numpy()
    E.temp_k[:] = 0
    E.temp_Ek[:] = 0
    has = cnt > 0
    E.temp_k[has] = ks[has] / cnt[has]
    E.temp_Ek[has] = ek[has] / cnt[has]


def _write_ek_csv(E, t, path="Ek_plot_data.csv"):
    import csv

    cols = {}
    if os.path.exists(path):
        with open(path, newline="") as f:
            rows = list(csv.reader(f))
        if rows:
            header = rows[0]
            for j, h in enumerate(header):
                cols[h] = [r[j] for r in rows[1:]]
    cols[f"k-{t:g}"] = [repr(float(x)) for x in E.temp_k]
    cols[f"Ek-{t:g}"] = [repr(float(x)) for x in E.temp_Ek]
    names = list(cols)
    n = max(len(v) for v in cols.values())
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(names)
        for i in range(n):
            w.writerow([cols[c][i] if i < len(cols[c]) else "" for c in names])


def _step(domain):
    if domain.method == "RK4":
        RK4.solveForNextTimestep(domain)
    elif domain.method == "ETD2":
        ETD2.solveForNextTimestep(domain)
    elif domain.method == "RK3":
        RK3.solveForNextTimestep(domain)
    else:
        raise ValueError("Invalid method name")


def spectralMethodSolver(domain, file_out_time: float = 0.0, No_Ek_plot: int = 5, verbose: bool = False,
                         max_steps: int | None = None):
    """Time loop of SM_solver.py:18-83 (no HDF5 snapshots).

    Returns the number of steps taken.  ``max_steps`` (not in the reference)
    stops after that many steps."""
    u_max = find_max_velocity(domain.U, domain.comm)
    domain.dt = compute_dt(domain.U, u_max, domain.nu, domain.comm, dx=domain.dx, CFL=0.3)
    t = domain.dt
    Ek_plot_interval = domain.T / No_Ek_plot if No_Ek_plot != 0 else 0
    Ek_plot_counter = 0
    start = time.time()
    domain.U_hat1.copy_(domain.U_hat)
    Ek_var = Ek_plot_variables(domain) if Ek_plot_interval > 0 else None
    steps = 0
    while True:
        if Ek_plot_interval > 0 and t >= Ek_plot_counter:
            calculate_Ek(domain, Ek_var)
            _write_ek_csv(Ek_var, t)
            Ek_plot_counter += Ek_plot_interval
        if (t >= domain.T) or (domain.dt < 1e-8):
            break
        if max_steps is not None and steps >= max_steps:
            break
        if verbose:
            print(f"t = {t:<10g} dt = {domain.dt:<10g} max u = {u_max:g}", flush=True)
        t += domain.dt
        _step(domain)
        u_max = find_max_velocity(domain.U, domain.comm)
        domain.dt = compute_dt(domain.U, u_max, domain.nu, domain.comm, dx=domain.dx, CFL=domain.CFL)
        if domain.dt != domain.dt:
            break
        steps += 1
    total = time.time() - start
    if verbose and steps:
        print(f"Time per gridpoint per timestep: {total / (domain.N ** 3 * steps)}", flush=True)
    return steps
