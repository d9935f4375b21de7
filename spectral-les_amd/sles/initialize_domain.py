"""InitializeDomain: the state bag of the solver, on one MI355X.

Mirrors ``InitializeDomain`` of the reference (src/initialize_domain.py:19-186)
for the 3D branch on one rank: same constructor keywords and attribute names,
with every field a torch tensor on the GPU in the reference layout

    U        list of 3 (N, N, N) real views      (initialize_domain.py:66)
    curl     (3, N, N, N) real                   (:108)
    U_hat, dU, U_hat0, U_hat1, temp1_Img, F0, F1, temp_I
             (3, N, N, N/2+1) complex            (:67-75,175)
    P_hat    (N, N, N/2+1) complex               (:70)
    nuR, aTemp, a11 .. a33   (N, N, N) real      (:174-186)

``fft`` is a :class:`SpectralFFT` with the PFFT ``forward``/``backward``
call surface.  The wavenumber arrays ``K``, ``K2``, ``K_over_K2`` and
``dealias`` are built on first access only: no kernel reads them, every
kernel computes wavenumbers from indices.

Extra keywords (not in the reference): ``N`` gives the grid directly (the
reference uses N = No_of_gridpoint / 4 when LES is on, :38-40),
``precision`` = "fp64" | "fp32", ``device``, ``fused`` selects the fused
time-step kernels for ``solveForNextTimestep`` (default) or the
operator-by-operator path that reproduces every intermediate array, and
``etd2_form`` = "cox_matthews" (default) | "reference" picks the ETD2 update.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .fft import SpectralFFT, Workspace

__all__ = ["InitializeDomain", "SingleRankComm"]


class SingleRankComm:
    """Stand-in for MPI.COMM_WORLD on one rank (initialize_domain.py:51-53)."""

    rank = 0
    size = 1

    def Get_size(self):
        return 1

    def Get_rank(self):
        return 0


_GRAD = ("a11", "a12", "a13", "a21", "a22", "a23", "a31", "a32", "a33")


class InitializeDomain:
    def __init__(
        self,
        nu: float = 0.01,
        final_time: float = 0.0,
        dimension: int = 3,
        time_stepSize: float = 0.01,
        No_of_gridpoint: int = 128,
        forcing=0.0,
        forcing_range=(0, 1),
        method="ETD2",
        LES_ON=0,
        CFL: float = 0.5,
        Cs: float = 0.14,
        LES_model=None,
        *,
        N=None,
        precision: str = "fp64",
        device=None,
        fused: bool = True,
        etd2_form: str = "cox_matthews",
        tuned: bool = True,
    ):
        if dimension != 3:
            raise NotImplementedError("only the 3D solver is implemented (the 2D branch is out of scope)")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if not torch.cuda.is_available():
            raise RuntimeError("InitializeDomain needs a ROCm GPU: the HIP kernels are the only compute path")
        self.nu = nu
        self.CFL = CFL
        self.T = final_time
        self.dt = time_stepSize
        self.LES_ON = True if LES_ON in (1, True) else False
        self.NN = 4 if self.LES_ON else 1
        self.N = int(N) if N is not None else int(No_of_gridpoint / self.NN)
        self.N_full = (No_of_gridpoint if N is None else 4 * self.N) if self.LES_ON else None
        self.dx = 2 * math.pi / self.N
        self.method = method
        self.firstIteration = True
        self.forcing = forcing
        self.forcing_range = forcing_range
        self.comm = SingleRankComm()
        self.num_processes = 1
        self.rank = 0
        self.Np = self.N
        self.dim = 3
        self.shape = [self.N] * 3
        self.precision = precision
        self.fused = fused
        if etd2_form not in ("cox_matthews", "reference"):
            raise ValueError("etd2_form must be 'cox_matthews' or 'reference'")
        self.etd2_form = etd2_form
        self.device = torch.device(device if device is not None else "cuda")
        self.rdtype = torch.float64 if precision == "fp64" else torch.float32
        self.cdtype = torch.complex128 if precision == "fp64" else torch.complex64
        N = self.N
        Nh = N // 2 + 1
        self.Nh = Nh
        prec = _lib.SLES_FP64 if precision == "fp64" else _lib.SLES_FP32
        self.plan = _lib.Plan(N, prec, self.device)
        self.ws = Workspace(N, self.rdtype, self.cdtype, self.device)
        self.fft = SpectralFFT(self.plan, self.ws, self.rdtype, self.cdtype, self.device)

        def rz(*s):
            return torch.zeros(*s, dtype=self.rdtype, device=self.device)

        def cz(*s):
            return torch.zeros(*s, dtype=self.cdtype, device=self.device)

        self._U = rz(3, N, N, N)
        self.U = [self._U[0], self._U[1], self._U[2]]
        self.U_hat = cz(3, N, N, Nh)
        self.temp1_Img = cz(3, N, N, Nh)
        self.P = rz(N, N, N)
        self.P_hat = cz(N, N, Nh)
        self.dU = cz(3, N, N, Nh)
        self.U_hat0 = cz(3, N, N, Nh)
        self.U_hat1 = cz(3, N, N, Nh)
        self.F0 = cz(3, N, N, Nh) if method == "ETD2" else None
        self.F1 = cz(3, N, N, Nh) if method == "ETD2" else None
        self.kx = np.fft.fftfreq(N, 1.0 / N)
        self.kz = self.kx[: (N // 2 + 1)].copy()
        self.kz[-1] *= -1
        self.kmax_dealias = 2.0 / 3.0 * (N // 2 + 1)
        self.w = rz(N, N, N)
        self.curl = rz(3, N, N, N)
        self.a = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]
        self.b = [0.5, 0.5, 1.0]
        self._K = None
        self._dealias = None
        self.Cs = Cs
        if self.LES_ON:
            if LES_model is None:
                raise ValueError("Please choose LES_model (Smagorinsky or Vreman)")
            self.LES_model = LES_model.upper()
            self.nuR = rz(N, N, N)
            self.temp_I = cz(3, N, N, Nh)
            self._grad = rz(9, N, N, N)
            for i, name in enumerate(_GRAD):
                setattr(self, name, self._grad[i])
            self.aTemp = rz(N, N, N)
        else:
            self.LES_model = None
        # kernel variants chosen by the autotuner for this (N, precision, model)
        self.tuning = None
        if tuned:
            from . import autotune

            self.tuning = autotune.apply(self.plan, N, precision, self.sgs_model)
        # device scalars of the fused path (u_max, dt) and reduction scratch
        self._scalars = torch.zeros(8, dtype=torch.float64, device=self.device)
        self._red_work = torch.empty(int(_lib.lib.sles_reduce_work_bytes()) // 8, dtype=torch.float64,
                                     device=self.device)

    # ------------------------------------------------------------------
    # wavenumber arrays (initialize_domain.py:99-114,162-165), built lazily
    # ------------------------------------------------------------------
    @property
    def K(self):
        if self._K is None:
            kx = torch.as_tensor(self.kx, dtype=torch.int64, device=self.device)
            kz = torch.as_tensor(self.kz, dtype=torch.int64, device=self.device)
            self._K = torch.stack(torch.meshgrid(kx, kx, kz, indexing="ij"))
        return self._K

    @property
    def K2(self):
        K = self.K
        return (K * K).sum(0)

    @property
    def K_over_K2(self):
        K2 = self.K2
        return self.K.to(torch.float64) / torch.where(K2 == 0, 1, K2).to(torch.float64)

    @property
    def dealias(self):
        if self._dealias is None:
            K = self.K.abs().to(torch.float64)
            m = self.kmax_dealias
            self._dealias = (K[0] < m) & (K[1] < m) & (K[2] < m)
        return self._dealias

    # ------------------------------------------------------------------
    @property
    def sgs_model(self) -> int:
        if not self.LES_ON:
            return _lib.SLES_LES_NONE
        return _lib.SLES_LES_SMAGORINSKY if self.LES_model[0] == "S" else _lib.SLES_LES_VREMAN

    def sgs_coeff(self) -> float:
        """(Cs dx)^2 for Smagorinsky (smagorinsky.py:39), 2.5 Cs^2 dx^2 for
        Vreman (vreman.py:123)."""
        if not self.LES_ON:
            return 0.0
        if self.LES_model[0] == "S":
            return (self.Cs * self.dx) ** 2
        return 2.5 * self.Cs * self.Cs * self.dx * self.dx

    @property
    def stream(self) -> int:
        return _lib.stream_handle(self.device)


_PLAN_CACHE = {}


def _plan_for(field):
    """(plan, reduction scratch, FP64 output scalar) for a standalone (N,N,N)
    real field, cached per (N, dtype, device)."""
    n = int(field.shape[-1])
    if field.dtype not in (torch.float64, torch.float32) or tuple(field.shape[-3:]) != (n, n, n):
        raise ValueError("expected (N, N, N) float64/float32 fields")
    key = (n, field.dtype, str(field.device))
    hit = _PLAN_CACHE.get(key)
    if hit is None:
        prec = _lib.SLES_FP64 if field.dtype == torch.float64 else _lib.SLES_FP32
        plan = _lib.Plan(n, prec, field.device)
        work = torch.empty(int(_lib.lib.sles_reduce_work_bytes()) // 8, dtype=torch.float64, device=field.device)
        out = torch.zeros(1, dtype=torch.float64, device=field.device)
        hit = (plan, work, out)
        _PLAN_CACHE[key] = hit
    return hit
