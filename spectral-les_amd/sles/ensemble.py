"""Ensembles of uneven grids advanced together under one captured HIP graph.

Each member is a full LES domain (its own grid N, seed, Cs and CFL dt,
computed on the device every step, so no host synchronisation is needed
between steps).  One ensemble step = one RK4 step of every member; the
members' ``sles_step`` launch sequences are recorded once, each member on its
own stream forked from the capture stream, into a HIP graph
(``torch.cuda.CUDAGraph``), and the graph is replayed per step.  Members of
different sizes therefore overlap on the GPU, and the launch cost of the
~20 kernels per member-step is paid once per graph replay.  The kernel
variants of each member come from the autotuner table (InitializeDomain
applies it).

    ens = Ensemble([(64, 8), (96, 8), (128, 8)], precision="fp32", seed=0)
    ens.capture(steps_per_graph=1)
    rate = ens.run(200)     # grid points x steps / s over all members
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import fused
from .initial import isotropic_spectrum, set_spectral_state
from .initialize_domain import InitializeDomain
from .SM_solver import compute_dt, find_max_velocity


class Ensemble:
    def __init__(self, groups=((64, 8), (96, 8), (128, 8)), nu=1e-4, cs_range=(0.10, 0.20), seed=0,
                 precision="fp32", cfl=0.5, k0=4.0, les_model="Smagorinsky"):
        rng = np.random.default_rng(seed)
        self.members = []
        self.cfl = cfl
        i = 0
        for N, count in groups:
            for _ in range(count):
                Cs = float(rng.uniform(*cs_range))
                d = InitializeDomain(N=N, nu=nu, method="RK4", LES_ON=1, LES_model=les_model, Cs=Cs, CFL=cfl,
                                     precision=precision)
                set_spectral_state(d, isotropic_spectrum(N, k0=k0, seed=seed * 1000 + i, urms=1.0, dtype=d.cdtype))
                # first dt with CFL 0.3 on the host (src/SM_solver.py:21-23); later ones on the device
                d.dt = compute_dt(d.U, find_max_velocity(d.U), nu, dx=d.dx, CFL=0.3)
                self.members.append(d)
                i += 1
        self.streams = [torch.cuda.Stream() for _ in self.members]
        self.graph = None
        self.steps_per_graph = 1
        self.first_done = False

    @property
    def points(self):
        return sum(d.N ** 3 for d in self.members)

    def _first_step(self):
        for d in self.members:
            fused.rk4_step(d, dt=d.dt, physical=False)
        self.first_done = True

    def _record(self, steps):
        cur = torch.cuda.current_stream()
        for _ in range(steps):
            for d, s in zip(self.members, self.streams):
                s.wait_stream(cur)
                with torch.cuda.stream(s):
                    fused.rk4_step(d, dt_device=True, cfl=self.cfl, physical=False)
            for s in self.streams:
                cur.wait_stream(s)

    def step_eager(self, steps=1):
        if not self.first_done:
            self._first_step()
        self._record(steps)

    def capture(self, steps_per_graph=1):
        if not self.first_done:
            self._first_step()
        # warm the workspaces outside the capture
        self._record(1)
        torch.cuda.synchronize()
        self.steps_per_graph = steps_per_graph
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._record(steps_per_graph)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def run(self, nsteps):
        """Advance nsteps ensemble steps by graph replays; returns grid
        points x steps / s over all members."""
        if self.graph is None:
            self.capture()
        reps = max(1, nsteps // self.steps_per_graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.graph.replay()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return self.points * reps * self.steps_per_graph / dt

    def dt_values(self):
        return [float(d._scalars[1].item()) for d in self.members]
