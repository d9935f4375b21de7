"""Ensembles of uneven grids advanced together under one captured HIP graph.

Each member is a full LES domain with its own seed, Cs and CFL dt.  The
members' runs are those of spectralMethodSolver's loop (src/SM_solver.py:18-68,
first dt at CFL 0.3, then at the domain's CFL) with the loop state on the
device, so no host synchronisation happens between steps.  Members of one grid
size form a group whose spectral state is packed member by member
(``group.U_hat[m]`` is member m's ``U_hat``; the member's domain holds views of
it), and one ``sles_step_drive_batch`` call advances the whole group: every
pass of an RK step is one launch over all its members (grid rows = members),
so a group of small grids fills the device like one larger grid.  Members of
one size are split into groups of ``batch_size`` members.  Each group runs on
its own stream forked from the capture stream; ``steps_per_graph`` ensemble
steps (one step of every member each) are recorded once into a HIP graph
(``torch.cuda.CUDAGraph``) and replayed: one graph launch per replay.  The kernel variants of each group
come from the autotuner table (InitializeDomain applies it).

``batched=False`` runs the same steps with one ``sles_step_drive`` call per
member (each on its own stream); the two forms give identical states.
``les_model=None`` runs the members without an SGS model.

    ens = Ensemble([(64, 8), (96, 8), (128, 8)], precision="fp32", seed=0)
    ens.capture(steps_per_graph=1)
    rate = ens.run(200)     # grid points x steps / s over all members
"""

from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from . import _lib, fused
from .initial import isotropic_spectrum, set_spectral_state
from .initialize_domain import InitializeDomain

# driver-state slots (SLES_DRIVE_STATE_* of sles.h)
S_UMAX, S_DT, S_T, S_STATUS, S_STEPS, S_DT_APPLY, S_T_PAUSE, S_T_FINAL = range(8)
STATE_SIZE = 8


def batch_size(N, count, batch=None):
    """Members per batched launch for `count` members of grid N: `batch` (an
    int, or a dict by N), else enough members for about 2^20 grid points per
    launch (16 at 32^3, 4 at 64^3, 1 from 96^3 up).  Small grids are launch
    bound alone and fill the device together; from about 2^20 points a
    member's passes fill it by themselves, and members on separate streams
    then overlap different passes, which measured faster than one launch over
    all of them (FP32, 8 members: 64^3 4.0e9 grid points x steps / s in
    batches of 4 against 3.3e9 in one batch and 2.6e9 one member per launch;
    128^3 3.7e9 one member per launch against 3.4e9 in one batch)."""
    if isinstance(batch, dict):
        batch = batch.get(N)
    if batch is None:
        batch = max(1, (1 << 20) // N ** 3)
    return max(1, min(int(batch), count))


class Group:
    """Members of one grid size, packed for sles_step_drive_batch."""

    def __init__(self, members, cfl_first):
        d0 = members[0]
        self.members = members
        self.N = d0.N
        self.plan = d0.plan
        self.cfl_first = float(cfl_first)
        M = len(members)
        dev = d0.device
        shape = (M,) + tuple(d0.U_hat.shape)
        self.U_hat = torch.empty(shape, dtype=d0.cdtype, device=dev)
        self.U_hat0 = torch.zeros(shape, dtype=d0.cdtype, device=dev)
        self.U_hat1 = torch.zeros(shape, dtype=d0.cdtype, device=dev)
        for m, d in enumerate(members):
            self.U_hat[m].copy_(d.U_hat)
            d.U_hat, d.U_hat0, d.U_hat1 = self.U_hat[m], self.U_hat0[m], self.U_hat1[m]
        self.coeff = torch.tensor([d.sgs_coeff() for d in members], dtype=torch.float64, device=dev)
        self.state = torch.zeros((M, STATE_SIZE), dtype=torch.float64, device=dev)
        self.state[:, S_T] = math.nan
        self.state[:, S_T_PAUSE] = math.inf
        self.state[:, S_T_FINAL] = math.inf
        self.realloc_work()

    def realloc_work(self):
        """Allocate the batch workspace for the plan's current layout (its
        size follows the plan's SLES_TUNE_PAD: call again after changing it)."""
        d0 = self.members[0]
        nbytes = int(_lib.lib.sles_step_batch_work_bytes(self.plan.handle, d0.sgs_model, len(self.members)))
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=d0.device)

    @staticmethod
    def _scheme(d):
        if d.method == "RK4":
            return fused.SCHEME_RK4
        if d.method == "RK3":
            return fused.SCHEME_RK3
        raise ValueError("ensembles run RK4 or RK3")

    def step_batched(self, nsteps):
        d0 = self.members[0]
        prm = fused.params(d0, cfl=d0.CFL)
        _lib.call("sles_step_drive_batch", self.plan.handle, self._scheme(d0), ctypes.byref(prm), len(self.members),
                  self.coeff.data_ptr(), self.cfl_first, int(nsteps), self.U_hat.data_ptr(),
                  self.U_hat0.data_ptr(), self.U_hat1.data_ptr(), None, self.state.data_ptr(),
                  self.work.data_ptr(), d0.stream)

    def step_member(self, m, nsteps):
        d = self.members[m]
        prm = fused.params(d, cfl=d.CFL)
        _lib.call("sles_step_drive", d.plan.handle, self._scheme(d), ctypes.byref(prm), self.cfl_first, int(nsteps),
                  d.U_hat.data_ptr(), d.U_hat0.data_ptr(), d.U_hat1.data_ptr(), None, None, None,
                  self.state[m].data_ptr(), fused.work(d).data_ptr(), d.stream)


class Ensemble:
    def __init__(self, groups=((64, 8), (96, 8), (128, 8)), nu=1e-4, cs_range=(0.10, 0.20), seed=0,
                 precision="fp32", cfl=0.5, cfl_first=0.3, k0=4.0, les_model="Smagorinsky", method="RK4",
                 batched=True, batch=None):
        rng = np.random.default_rng(seed)
        self.members = []
        self.groups = []
        self.batched = bool(batched)
        i = 0
        for N, count in groups:
            ds = []
            for _ in range(count):
                Cs = float(rng.uniform(*cs_range))
                d = InitializeDomain(N=N, nu=nu, method=method, LES_ON=1 if les_model else 0, LES_model=les_model,
                                     Cs=Cs, CFL=cfl, precision=precision)
                # members run side by side (batched launches, or one stream
                # each), which fills the CUs: one strided tile per CU
                d.plan.set_tuning(narrow=0)
                set_spectral_state(d, isotropic_spectrum(N, k0=k0, seed=seed * 1000 + i, urms=1.0, dtype=d.cdtype))
                ds.append(d)
                i += 1
            k = batch_size(N, count, batch)
            for j in range(0, count, k):
                self.groups.append(Group(ds[j:j + k], cfl_first))
            self.members.extend(ds)
        n_streams = len(self.groups) if self.batched else len(self.members)
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.graph = None
        self.steps_per_graph = 1

    @property
    def points(self):
        return sum(d.N ** 3 for d in self.members)

    def _record(self, steps):
        cur = torch.cuda.current_stream()
        if self.batched:
            work = [(s, g.step_batched, ()) for g, s in zip(self.groups, self.streams)]
        else:
            pairs = [(g, m) for g in self.groups for m in range(len(g.members))]
            work = [(s, g.step_member, (m,)) for (g, m), s in zip(pairs, self.streams)]
        for s, fn, args in work:
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                fn(*args, steps)
        for s in self.streams:
            cur.wait_stream(s)

    def step_eager(self, steps=1):
        self._record(steps)

    def capture(self, steps_per_graph=1):
        """Record steps_per_graph ensemble steps into a HIP graph (the capture
        runs nothing; each replay advances every member)."""
        if not self.batched:
            for d in self.members:
                fused.work(d)  # per-member workspaces exist before the capture
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

    def states(self):
        """Driver state of every member (rows in member order; S_* slots)."""
        return torch.cat([g.state for g in self.groups]).cpu()

    def dt_values(self):
        """dt of each member's last step."""
        return self.states()[:, S_DT_APPLY].tolist()
