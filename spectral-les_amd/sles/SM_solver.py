"""Driver-side reductions and the time loop (reference src/SM_solver.py).

* ``find_max_velocity`` / ``compute_dt`` (SM_solver.py:245-277): the CFL rule
  on a device FP64 max reduction (signed max over all components, as
  ``numpy.max`` of the list U).
* ``Ek_plot_variables`` / ``calculate_Ek`` (SM_solver.py:91-117,154-192): the
  binned mean spectrum from one device pass with FP64 bin sums.
* ``spectralMethodSolver`` (SM_solver.py:18-83): the time loop with the same
  dt rule, stop conditions and NaN guard.  With the fused kernels the loop
  runs on the device (``sles_step_drive``): dt, t and the stop rules are
  evaluated per step on the GPU, the host enqueues chunks of steps (captured
  into a HIP graph when ``graph=True``) and reads the loop state once per
  chunk; it stops exactly at the loop tops where outputs are due.  E(k) is
  written to ``Ek_plot_data.csv`` with the reference's ``k-{t}`` / ``Ek-{t}``
  columns when ``No_Ek_plot > 0``; velocity snapshots (``output_h5``,
  SM_solver.py:280-302) are written every ``file_out_time`` as raw binary +
  XDMF pairs in ``dumpfile/`` (h5py is not available; see
  ``output_snapshot``).  Plotting is not part of this package.
"""

from __future__ import annotations

import ctypes
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
    "output_snapshot",
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
    """Bin bookkeeping of the E(k) spectrum (SM_solver.py:91-117).  Like the
    reference (:114-116) it removes a ``Ek_plot_data.csv`` left in the working
    directory by an earlier run (``path=None``: no file is touched)."""

    def __init__(self, domain, path="Ek_plot_data.csv"):
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
        self._work = None  # binning partials, allocated by the first calculate_Ek
        if path is not None and os.path.exists(path):
            os.remove(path)


def calculate_Ek(domain, E):
    """temp_k[i] = mean |k| and temp_Ek[i] = mean Ek over modes with K^2 in
    [(5i)^2, (5i+5)^2), Ek = sum_i |U_hat_i|^2 doubled for kz >= 1
    (SM_solver.py:154-192)."""
    nb = E.smoothing_para + 1
    dev = E._dev
    uh = domain.U_hat
    if not uh.is_contiguous():
        uh = uh.contiguous()
    if E._work is None:
        E._work = torch.empty(int(_lib.lib.sles_spectrum_work_bytes(domain.plan.handle, nb)), dtype=torch.uint8,
                              device=domain.device)
    _lib.call("sles_spectrum", domain.plan.handle, uh.data_ptr(), nb, dev[0].data_ptr(), dev[1].data_ptr(),
              dev[2].data_ptr(), E._work.data_ptr(), domain.stream)
    ek, ks, cnt = dev.cpu().numpy()
    E.temp_k[:] = 0
    E.temp_Ek[:] = 0
    has = cnt > 0
    E.temp_k[has] = ks[has] / cnt[has]
    E.temp_Ek[has] = ek[has] / cnt[has]


def _write_ek_csv(E, t, path="Ek_plot_data.csv"):
    """Add the columns k-{t:g} / Ek-{t:g} to the CSV (SM_solver.py:137-145:
    read_csv, add two columns, write back; a new file starts with them)."""
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


def make_directory(name, delete_old=True):
    """utilites.py:8-15: create the directory, or empty it."""
    if not os.path.isdir(name):
        os.makedirs(name, exist_ok=True)
    elif delete_old:
        for f in os.listdir(name):
            os.remove(os.path.join(name, f))


_XDMF = """<?xml version="1.0" ?>
<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>
<Xdmf Version="2.0">
  <Domain>
    <Grid Name="dump{tag}" GridType="Uniform">
      <Time Value="{step}"/>
      <Topology TopologyType="3DCoRectMesh" Dimensions="{n} {n} {n}"/>
      <Geometry GeometryType="ORIGIN_DXDYDZ">
        <DataItem Format="XML" NumberType="Float" Precision="8" Dimensions="3">0 0 0</DataItem>
        <DataItem Format="XML" NumberType="Float" Precision="8" Dimensions="3">{dx!r} {dx!r} {dx!r}</DataItem>
      </Geometry>
{attrs}    </Grid>
  </Domain>
</Xdmf>
"""
_XDMF_ATTR = """      <Attribute Name="{name}" AttributeType="Scalar" Center="Node">
        <DataItem Format="Binary" NumberType="Float" Precision="{prec}" Endian="Little" Seek="{seek}" Dimensions="{n} {n} {n}">{file}</DataItem>
      </Attribute>
"""


def output_snapshot(domain, t, file_out_time, directory="dumpfile"):
    """Velocity snapshot at time t, the counterpart of output_h5
    (SM_solver.py:280-302), which writes u, v, w as step round(t //
    file_out_time) of dump{t // file_out_time:g}.h5 (utilites.py:27-34 adds an
    XDMF description and moves both into dumpfile/).  h5py is not available
    here, so the same fields are written as one raw little-endian binary file
    dump{tag}.bin (u, then v, then w; each (N, N, N) C order, z fastest, in
    the run's precision) with an XDMF file dump{tag}.xdmf that describes them
    on the uniform grid x = i 2 pi / N.  Returns the two paths."""
    d = domain
    tag = f"{t // file_out_time:g}"
    step = round(t // file_out_time)
    os.makedirs(directory, exist_ok=True)
    binname = f"dump{tag}.bin"
    U = d._U.detach().cpu().numpy()
    U.astype(U.dtype.newbyteorder("<"), copy=False).tofile(os.path.join(directory, binname))
    n = d.N
    prec = U.dtype.itemsize
    attrs = "".join(_XDMF_ATTR.format(name=nm, prec=prec, seek=i * n ** 3 * prec, n=n, file=binname)
                    for i, nm in enumerate(("u", "v", "w")))
    xpath = os.path.join(directory, f"dump{tag}.xdmf")
    with open(xpath, "w") as f:
        f.write(_XDMF.format(tag=tag, step=step, n=n, dx=float(d.dx), attrs=attrs))
    return os.path.join(directory, binname), xpath


def _step(domain):
    if domain.method == "RK4":
        RK4.solveForNextTimestep(domain)
    elif domain.method == "ETD2":
        ETD2.solveForNextTimestep(domain)
    elif domain.method == "RK3":
        RK3.solveForNextTimestep(domain)
    else:
        raise ValueError("Invalid method name")


class _Outputs:
    """The loop-top outputs of SM_solver.py:40-45 and their counters."""

    def __init__(self, domain, file_out_time, No_Ek_plot, directory):
        self.d = domain
        self.file_out_time = file_out_time
        self.file_time = 0.0
        self.Ek_interval = domain.T / No_Ek_plot if No_Ek_plot != 0 else 0
        self.Ek_counter = 0
        self.directory = directory
        self.Ek_var = Ek_plot_variables(domain) if self.Ek_interval > 0 else None
        if file_out_time > 0:
            make_directory(directory, delete_old=True)

    def next_time(self):
        """Earliest t at which a loop top writes something (inf: never)."""
        t = math.inf
        if self.file_out_time > 0:
            t = min(t, self.file_time)
        if self.Ek_interval > 0:
            t = min(t, self.Ek_counter)
        return t

    def write(self, t):
        if self.file_out_time > 0 and t >= self.file_time:
            output_snapshot(self.d, t, self.file_out_time, self.directory)
            self.file_time += self.file_out_time
        if self.Ek_interval > 0 and t >= self.Ek_counter:
            calculate_Ek(self.d, self.Ek_var)
            _write_ek_csv(self.Ek_var, t)
            self.Ek_counter += self.Ek_interval


# states of sles_step_drive (include/sles.h)
_RUNNING, _DONE, _NAN, _DT_TOO_SMALL, _PAUSED, _RESUME = range(6)
_S_UMAX, _S_DT, _S_T, _S_STATUS, _S_STEPS, _S_DT_APPLY, _S_T_PAUSE, _S_T_FINAL = range(8)


def _device_loop_supported(domain):
    if not getattr(domain, "fused", False):
        return False
    if domain.method == "ETD2":
        return getattr(domain, "etd2_form", "cox_matthews") == "cox_matthews"
    return domain.method in ("RK4", "RK3")


class _DeviceLoop:
    """Chunks of fused steps with the loop state on the device."""

    def __init__(self, domain, chunk, graph):
        from . import fused

        self.d = domain
        self.fused = fused
        self.chunk = max(1, int(chunk))
        self.graph = bool(graph)
        self.graphs = {}
        self.host_reads = 0
        self.replays = 0
        self.state = torch.zeros(8, dtype=torch.float64, device=domain.device)
        self.state[_S_T] = math.nan
        self.state[_S_T_PAUSE] = math.inf
        self.state[_S_T_FINAL] = domain.T

    def _scheme(self, first):
        f = self.fused
        m = self.d.method
        if m == "RK4":
            return f.SCHEME_RK4, None
        if m == "RK3":
            return f.SCHEME_RK3, None
        # ETD2: the first step is an RK4 step that also stores F(u_0) in F0
        return (f.SCHEME_RK4 if first else f.SCHEME_ETD2), self.d.F0

    def _enqueue(self, nsteps, first):
        d = self.d
        scheme, aux = self._scheme(first)
        prm = self.fused.params(d, cfl=d.CFL)
        w = self.fused.work(d)
        _lib.call("sles_step_drive", d.plan.handle, scheme, ctypes.byref(prm), 0.3, int(nsteps),
                  d.U_hat.data_ptr(), d.U_hat0.data_ptr(), d.U_hat1.data_ptr(),
                  aux.data_ptr() if aux is not None else None, d._U.data_ptr(), d.curl.data_ptr(),
                  self.state.data_ptr(), w.data_ptr(), d.stream)

    def run_chunk(self, nsteps, first):
        """Enqueue nsteps steps (one graph replay when captured) and return
        the loop state read back once."""
        if self.graph and not first and nsteps == self.chunk:
            g = self.graphs.get(nsteps)
            if g is None:
                self.fused.work(self.d)  # the workspace exists before the capture
                torch.cuda.synchronize(self.d.device)
                g = torch.cuda.CUDAGraph()
                # capture records the launches without running them; the
                # replay below runs the chunk
                with torch.cuda.graph(g):
                    self._enqueue(nsteps, first)
                self.graphs[nsteps] = g
            g.replay()
            self.replays += 1
        else:
            self._enqueue(nsteps, first)
        self.host_reads += 1
        return self.state.cpu().tolist()

    def set(self, slot, value):
        self.state[slot : slot + 1].fill_(float(value))


def spectralMethodSolver(domain, file_out_time: float = 0.0, No_Ek_plot: int = 5, verbose: bool = False,
                         max_steps: int | None = None, chunk: int = 10, graph: bool = False,
                         dumpdir: str = "dumpfile"):
    """Time loop of SM_solver.py:18-83.

    Returns the number of steps taken (the reference's counter1: a step
    after which dt is NaN is not counted).  ``max_steps`` (not in the
    reference) stops after that many steps.  With the fused kernels the loop
    runs on the device in chunks of ``chunk`` steps (one host read of the
    loop state per chunk; ``graph=True`` replays each full chunk as a
    captured HIP graph); the operator-by-operator path (``fused=False``) and
    the literal reference ETD2 run the same loop on the host."""
    if _device_loop_supported(domain):
        return _device_solver(domain, file_out_time, No_Ek_plot, verbose, max_steps, chunk, graph, dumpdir)
    return _host_solver(domain, file_out_time, No_Ek_plot, verbose, max_steps, dumpdir)


def _device_solver(domain, file_out_time, No_Ek_plot, verbose, max_steps, chunk, graph, dumpdir):
    d = domain
    start = time.time()
    out = _Outputs(d, file_out_time, No_Ek_plot, dumpdir)
    d.U_hat1.copy_(d.U_hat)  # SM_solver.py:33
    loop = _DeviceLoop(d, chunk, graph)
    d._drive = loop
    t_pause = out.next_time()
    loop.set(_S_T_PAUSE, t_pause)
    taken = 0
    t = dt = math.nan
    st = None
    while True:
        first = d.method == "ETD2" and d.firstIteration
        n = 1 if first else loop.chunk
        if max_steps is not None:
            n = min(n, max_steps - taken)
            if n <= 0:
                break
        if not first and math.isfinite(t) and math.isfinite(dt) and dt > 0:
            # end the chunk near the next stop (output time or T), where the
            # device pauses: later steps of a chunk would be skipped there
            horizon = min(t_pause, d.T)
            if math.isfinite(horizon):
                n = max(1, min(n, math.ceil((horizon - t) / dt) + 1))
        st = loop.run_chunk(n, first)
        status = int(st[_S_STATUS])
        taken = int(st[_S_STEPS])
        t, dt = st[_S_T], st[_S_DT]
        if first and taken > 0:
            d.firstIteration = False
        if verbose:
            print(f"t = {t:<10g} dt = {dt:<10g} max u = {st[_S_UMAX]:g} steps = {taken}", flush=True)
        if status in (_RUNNING, _RESUME):
            continue
        if status == _PAUSED:
            out.write(t)
            t_pause = out.next_time()
            loop.set(_S_T_PAUSE, t_pause)
            loop.set(_S_STATUS, _RESUME)
            continue
        if status == _DT_TOO_SMALL:
            d.dt = dt
            raise DtTooSmall(f"Time step (dt = {dt:g}) is smaller than 1e-6")
        break  # _DONE or _NAN
    status = int(st[_S_STATUS]) if st is not None else _RUNNING
    if st is not None:
        d.dt = st[_S_DT]
    if status in (_RUNNING, _RESUME) and taken > 0:
        # stopped by max_steps: the reference computes dt after every step
        # (SM_solver.py:60-67), the device only at the next loop top
        d.dt = compute_dt(d.U, find_max_velocity(d.U, d.comm), d.nu, d.comm, dx=d.dx, CFL=d.CFL)
        if d.dt != d.dt:
            status = _NAN
    total = time.time() - start
    if verbose and taken:
        print(f"Time per gridpoint per timestep: {total / (d.N ** 3 * taken)}", flush=True)
    if status == _NAN and taken > 0:
        return taken - 1
    return taken


def _host_solver(domain, file_out_time, No_Ek_plot, verbose, max_steps, dumpdir):
    d = domain
    u_max = find_max_velocity(d.U, d.comm)
    d.dt = compute_dt(d.U, u_max, d.nu, d.comm, dx=d.dx, CFL=0.3)
    t = d.dt
    start = time.time()
    out = _Outputs(d, file_out_time, No_Ek_plot, dumpdir)
    d.U_hat1.copy_(d.U_hat)
    steps = 0
    while True:
        out.write(t)
        if (t >= d.T) or (d.dt < 1e-8):
            break
        if max_steps is not None and steps >= max_steps:
            break
        if verbose:
            print(f"t = {t:<10g} dt = {d.dt:<10g} max u = {u_max:g}", flush=True)
        t += d.dt
        _step(d)
        u_max = find_max_velocity(d.U, d.comm)
        d.dt = compute_dt(d.U, u_max, d.nu, d.comm, dx=d.dx, CFL=d.CFL)
        if d.dt != d.dt:
            break
        steps += 1
    total = time.time() - start
    if verbose and steps:
        print(f"Time per gridpoint per timestep: {total / (d.N ** 3 * steps)}", flush=True)
    return steps
