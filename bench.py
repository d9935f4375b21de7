#!/usr/bin/env python3
"""Benchmark: RK4 steps/s of the 256^3 decaying isotropic Smagorinsky LES, FP64.

Workload (one step = one iteration of the reference driver loop,
src/SM_solver.py:39-68, without I/O): N = 256, Smagorinsky Cs = 0.17,
Delta = dx = 2 pi/N, nu = 1e-4, RK4, dt from the CFL rule (0.3 for the first
step, then 0.5) evaluated on the device from max(u) of the state.  The loop
runs as the package's driver does (sles_step_drive): chunks of 10 steps with
the loop state on the device, U / curl of the state written at the end of
each chunk (the only point where the host can read them; each step
recomputes the physical fields it needs from U_hat).  The same steps with
one sles_step call each and U / curl written after every step are reported
in extra.  Initial field: seeded solenoidal random field with
E(k) ~ k^4 exp(-2 (k/4)^2), u_rms = 1 (synthetic).

Prints one JSON line (rank 0, one GPU).  Timing: W warm-up steps, then K
steps between two torch.cuda events on the stream the kernels run on.  The
roofline object reports the dominant kernel's algorithmic bytes per launch
over its mean launch time (hipEvents around each launch during K extra
profiled steps).  The CPU baseline times the oracle (numpy/scipy FP64
restatement of the reference, same FFT count and temporaries) for one
warm-up and three RK4 steps of the same 256^3 LES from the same seeded field
on the host cores (all cores and one core).  The FP32 form's velocity rel-L2
against the FP64 form after 100 steps (256^3 and 512^3) is reported beside
its speed-up.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "spectral-les_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
KINDS = ["x_inverse", "y_inverse", "z_fused", "y_forward", "x_forward_update", "cfl_reduce", "physical_out"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def field_bytes(N, precision):
    s = 8 if precision == "fp64" else 4
    R = N ** 3 * s
    C = N * N * (N // 2 + 1) * 2 * s
    return R, C


def kernel_traffic_fields(les):
    """(complex fields read, complex fields written) per launch of each pass;
    the compulsory traffic of the pass (every operand read once, every
    result written once)."""
    if les:
        return {
            "x_inverse": (3, 6),
            "y_inverse": (6, 9),
            "z_fused": (9, 8),
            "y_forward": (8, 6),
            # stages 1, 2 (the mean over the 4 stages is used below)
            "x_forward_update": (6 + 9, 6),
        }
    return {
        "x_inverse": (3, 6),
        "y_inverse": (6, 6),
        "z_fused": (6, 3),
        "y_forward": (3, 3),
        "x_forward_update": (3 + 9, 6),
    }


def xfwd_mean_fields(les, fused=False, steps=3):
    """Mean (read + write) complex fields per x-forward/update launch over one
    driver call of `steps` RK4 steps (sles_step_drive, U and curl written at
    its end).  Stage 0 reads W + U_hat and writes U_hat0, U_hat1, U_hat;
    stages 1-2 read W + U_hat, U_hat0, U_hat1 and write U_hat1, U_hat;
    stage 3 reads W + U_hat, U_hat1 and writes U_hat1, U_hat.
    fused (the plan's fuse_x): every launch also writes the output of the
    x-inverse that follows it (6 fields with LES, 5 without) -- the next
    stage's, or at stage 3 the next step's -- except the call's last stage 3,
    which writes the physical-output pass's (5 fields); those x-inverse
    passes then read nothing more."""
    w = 6 if les else 3
    per_stage = [w + 3 + 9, w + 9 + 6, w + 9 + 6, w + 6 + 6]
    total = steps * sum(per_stage)
    if fused:
        xi = 6 if les else 5
        total += (4 * steps - 1) * xi + 5
    return total / (4.0 * steps)


def step_bytes_minimal(N, precision, les):
    """Algorithmic bytes per RK4 step, 4 [n_T (R + C) + 15 C] (minimal
    transform set, each transform one read + one write; RK state 15 C)."""
    R, C = field_bytes(N, precision)
    nT = 21 if les else 9
    return 4 * (nT * (R + C) + 15 * C)


def make_domain(N, precision, seed=1):
    from sles.initial import isotropic_spectrum, set_spectral_state
    from sles.initialize_domain import InitializeDomain

    import torch

    d = InitializeDomain(
        N=N, nu=1e-4, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17, CFL=0.5, precision=precision
    )
    Uh = isotropic_spectrum(N, k0=4.0, seed=seed, urms=1.0, dtype=d.cdtype)
    set_spectral_state(d, Uh)
    del Uh
    torch.cuda.synchronize()
    return d


def run_gpu(N, precision, steps, warmup, profile_steps=3, chunk=10):
    """The driver loop on the device (sles_step_drive: dt from the CFL rule
    and the loop's stop rules evaluated per step on the GPU, first dt with
    CFL 0.3, then 0.5), enqueued in chunks of `chunk` steps with U and curl
    of the state written at the end of every chunk; `steps` timed steps
    after `warmup` untimed ones.  Also times the same steps one sles_step
    call each with U and curl written after every step."""
    import math

    import torch

    from sles import _lib, fused

    d = make_domain(N, precision)
    state = torch.zeros(8, dtype=torch.float64, device=d.device)
    state[2] = math.nan  # t starts at the first dt
    state[6] = math.inf  # no output pauses
    state[7] = math.inf  # no final time: the steps are counted here
    prm = fused.params(d, cfl=0.5)
    w = fused.work(d)

    def drive(n):
        _lib.call("sles_step_drive", d.plan.handle, fused.SCHEME_RK4, ctypes.byref(prm), 0.3, int(n),
                  d.U_hat.data_ptr(), d.U_hat0.data_ptr(), d.U_hat1.data_ptr(), None, d._U.data_ptr(),
                  d.curl.data_ptr(), state.data_ptr(), w.data_ptr(), d.stream)

    def run(n):
        while n > 0:
            k = min(chunk, n)
            drive(k)
            n -= k

    d.U_hat1.copy_(d.U_hat)
    run(warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    run(steps)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    st = state.cpu().numpy()
    ok = bool(st[3] == 0 and st[4] == warmup + steps)
    if not ok:
        log(f"[bench] WARNING: driver state {st.tolist()}: steps were not all taken")
    # the same steps, one sles_step call each, U and curl written after every step
    e0.record(stream)
    for _ in range(steps):
        fused.rk4_step(d, dt_device=True, cfl=0.5)
    e1.record(stream)
    e1.synchronize()
    ms_each = e0.elapsed_time(e1)
    # per-kernel timing over a few more steps (one chunk)
    ms_k = (ctypes.c_double * 8)()
    cnt_k = (ctypes.c_long * 8)()
    kern = {}
    if profile_steps:
        _lib.call("sles_profile_enable", 1)
        drive(profile_steps)
        torch.cuda.synchronize()
        _lib.call("sles_profile_read", ms_k, cnt_k, 7)
        _lib.call("sles_profile_enable", 0)
        kern = {KINDS[i]: (ms_k[i], cnt_k[i]) for i in range(7)}
    out = dict(ms_total=ms, ms_total_step_each=ms_each, steps=steps, dt_last=float(st[1]),
               umax_last=float(st[0]), kernels=kern, ok=ok, tuning=d.plan.get_tuning(), profile_steps=profile_steps)
    del d
    torch.cuda.empty_cache()
    return out


def measured_traffic(kernel, N, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/hbm_traffic_rk4_smagorinsky_<N>_<precision>.txt,
    scaled by the plain-copy calibration recorded there), or None."""
    f = os.path.join(ROOT, "profiles", f"hbm_traffic_rk4_smagorinsky_{N}_{precision}.txt")
    key = f"traffic_bytes_per_launch {kernel} {precision} {N}:"
    try:
        with open(f) as fh:
            for line in fh:
                if line.startswith(key):
                    return int(line[len(key):].strip())
    except OSError:
        return None
    return None


def tgv_gpu(N=64, steps=200, warmup=5, graph_steps=10):
    """The CPU-runnable Taylor-Green case: 64^3, nu = 1/1600, dt = 0.01, RK4,
    no LES, FP64; fused steps timed with events (U and curl written each step),
    launched one by one and as replays of a HIP graph of graph_steps steps."""
    import numpy as np
    import torch

    from sles import fused
    from sles.initial import initial_velocity
    from sles.initialize_domain import InitializeDomain

    d = InitializeDomain(N=N, nu=1 / 1600, time_stepSize=0.01, method="RK4")
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    initial_velocity(d, np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X))
    d.U_hat1.copy_(d.U_hat)
    for _ in range(warmup):
        fused.rk4_step(d)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fused.rk4_step(d)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    out = {"steps_per_s": round(1e3 / ms, 1), "ms_per_step": round(ms, 4), "steps": steps}
    # the same steps captured into one graph (the step is launch-bound at 64^3)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fused.rk4_step(d)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(graph_steps):
            fused.rk4_step(d)
    g.replay()
    torch.cuda.synchronize()
    reps = max(1, steps // graph_steps)
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    e1.synchronize()
    msg = e0.elapsed_time(e1) / (reps * graph_steps)
    out["graph_steps_per_s"] = round(1e3 / msg, 1)
    out["graph_ms_per_step"] = round(msg, 4)
    return out


def fft_rates(N, precision, reps=10):
    """Generic 3D transforms (sles_fft_r2c / sles_fft_c2r) of a batch of 3
    fields: ms per call and algorithmic GB/s, 3 (R + 5 C) bytes per call (z
    pass R in / C out, y and x passes C in / C out, each field once)."""
    import torch

    from sles import _lib

    prec = _lib.SLES_FP64 if precision == "fp64" else _lib.SLES_FP32
    rdt = torch.float64 if precision == "fp64" else torch.float32
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    plan = _lib.Plan(N, prec, torch.device("cuda"))
    u = torch.randn(3, N, N, N, dtype=rdt, device="cuda")
    U = torch.empty(3, N, N, N // 2 + 1, dtype=cdt, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    R, C = field_bytes(N, precision)
    out = {}
    for name, xcd, fn in (
        ("r2c", 0, lambda: _lib.call("sles_fft_r2c", plan.handle, u.data_ptr(), U.data_ptr(), 3, 1.0, s)),
        ("c2r", 0, lambda: _lib.call("sles_fft_c2r", plan.handle, U.data_ptr(), u.data_ptr(), U.data_ptr(), 3,
                                     1.0 / N ** 3, s)),
        ("r2c_xcd", 1, lambda: _lib.call("sles_fft_r2c", plan.handle, u.data_ptr(), U.data_ptr(), 3, 1.0, s)),
        ("c2r_xcd", 1, lambda: _lib.call("sles_fft_c2r", plan.handle, U.data_ptr(), u.data_ptr(), U.data_ptr(),
                                         3, 1.0 / N ** 3, s)),
    ):
        plan.set_tuning(xcd=xcd)
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms": round(ms, 4), "algorithmic_GBps": round(3 * (R + 5 * C) / (ms * 1e-3) / 1e9, 1)}
    return out


def ensemble_rates(steps=100, steps_per_graph=10):
    """8 x 64^3 + 8 x 96^3 + 8 x 128^3 FP32 Smagorinsky RK4 members (own seeds,
    Cs ~ U[0.10, 0.20], CFL dt on the device), graph replays of 10 ensemble
    steps: aggregate grid points x steps / s with one launch per pass per grid
    size (batched) and with one sles_step_drive call per member."""
    import torch

    from sles.studies import ensemble_rate

    out = {"members": "8 x 64^3 + 8 x 96^3 + 8 x 128^3", "steps": steps, "steps_per_graph": steps_per_graph}
    for batched in (True, False):
        log(f"[bench] ensemble {'batched' if batched else 'one call per member'}")
        r = ensemble_rate(steps=steps, steps_per_graph=steps_per_graph, batched=batched)
        key = "batched" if batched else "per_member"
        out[key] = {"grid_points_x_steps_per_s": round(r["grid_points_x_steps_per_s"], 1),
                    "all_members_stepped": r["all_members_stepped"]}
        torch.cuda.empty_cache()
    out["batched_speedup"] = round(out["batched"]["grid_points_x_steps_per_s"] /
                                   out["per_member"]["grid_points_x_steps_per_s"], 3)
    return out


def tgv_cpu(N=64, steps=20):
    """The same Taylor-Green case on the oracle (numpy FP64, one process)."""
    import numpy as np

    from oracle import sles_oracle as O

    o = O.Domain(N=N, nu=1 / 1600, time_stepSize=0.01, method="RK4")
    x, y, z = O.grid(N)
    O.initial_velocity(o, np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), np.zeros_like(x))
    o.U_hat1[:] = o.U_hat
    t0 = time.perf_counter()
    for _ in range(steps):
        O.rk4_step(o)
    t = time.perf_counter() - t0
    return {"steps_per_s": round(steps / t, 3), "steps": steps, "cores": 1,
            "sample": f"{steps} RK4 steps of the 64^3 Taylor-Green case, numpy pocketfft, {t:.2f} s"}


def x_forward_pattern_ceiling(N, precision, pad, fused=True):
    """The x-forward + update pass's own access pattern with no arithmetic
    (sles_stream_probe, the radix-4 x-forward tile, XCD order, the step's
    padded plane layout): the RK4 stages' traffic mixes -- 9+9, 15+6, 15+6,
    12+6 fields read + written, and 6 more written by each launch that also
    runs the next x-inverse (fused) -- in algorithmic GB/s: the bandwidth
    this tile shape and traffic reach on this device, beside the 8 TB/s
    peak."""
    from sles.studies import stream_ceiling

    xi = 6 if fused else 0
    mixes = ((9, 9 + xi), (15, 6 + xi), (12, 6 + xi))
    r = stream_ceiling(N, precision, mixes=mixes, tiles=(0,), verbose=False, pad=pad)
    R, C = field_bytes(N, precision)
    per_stage = (mixes[0], mixes[1], mixes[1], mixes[2])  # stages 0-3
    ms = sum(r[f"tile0_{a}r{b}w"]["ms"] for a, b in per_stage) / 4
    nbytes = sum(a + b for a, b in per_stage) / 4 * C
    return {"GBps": round(nbytes / (ms * 1e-3) / 1e9, 1), "avg_launch_ms": round(ms, 4), "pad_lines": pad,
            "fields_per_launch": sum(a + b for a, b in per_stage) / 4,
            "kernel": "sles_stream_probe tile 0 (loads and stores of the x-forward tile, no arithmetic)"}


def roofline_of(res, N, precision, les=True):
    R, C = field_bytes(N, precision)
    kern = res["kernels"]
    fields = kernel_traffic_fields(les)
    best = max(fields, key=lambda k: kern[k][0])
    tot_ms, cnt = kern[best]
    avg_s = tot_ms / cnt / 1e3
    # x-forward launches that also run the next x-inverse (plan knob fuse_x)
    fused = bool(res["tuning"].get("fuse_x", 0))
    steps = res["profile_steps"]
    if best == "x_forward_update":
        nbytes = xfwd_mean_fields(les, fused, steps) * C
    else:
        nbytes = sum(fields[best]) * C
    achieved = nbytes / avg_s / 1e9
    per_kernel = {}
    for k, (t, c) in kern.items():
        if c == 0:
            continue
        entry = {"ms_per_launch": round(t / c, 4), "launches": c}
        if k in fields:
            b = xfwd_mean_fields(les, fused, steps) * C if k == "x_forward_update" else sum(fields[k]) * C
            entry["algorithmic_GBps"] = round(b / (t / c / 1e3) / 1e9, 1)
        per_kernel[k] = entry
    return {
        "bound": "hbm",
        "kernel": best,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(best, N, precision),
        "algorithmic_bytes_per_launch": int(nbytes),
        "avg_launch_ms": round(avg_s * 1e3, 4),
    }, per_kernel


def cpu_baseline(N, U_hat, steps=3):
    """The oracle (numpy / scipy FP64 restatement of the reference, with its
    108 FFTs per LES RK4 step and its numpy temporaries) on the same 256^3
    Smagorinsky workload as the GPU: the same seeded initial spectrum U_hat
    (complex128, host), dt from the CFL rule (0.3 for the first step, then
    0.5).  Two variants, each one warm-up step then `steps` timed steps
    (mean +- sd): all host cores (scipy.fft workers = cores, the reported
    value) and one core (numpy pocketfft); the pointwise numpy work is
    single-threaded in both."""
    import statistics

    import numpy as np

    from oracle import sles_oracle as O

    cores = min(O.cpu_count(), 16)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    res = {}
    for name, workers in (("all_cores", cores), ("one_core", None)):
        d = O.Domain(N=N, nu=1e-4, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17, CFL=0.5,
                     fft_workers=workers)
        d.U_hat[:] = U_hat
        for i in range(3):
            d.U[i][:] = d.fft.backward(d.U_hat[i], d.U[i], normalize=True)
        d.curl[:] = O.del_cross_U(d.fft, 3, d.U_hat, d.curl, d.K)
        d.U_hat1[:] = d.U_hat
        times = []
        for k in range(steps + 1):
            d.dt = O.compute_dt(d.U, O.find_max_velocity(d.U), d.nu, dx=d.dx, CFL=0.3 if k == 0 else 0.5)
            t0 = time.perf_counter()
            O.rk4_step(d)
            times.append(time.perf_counter() - t0)
        timed = times[1:]
        m = statistics.mean(timed)
        sd = statistics.stdev(timed) if len(timed) > 1 else 0.0
        res[name] = {"steps_per_s": round(1.0 / m, 5), "s_per_step_mean": round(m, 3), "s_per_step_sd": round(sd, 3),
                     "warmup_s": round(times[0], 3), "timed_steps": len(timed),
                     "threads": cores if workers else 1}
        log(f"[bench] cpu {name}: {m:.2f} +- {sd:.2f} s per step ({len(timed)} timed after 1 warm-up)")
        del d
    a = res["all_cores"]
    return {
        "value": a["steps_per_s"],
        "unit": "RK4 steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"1 warm-up + {steps} timed RK4 steps of the {N}^3 Smagorinsky LES from the GPU run's seeded "
        f"initial field (CFL dt), oracle FP64 (108 FFTs + numpy temporaries per step) with scipy.fft "
        f"workers={cores}, pointwise numpy single-threaded: {a['s_per_step_mean']} +- {a['s_per_step_sd']} s "
        f"per step; one-core variant (numpy pocketfft): {res['one_core']['s_per_step_mean']} +- "
        f"{res['one_core']['s_per_step_sd']} s per step",
        "cpu": cpu,
        "variants": res,
    }


def fp32_error(N, steps=100):
    """Velocity rel-L2 of the FP32 form against the FP64 form after `steps`
    RK4 steps on the same seeded field and the FP64 run's dt sequence."""
    from sles.studies import precision_study

    r = precision_study(N=N, checkpoints=(steps,), seed=1)["results"][-1]
    return {"steps": steps, "velocity_rel_l2": r["velocity_rel_l2"],
            "max_rel_Ek_error_resolved_shells": r["max_rel_Ek_error_resolved_shells"]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed oracle steps per CPU-baseline variant")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-form comparison run")
    ap.add_argument("--no-512", action="store_true", help="skip the 512^3 FP64 / FP32 step timings in extra")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble throughput in extra")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy 5 times (reference kernel for HBM counters)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(
            f"bench.py runs on one GPU only: --gpus {a.gpus} is not supported "
            "(single-GPU implementation: no multi-GPU path)",
            file=sys.stderr,
        )
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    N = a.n
    if a.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 5 x 1 GiB device copies done")
        del src, dst
    log(f"[bench] N={N} FP64 Smagorinsky RK4: warmup {a.warmup}, timed {a.steps}")
    r64 = run_gpu(N, "fp64", a.steps, a.warmup)
    ms_step = r64["ms_total"] / a.steps
    value = 1e3 / ms_step
    roof, per_kernel = roofline_of(r64, N, "fp64")
    roof["pattern_ceiling"] = x_forward_pattern_ceiling(N, "fp64", r64["tuning"]["pad"],
                                                        bool(r64["tuning"].get("fuse_x", 0)))
    step_bytes = step_bytes_minimal(N, "fp64", True)
    log(f"[bench] fp64 {value:.2f} steps/s ({ms_step:.3f} ms/step); dominant {roof['kernel']} "
        f"{roof['achieved']} GB/s")
    extra = {
        "driver_chunk_steps": 10,
        "steps_per_s_one_call_per_step": round(1e3 * a.steps / r64["ms_total_step_each"], 3),
        "grid_points_x_steps_per_s": round(value * N ** 3, 1),
        "step_roofline_frac": round(step_bytes * value / (HBM_PEAK_GBS * 1e9), 4),
        "step_algorithmic_GB": round(step_bytes / 1e9, 3),
        "kernels_fp64": per_kernel,
        "tuning_fp64": r64["tuning"],
        "dt_last": r64["dt_last"],
        "fp64_valid": r64["ok"],
    }
    if not a.no_fp32:
        log("[bench] FP32 form")
        r32 = run_gpu(N, "fp32", a.steps, a.warmup)
        v32 = 1e3 * a.steps / r32["ms_total"]
        extra["fp32_steps_per_s"] = round(v32, 3)
        extra["fp32_steps_per_s_one_call_per_step"] = round(1e3 * a.steps / r32["ms_total_step_each"], 3)
        extra["fp32_valid"] = r32["ok"]
        extra["fp32_speedup_vs_fp64"] = round(v32 / value, 3)
        extra["kernels_fp32"] = roofline_of(r32, N, "fp32")[1]
        log(f"[bench] fp32 {v32:.2f} steps/s")
        extra["fp32_vs_fp64_error"] = fp32_error(N)
        log(f"[bench] fp32 vs fp64 after 100 steps: {extra['fp32_vs_fp64_error']}")
    if not a.no_512:
        for pr in ("fp64", "fp32"):
            log(f"[bench] 512^3 {pr}")
            r = run_gpu(512, pr, 10, 2, profile_steps=0)
            extra[f"grid512_{pr}"] = {"steps_per_s": round(1e4 / r["ms_total"], 3),
                                       "ms_per_step": round(r["ms_total"] / 10, 3), "steps": 10, "valid": r["ok"],
                                       "steps_per_s_one_call_per_step": round(1e4 / r["ms_total_step_each"], 3)}
            log(f"[bench] 512^3 {pr}: {extra[f'grid512_{pr}']}")
        extra["grid512_fp32_speedup_vs_fp64"] = round(extra["grid512_fp32"]["steps_per_s"] /
                                                      extra["grid512_fp64"]["steps_per_s"], 3)
        if not a.no_fp32:
            extra["grid512_fp32_vs_fp64_error"] = fp32_error(512)
            log(f"[bench] 512^3 fp32 vs fp64 after 100 steps: {extra['grid512_fp32_vs_fp64_error']}")
    sizes = (N,) if a.no_512 else (N, 512)
    extra["fft_3d_batch3"] = {f"{n}_{pr}": fft_rates(n, pr) for n in sizes for pr in ("fp64", "fp32")}
    log(f"[bench] generic 3D FFT: {extra['fft_3d_batch3']}")
    log("[bench] 64^3 Taylor-Green (no LES) FP64")
    extra["taylor_green_64_fp64"] = tgv_gpu()
    log(f"[bench] tgv64 {extra['taylor_green_64_fp64']['steps_per_s']} steps/s")
    if not a.no_ensemble:
        extra["ensemble_fp32"] = ensemble_rates()
        log(f"[bench] ensemble: {extra['ensemble_fp32']}")
    cpu = None
    if not a.no_cpu:
        extra["taylor_green_64_cpu_oracle"] = tgv_cpu()
        log("[bench] CPU baseline (oracle, 1 warm-up + 3 RK4 steps, all cores and one core)")
        from sles.initial import isotropic_spectrum

        Uh = isotropic_spectrum(N, k0=4.0, seed=1, urms=1.0).cpu().numpy()  # the GPU run's field (make_domain)
        torch.cuda.empty_cache()
        cpu = cpu_baseline(N, Uh, steps=a.cpu_steps)
        del Uh
        log(f"[bench] cpu {cpu['value']} steps/s on {cpu['cores']} cores")
    line = {
        "metric": "RK4 steps/s at 256^3 (FP64 vs FP32 @ rel-L2<=1e-4); % HBM peak",
        "value": round(value, 3),
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^4 exp(-2(k/4)^2) solenoidal field, u_rms=1)",
        "config": {
            "workload": f"decaying isotropic LES {N}^3, Smagorinsky Cs=0.17, nu=1e-4, RK4, CFL 0.3 then 0.5, "
            "driver loop on the device (dt and stop rules per step on the GPU, chunks of 10 steps, U/curl of the "
            "state written at each chunk end)",
            "grid": N,
            "les_model": "Smagorinsky",
            "scheme": "RK4",
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
