"""Autotuner for the fused time-step kernels and the generic transforms.

For one (N, precision, LES model) it times every combination of the kernel
variants a plan exposes (``sles_plan_set_tuning``: radix-8 or radix-4 line
plans for the strided passes, the x-forward/update pass and the z kernel --
which also fixes each pass's tile: the columns per block follow from the
plan's threads per line and the 64 KiB LDS tile -- the XCD-aware tile order
and the fusion of the x-forward with the next x-inverse) on a few RK4 steps
of a synthetic field through the device driver loop, and keeps the fastest;
then, with those variants, it times the padding of the workspace's kx planes
(``pad``, in 128-byte lines, N >= 64) and, for N = 64, 96 and 128, the tile
width of the strided passes (``narrow``: one or four tiles per CU).  For
N = 256 and 512 it also
times the four tile widths of the generic transforms' strided passes
(``fft_wide``: twice the columns per tile in the x and/or y pass) on a
batched forward + backward transform.  Results persist in a JSON table
(default: ``tuning.json`` next to this file, override with
``SLES_TUNING_TABLE``) that ``InitializeDomain`` applies to its plan.

    python -m sles.autotune --n 256 --precision fp64 --les smagorinsky

Every variant computes the same arithmetic up to round-off order, so the
choice never changes results beyond FP rounding (tests/test_gpu_fused.py
checks each variant against the oracle).
"""

from __future__ import annotations

import argparse
import itertools
import json
import os
import time

import torch

from . import _lib

TABLE = os.environ.get("SLES_TUNING_TABLE", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json"))
KNOBS = ("strided_radix4", "xfwd_radix4", "z_radix4", "xcd", "fuse_x", "fft_wide", "pad", "narrow")
WIDE_SIZES = (256, 512)  # sizes with wide generic-transform tiles
DEFAULTS = dict(strided_radix4=0, xfwd_radix4=0, z_radix4=0, xcd=1, fuse_x=1)  # a new plan's settings
PADS = (0, 5, 17, 33)  # workspace plane padding tried, in 128-byte lines
PAD_MIN_N = 64  # smaller grids keep the compact layout (launch bound)
NARROW_SIZES = (64, 96, 128)  # sizes with narrow small-grid tiles to choose
RADIX4_SIZES = (12, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512)  # sizes with an alternative line plan


def key(N, precision, les_model):
    return f"{int(N)}-{precision}-{int(les_model)}"


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def apply(plan, N, precision, les_model, path=TABLE):
    """Set the tuned variant of (N, precision, les_model) on ``plan`` if the
    table has one; returns the settings applied (or None)."""
    entry = load_table(path).get(key(N, precision, les_model))
    if entry is None:
        return None
    plan.set_tuning(**{k: entry["settings"][k] for k in KNOBS if k in entry["settings"]})  # fft_wide: N 256, 512
    return entry["settings"]


def candidates(N, precision=None):
    """Every kernel variant of a grid size (the same set in both precisions)."""
    radix = (0, 1) if N in RADIX4_SIZES else (0,)
    for s, x, z, c, f in itertools.product(radix, radix, radix, (0, 1), (0, 1)):
        yield dict(strided_radix4=s, xfwd_radix4=x, z_radix4=z, xcd=c, fuse_x=f)


def time_fft_wide(domain, wide, reps=5):
    """Mean ms of a batched (3 fields) forward + backward generic transform
    with fft_wide = `wide`."""
    d = domain
    d.plan.set_tuning(fft_wide=wide)
    u = d._U
    X = d.fft.forward(u).clone()
    d.fft.backward(X)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        d.fft.forward(u)
        d.fft.backward(X)
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def time_variant(domain, settings, steps=3, warmup=1):
    """Mean ms per RK4 step of the device driver loop (sles_step_drive, the
    path of spectralMethodSolver: CFL dt on the device, U and curl written
    at the end of the call) with `settings` on the plan; the initial state
    is restored afterwards."""
    import ctypes
    import math

    from . import fused

    d = domain
    d.plan.set_tuning(**settings)
    u0 = d.U_hat.clone()
    state = torch.zeros(8, dtype=torch.float64, device=d.device)
    state[2] = math.nan  # t starts at the first dt
    state[6] = math.inf  # no output pauses
    state[7] = math.inf  # no final time
    prm = fused.params(d, cfl=0.5)
    w = fused.work(d)  # sized for this variant's layout

    def drive(n):
        _lib.call("sles_step_drive", d.plan.handle, fused.SCHEME_RK4, ctypes.byref(prm), 0.3, int(n),
                  d.U_hat.data_ptr(), d.U_hat0.data_ptr(), d.U_hat1.data_ptr(), None, d._U.data_ptr(),
                  d.curl.data_ptr(), state.data_ptr(), w.data_ptr(), d.stream)

    drive(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    drive(steps)
    e1.record()
    e1.synchronize()
    d.U_hat.copy_(u0)
    return e0.elapsed_time(e1) / steps


def tune(N, precision="fp64", les="smagorinsky", steps=3, path=TABLE, verbose=True):
    from .initial import isotropic_spectrum, set_spectral_state
    from .initialize_domain import InitializeDomain

    model = {"none": None, "smagorinsky": "Smagorinsky", "vreman": "Vreman"}[les.lower()]
    kw = dict(LES_ON=1, LES_model=model, Cs=0.17) if model else {}
    d = InitializeDomain(N=N, nu=1e-4, method="RK4", precision=precision, tuned=False, **kw)
    set_spectral_state(d, isotropic_spectrum(N, seed=0, dtype=d.cdtype))
    results = []
    for st in candidates(N, precision):
        ms = time_variant(d, st, steps=steps)
        results.append((ms, st))
        if verbose:
            print(f"[autotune] N={N} {precision} {les}: {st} -> {ms:.3f} ms/step", flush=True)
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    best = dict(best)
    default_ms = next(ms for ms, st in results if st == DEFAULTS)
    # the workspace layout, with the best kernel variants (the default
    # padding of the plan was on while those were timed)
    if N >= PAD_MIN_N:
        pads = [(time_variant(d, dict(best, pad=pad), steps=steps), pad) for pad in PADS]
        if verbose:
            for ms, pad in pads:
                print(f"[autotune] N={N} {precision} {les}: pad={pad} -> {ms:.3f} ms/step", flush=True)
        best_ms, best["pad"] = min(pads)
    else:
        best["pad"] = 0
    # the strided passes' tile width on small grids (one or four tiles per
    # CU), with the best variants and padding
    if N in NARROW_SIZES:
        narrow = [(time_variant(d, dict(best, narrow=w), steps=steps), w) for w in (0, 1)]
        if verbose:
            for ms, w in narrow:
                print(f"[autotune] N={N} {precision} {les}: narrow={w} -> {ms:.3f} ms/step", flush=True)
        best_ms, best["narrow"] = min(narrow)
    if N in WIDE_SIZES:
        d.plan.set_tuning(**best)
        wide = [(time_fft_wide(d, w), w) for w in range(4)]
        if verbose:
            for ms, w in wide:
                print(f"[autotune] N={N} {precision} generic transforms fft_wide={w} -> {ms:.3f} ms", flush=True)
        best["fft_wide"] = min(wide)[1]
    table = load_table(path)
    table[key(N, precision, d.sgs_model)] = {
        "settings": best,
        "ms_per_step": round(best_ms, 4),
        "default_ms_per_step": round(default_ms, 4),
        "device": torch.cuda.get_device_name(),
        "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S"),
    }
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return best, best_ms, results


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs="+", default=[256])
    ap.add_argument("--precision", nargs="+", default=["fp64", "fp32"])
    ap.add_argument("--les", nargs="+", default=["smagorinsky"])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--table", default=TABLE)
    a = ap.parse_args(argv)
    for N in a.n:
        for prec in a.precision:
            for les in a.les:
                best, ms, _ = tune(N, prec, les, steps=a.steps, path=a.table)
                print(f"[autotune] best N={N} {prec} {les}: {best} {ms:.3f} ms/step", flush=True)


if __name__ == "__main__":
    main()
