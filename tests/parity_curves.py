"""Step-by-step parity of the fused FP64 GPU step against the oracle on the
decaying isotropic LES workload (seeded k^4 exp(-2 (k/4)^2) field, u_rms = 1,
Smagorinsky Cs = 0.17, nu = 1e-4, RK4, CFL rule: first dt 0.3, then 0.5).

Each side follows the reference loop with its own dt (the GPU's from its
device CFL reduction, the oracle's from compute_dt on its own U); every
`every` steps it prints the velocity rel-L2 (FP64 sums over the (3, N, N, N)
field), the relative dt difference and the largest relative difference of
the calculate_Ek bins (src/SM_solver.py:154-192): against the peak bin over
all bins, and relative over the bins above 1e-8 of the peak.  With --driver
the GPU side advances through sles_step_drive calls of `every` steps (the
path of spectralMethodSolver and bench.py: padded copies of the state, the
x-inverse fused across steps, U written at each call's end); otherwise one
sles_step call per step.

    python tests/parity_curves.py --n 64 --steps 500 --every 25 [--driver]

Test infrastructure (it runs the oracle): not collected by pytest; run on a
machine with a GPU.
"""

import argparse
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "spectral-les_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import sles_oracle as O  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--les", default="Smagorinsky")
    ap.add_argument("--driver", action="store_true", help="GPU steps through sles_step_drive calls of --every steps")
    a = ap.parse_args(argv)

    from sles import fused
    from sles.initial import initial_velocity, isotropic_spectrum
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import Ek_plot_variables, calculate_Ek

    N = a.n
    kw = dict(nu=1e-4, method="RK4", LES_ON=1, LES_model=a.les, Cs=0.17, CFL=0.5)
    g = InitializeDomain(N=N, **kw)
    Uh = isotropic_spectrum(N, k0=4.0, seed=a.seed, urms=1.0)
    u = g.fft.backward(Uh, normalize=True).cpu().numpy().copy()
    del Uh
    initial_velocity(g, *u)
    g.U_hat1.copy_(g.U_hat)
    o = O.Domain(N=N, fft_workers=min(O.cpu_count(), 16), **kw)
    O.initial_velocity(o, *u)
    o.U_hat1[:] = o.U_hat
    E = Ek_plot_variables(g, path=None)
    Eo = O.EkBins(o)
    if a.driver:
        import ctypes
        import math

        from sles import _lib

        st = torch.zeros(8, dtype=torch.float64, device=g.device)
        st[2] = math.nan  # t starts at the first dt
        st[6] = math.inf  # no output pauses
        st[7] = math.inf  # no final time
        prm = fused.params(g, cfl=0.5)
        w = fused.work(g)

        def drive(n):
            _lib.call("sles_step_drive", g.plan.handle, fused.SCHEME_RK4, ctypes.byref(prm), 0.3, int(n),
                      g.U_hat.data_ptr(), g.U_hat0.data_ptr(), g.U_hat1.data_ptr(), None, g._U.data_ptr(),
                      g.curl.data_ptr(), st.data_ptr(), w.data_ptr(), g.stream)
    path = f"sles_step_drive, {a.every} steps per call" if a.driver else "one sles_step call per step"
    print(f"# N={N} {a.les} Cs=0.17 nu=1e-4 RK4 CFL 0.3 then 0.5, seed {a.seed}, FP64 GPU (fused step, device "
          f"dt; {path}; tuning {g.plan.get_tuning()}) vs oracle (numpy FP64, own dt)")
    print("# E(k): calculate_Ek bins; |dE|/peak over all bins, and the largest relative difference over the bins")
    print("# holding >= 1e-8 of the peak bin (the others hold round-off of modes far below the largest)")
    print("# step   t             dt_gpu        |dt_g/dt_o-1|  vel_rel_L2  Ek_absdiff/peak  nres  Ek_rel_resolved")
    t = 0.0
    t0 = time.time()
    worst = 0.0
    done = 0  # GPU steps taken (driver mode)
    for k in range(1, a.steps + 1):
        cfl = 0.3 if k == 1 else 0.5
        sample = k % a.every == 0 or k == a.steps
        if not a.driver:
            fused.rk4_step(g, dt_device=True, cfl=cfl)
            dtg = float(g._scalars[1])
        elif sample:
            drive(k - done)
            done = k
            dtg = float(st[1])
            assert int(st[4]) == k and int(st[3]) == 0, st.tolist()
        o.dt = O.compute_dt(o.U, O.find_max_velocity(o.U), o.nu, dx=o.dx, CFL=cfl)
        O.rk4_step(o)
        t += o.dt
        if sample:
            Ug = g._U.cpu().numpy()
            Uo = np.array(o.U)
            r = float(np.linalg.norm((Ug - Uo).ravel()) / np.linalg.norm(Uo.ravel()))
            worst = max(worst, r)
            calculate_Ek(g, E)
            O.calculate_Ek(o, Eo)
            peak = float(Eo.temp_Ek.max())
            ek_abs = float(np.max(np.abs(E.temp_Ek - Eo.temp_Ek))) / peak
            res = Eo.temp_Ek >= 1e-8 * peak
            ek = float(np.max(np.abs(E.temp_Ek[res] - Eo.temp_Ek[res]) / Eo.temp_Ek[res]))
            print(f"{k:6d}  {t:.6e}  {dtg:.6e}  {abs(dtg / o.dt - 1):.3e}  {r:.3e}  {ek_abs:.3e}  {int(res.sum()):3d}  "
                  f"{ek:.3e}", flush=True)
            print(f"[parity] N={N} step {k}/{a.steps} rel-L2 {r:.3e} ({time.time() - t0:.0f} s)", file=sys.stderr,
                  flush=True)
    print(f"# max velocity rel-L2 over the sampled steps: {worst:.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
