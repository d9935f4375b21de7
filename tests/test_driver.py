"""Driver-level behaviour: CFL rule, stop conditions and failure guards
(SM_solver.py:18-83, 245-277), the seeded isotropic initial field, and the
E(k) CSV output."""

import math
import os

import numpy as np
import pytest
import torch

from oracle import sles_oracle as O


def test_compute_dt_rule_and_exit_cpu():
    from sles.SM_solver import DtTooSmall, compute_dt

    dx = 2 * math.pi / 32
    assert compute_dt(None, 2.0, 0.01, dx=dx, CFL=0.5) == min(0.5 * dx / 2.0, 0.1 * dx ** 2 / 0.01)
    assert compute_dt(None, 2.0, 0.0, dx=dx, CFL=0.5) == 0.5 * dx / 2.0
    with pytest.raises(DtTooSmall):
        compute_dt(None, 1e9, 0.01, dx=dx, CFL=0.5)
    assert isinstance(DtTooSmall("x"), SystemExit)
    # NaN max velocity gives a NaN dt (the driver's break condition), not an exit
    assert compute_dt(None, float("nan"), 0.01, dx=dx, CFL=0.5) != compute_dt(None, float("nan"), 0.01, dx=dx,
                                                                              CFL=0.5)


def test_get_index_partition():
    from sles.utilites import get_index

    for N in (7, 64, 129):
        for n in (1, 2, 3, 8):
            parts = [get_index(N, n, r) for r in range(n)]
            assert parts[0][0] == 0 and parts[-1][1] == N
            assert all(parts[i][1] == parts[i + 1][0] for i in range(n - 1))
            sizes = [b - a for a, b in parts]
            assert max(sizes) - min(sizes) <= 1 and sizes == sorted(sizes, reverse=True)


gpu = pytest.mark.gpu


def _read_csv(path):
    import csv

    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    header = rows[0]
    cols = {h: np.array([float(r[j]) for r in rows[1:] if r[j] != ""]) for j, h in enumerate(header)}
    return header, cols


def _driver_case(tmp_path, method, chunk, graph, N=16, final_time=1.0, file_out_time=0.0, No_Ek_plot=2,
                 les="Smagorinsky", precision="fp64", CFL=0.5):
    """(GPU domain after spectralMethodSolver, its return value, oracle domain
    after the oracle loop, oracle (counter1, events, dts), output dir)."""
    from helpers import oracle_loop, random_field
    from sles.initial import initial_velocity
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import spectralMethodSolver

    u = random_field(N, 31, k_cut=3.0)
    kw = dict(nu=0.01, final_time=final_time, method=method, CFL=CFL)
    if les:
        kw.update(LES_ON=1, LES_model=les, Cs=0.17)
    g = InitializeDomain(N=N, precision=precision, **kw)
    initial_velocity(g, *u)
    out = tmp_path / f"{method}-{chunk}-{int(graph)}"
    out.mkdir()
    cwd = os.getcwd()
    os.chdir(out)
    try:
        steps = spectralMethodSolver(g, file_out_time, No_Ek_plot=No_Ek_plot, chunk=chunk, graph=graph)
    finally:
        os.chdir(cwd)
    o = O.Domain(N=N, **kw)
    O.initial_velocity(o, *u)
    ref = oracle_loop(o, file_out_time, No_Ek_plot)
    return g, steps, o, ref, out


@gpu
@pytest.mark.parametrize("method,chunk,graph,nek,cfl", [("RK4", 1, False, 2, 0.5), ("RK4", 4, False, 3, 0.5),
                                                        ("RK4", 10, True, 1, 0.5), ("RK3", 3, False, 2, 0.3),
                                                        ("ETD2", 4, True, 1, 0.2)])
def test_driver_loop_matches_oracle_loop(tmp_path, method, chunk, graph, nek, cfl):
    """spectralMethodSolver on the device (dt, t and the stop rules per step
    on the GPU, chunks of steps, optionally replayed HIP graphs) against the
    oracle restatement of the reference loop: the same number of steps, the
    state to 1e-10, and the E(k) CSV written at the same loop tops (same
    t labels, same columns, bins to 1e-10)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from helpers import rel

    g, steps, o, (n, events, dts), out = _driver_case(tmp_path, method, chunk, graph, No_Ek_plot=nek, CFL=cfl)
    assert steps == n and n > 8
    assert rel(g.U_hat, o.U_hat) < 1e-10
    assert rel(torch.stack(g.U), np.array(o.U)) < 1e-10
    assert math.isclose(g.dt, o.dt, rel_tol=1e-10)
    header, cols = _read_csv(out / "Ek_plot_data.csv")
    ek = [e for e in events if e[0] == "Ek"]
    assert header == [h for _, t, _ in ek for h in (f"k-{t:g}", f"Ek-{t:g}")]
    for _, t, (k, e) in ek:
        assert np.allclose(cols[f"k-{t:g}"], k, rtol=1e-12, atol=0)
        assert np.allclose(cols[f"Ek-{t:g}"], e, rtol=1e-10, atol=1e-300)
    loop = g._drive
    if graph:
        assert loop.replays > 0
    # one read of the device loop state per chunk (plus one per output pause)
    assert loop.host_reads <= math.ceil(n / chunk) + 2 * len(ek) + 4


@gpu
def test_driver_snapshots_match_oracle_loop(tmp_path):
    """file_out_time > 0: a raw-binary + XDMF snapshot of U at every loop top
    where output_h5 would write one (src/SM_solver.py:40-42,280-302), with
    the reference's file numbering t // file_out_time, holding U of that
    state."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import xml.etree.ElementTree as ET

    N = 16
    g, steps, o, (n, events, dts), out = _driver_case(tmp_path, "RK4", 4, False, N=N, final_time=0.6,
                                                      file_out_time=0.2, No_Ek_plot=0)
    snaps = [e for e in events if e[0] == "snapshot"]
    assert len(snaps) >= 3
    files = sorted(os.listdir(out / "dumpfile"))
    assert files == sorted({f"dump{t // 0.2:g}.{x}" for _, t, _ in snaps for x in ("bin", "xdmf")})
    for _, t, U in snaps:
        tag = f"{t // 0.2:g}"
        got = np.fromfile(out / "dumpfile" / f"dump{tag}.bin", dtype="<f8").reshape(3, N, N, N)
        assert np.linalg.norm(got - U) / np.linalg.norm(U) < 1e-10
        root = ET.parse(out / "dumpfile" / f"dump{tag}.xdmf").getroot()
        attrs = root.findall(".//Attribute")
        assert [a.get("Name") for a in attrs] == ["u", "v", "w"]
        seeks = [int(a.find("DataItem").get("Seek")) for a in attrs]
        assert seeks == [0, 8 * N ** 3, 16 * N ** 3]
        assert root.find(".//Topology").get("Dimensions") == f"{N} {N} {N}"
        assert root.find(".//Time").get("Value") == str(round(t // 0.2))


@gpu
def test_driver_dt_too_small_raises():
    """compute_dt's exit for dt < 1e-6 (src/SM_solver.py:273-276) ends the
    device loop with DtTooSmall, before any step."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from helpers import random_field
    from sles.initial import initial_velocity
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import DtTooSmall, spectralMethodSolver

    g = InitializeDomain(N=16, nu=0.01, final_time=1.0, method="RK4")
    initial_velocity(g, *(1e7 * random_field(16, 3, k_cut=3.0)))
    before = g.U_hat.clone()
    with pytest.raises(DtTooSmall):
        spectralMethodSolver(g, 0.0, No_Ek_plot=0)
    assert torch.equal(g.U_hat, before)


def test_stale_ek_csv_is_removed(tmp_path):
    """Ek_plot_variables removes an Ek_plot_data.csv left by an earlier run
    (src/SM_solver.py:114-116), so the new run's file holds its own columns
    only."""
    from sles.SM_solver import Ek_plot_variables

    class D:
        N = 8
        kx = np.fft.fftfreq(8, 1.0 / 8)
        kz = np.abs(kx[:5])
        device = "cpu"

    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with open("Ek_plot_data.csv", "w") as f:
            f.write("k-0.5,Ek-0.5\n1,2\n")
        Ek_plot_variables(D())
        assert not os.path.exists("Ek_plot_data.csv")
    finally:
        os.chdir(cwd)


@gpu
def test_driver_nan_guard_breaks():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sles.initialize_domain import InitializeDomain
    from sles.SM_solver import spectralMethodSolver

    g = InitializeDomain(N=16, nu=0.01, final_time=1.0, method="RK4")
    g.U_hat.fill_(float("nan"))
    g._U.fill_(1.0)  # finite first dt, NaN after the first step
    assert spectralMethodSolver(g, 0.0, No_Ek_plot=0) == 0


@gpu
def test_isotropic_initial_field():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sles.initial import isotropic_spectrum

    N = 32
    Uh = isotropic_spectrum(N, k0=4.0, seed=5, urms=1.0)
    Uh2 = isotropic_spectrum(N, k0=4.0, seed=5, urms=1.0)
    assert torch.equal(Uh, Uh2)  # seeded, deterministic
    U = Uh.cpu().numpy()
    u = np.fft.irfftn(U, s=(N, N, N), axes=(1, 2, 3))
    assert math.isclose(np.sqrt(np.mean(u * u)), 1.0, rel_tol=1e-12)  # u_rms over components
    # the stored spectrum is exactly the transform of a real field
    assert np.max(np.abs(np.fft.rfftn(u, axes=(1, 2, 3)) - U)) < 1e-9 * np.max(np.abs(U))
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    K = np.array(np.meshgrid(k, k, kz, indexing="ij"))
    div = np.sum(K * U, axis=0)
    assert np.max(np.abs(div)) < 1e-9 * np.max(np.abs(U))  # solenoidal
    kmag = np.sqrt(np.sum(K * K, 0))
    assert np.all(np.abs(U[:, kmag == 0]) == 0)
    assert np.all(U[:, np.abs(K[0]) == N // 2] == 0)
    keep = np.all(np.abs(K) < 2.0 / 3.0 * (N // 2 + 1), axis=0)
    assert np.all(U[:, ~keep] == 0)
    # |U_hat|^2 per mode ~ k^2 exp(-2 (k/k0)^2) away from the kz = 0 plane
    sel = (kz[None, None, :] * np.ones_like(kmag) > 0) & keep & (kmag > 0)
    ratio = np.sum(np.abs(U) ** 2, 0)[sel] / (kmag[sel] ** 2 * np.exp(-2 * (kmag[sel] / 4.0) ** 2))
    assert np.std(ratio) / np.mean(ratio) < 1e-6
