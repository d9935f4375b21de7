"""Large-grid checks of the fused step (kernel configurations of the
benchmark sizes): FP64 fused vs the FP64 operator path, and the FP32 form vs
FP64, on the seeded isotropic field."""

import pytest
import torch

from helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _domain(N, precision, fused, les="Smagorinsky"):
    from sles.initial import isotropic_spectrum, set_spectral_state
    from sles.initialize_domain import InitializeDomain

    d = InitializeDomain(N=N, nu=1e-4, method="RK4", LES_ON=1, LES_model=les, Cs=0.17, time_stepSize=1e-3,
                         precision=precision, fused=fused)
    Uh = isotropic_spectrum(N, k0=4.0, seed=3, urms=1.0)  # FP64, cast for FP32
    set_spectral_state(d, Uh.to(d.cdtype))
    return d


@pytest.mark.parametrize(
    "N,les",
    [(128, "Smagorinsky"), (128, "Vreman"), (256, "Smagorinsky"), (256, "Vreman"),
     (192, "Smagorinsky"), (384, "Vreman")],  # 192, 384: 12-point line plans (radices 4, 4, 4, 3 [, 2])
)
def test_fused_vs_operator_path_fp64(N, les):
    import sles.RK4_method as R

    a = _domain(N, "fp64", True, les)
    b = _domain(N, "fp64", False, les)
    for _ in range(2):
        R.solveForNextTimestep(a)
        R.solveForNextTimestep(b)
    assert rel(a.U_hat, b.U_hat) < 1e-12
    assert rel(a.curl, b.curl) < 1e-12


@pytest.mark.parametrize("N", [128, 256])
@pytest.mark.parametrize("les", ["Smagorinsky", "Vreman"])
def test_fp32_form_vs_fp64(N, les):
    import sles.RK4_method as R

    a = _domain(N, "fp64", True, les)
    b = _domain(N, "fp32", True, les)
    for _ in range(3):
        R.solveForNextTimestep(a)
        R.solveForNextTimestep(b)
    assert rel(b.U_hat.to(torch.complex128), a.U_hat) < 1e-5
    assert rel(b.curl.double(), a.curl) < 1e-5


def _div_and_energy(U_hat, N):
    """max |K . U_hat| / max(|K| |U_hat|) and the FP64 kinetic energy sum."""
    dev = U_hat.device
    k = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=dev)
    kz = torch.arange(N // 2 + 1, dtype=torch.float64, device=dev)
    U = U_hat.to(torch.complex128)
    div = k[:, None, None] * U[0] + k[None, :, None] * U[1] + kz[None, None, :] * U[2]
    kmag = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    scale = float((kmag * U.abs().pow(2).sum(0).sqrt()).max())
    w = torch.full_like(kmag, 2.0)
    w[..., 0] = 1.0
    w[..., -1] = 1.0
    energy = float((w * U.abs().pow(2).sum(0)).sum())
    return float(div.abs().max()) / scale, energy


@pytest.mark.parametrize("precision,les", [("fp64", "Smagorinsky"), ("fp32", "Vreman")])
def test_256_steps_solenoidal_and_dissipative(precision, les):
    """Size-independent properties at the benchmark size: after each fused RK4
    step (device CFL dt) the velocity stays divergence-free (projection) and
    the kinetic energy decreases (nu > 0, eddy viscosity >= 0, no forcing)."""
    from sles import fused

    N = 256
    d = _domain(N, precision, True, les)
    div0, e_prev = _div_and_energy(d.U_hat, N)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for _ in range(4):
        fused.rk4_step(d, dt_device=True, cfl=0.5)
        div, e = _div_and_energy(d.U_hat, N)
        assert div < tol, div
        assert e < e_prev
        e_prev = e
    assert float(d._scalars[1]) >= 1e-6  # the step was taken


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_512_fft_round_trip(precision):
    """Generic 3D r2c then c2r at 512^3 returns the input (FP64 1e-13, FP32 5e-6
    relative), and the r2c of a real field is Hermitian on the kz = 0 plane
    (to the same tolerance)."""
    from sles import _lib

    N = 512
    prec = _lib.SLES_FP64 if precision == "fp64" else _lib.SLES_FP32
    rdt = torch.float64 if precision == "fp64" else torch.float32
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    plan = _lib.Plan(N, prec, torch.device("cuda"))
    g = torch.Generator(device="cuda").manual_seed(5)
    u = torch.randn(1, N, N, N, generator=g, device="cuda", dtype=rdt)
    U = torch.empty(1, N, N, N // 2 + 1, dtype=cdt, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    _lib.call("sles_fft_r2c", plan.handle, u.data_ptr(), U.data_ptr(), 1, 1.0, s)
    plane = U[0, :, :, 0]
    neg = (-torch.arange(N, device="cuda")) % N
    herm = float((plane - plane[neg][:, neg].conj()).abs().max() / plane.abs().max())
    v = torch.empty_like(u)
    _lib.call("sles_fft_c2r", plan.handle, U.data_ptr(), v.data_ptr(), U.data_ptr(), 1, 1.0 / N ** 3, s)
    torch.cuda.synchronize()
    err = float((v - u).abs().max() / u.abs().max())
    tol = 1e-13 if precision == "fp64" else 5e-6
    assert err < tol, err
    assert herm < tol, herm


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_512_z_plans_agree(precision):
    """At 512 points the radix-4 z lines span two wavefronts (one line per
    block, ping-pong exchange buffers, block barriers) while the radix-8 lines
    fit one wavefront: one Smagorinsky RHS at 512^3 through either z plan
    agrees to rounding (FP64 1e-12, FP32 1e-5 relative)."""
    from sles import fused
    from sles.initialize_domain import InitializeDomain

    N = 512
    d = InitializeDomain(N=N, nu=1e-4, method="RK4", LES_ON=1, LES_model="Smagorinsky", Cs=0.17,
                         time_stepSize=1e-3, precision=precision, tuned=False)
    g = torch.Generator(device="cuda").manual_seed(11)
    shape = (3, N, N, N // 2 + 1)
    d.U_hat.copy_(torch.complex(torch.randn(shape, generator=g, device="cuda", dtype=d.rdtype),
                                torch.randn(shape, generator=g, device="cuda", dtype=d.rdtype)))
    out4 = torch.empty_like(d.U_hat)
    out8 = torch.empty_like(d.U_hat)
    d.plan.set_tuning(z_radix4=1)
    fused.rhs(d, out4)
    d.plan.set_tuning(z_radix4=0)
    fused.rhs(d, out8)
    torch.cuda.synchronize()
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert torch.isfinite(out4).all()
    num = sum(float((out4[i] - out8[i]).abs().double().pow(2).sum()) for i in range(3))
    den = sum(float(out8[i].abs().double().pow(2).sum()) for i in range(3))
    assert (num / den) ** 0.5 < tol
