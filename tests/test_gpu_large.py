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


@pytest.mark.parametrize("N", [128, 256])
@pytest.mark.parametrize("les", ["Smagorinsky", "Vreman"])
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
