"""sles_step_noalloc (the fused step with its workspace in the library's
static device scratch) against sles_step with a caller workspace: the same
kernels on the same inputs, so every output must agree bit for bit.  Two
calls in a row per case, the second on a smaller grid (the scratch keeps the
first call's contents), and a grid whose workspace exceeds the scratch is
rejected on the host before anything is launched."""

import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _domain(N, precision, les, seed):
    from sles.initial import isotropic_spectrum, set_spectral_state
    from sles.initialize_domain import InitializeDomain

    kw = dict(LES_ON=1, LES_model=les, Cs=0.17) if les else {}
    d = InitializeDomain(N=N, nu=1e-3, method="RK4", time_stepSize=2e-3, precision=precision, **kw)
    set_spectral_state(d, isotropic_spectrum(N, k0=3.0, seed=seed, urms=1.0, dtype=d.cdtype))
    return d


def _call(name, d, scheme, prm, aux, work=None):
    from sles import _lib

    args = [d.plan.handle, scheme, ctypes.byref(prm), d.U_hat.data_ptr(), d.U_hat0.data_ptr(),
            d.U_hat1.data_ptr(), aux.data_ptr() if aux is not None else None, d._U.data_ptr(),
            d.curl.data_ptr(), d._scalars.data_ptr()]
    if work is not None:
        args.append(work.data_ptr())
    args.append(d.stream)
    _lib.call(name, *args)


def _step_pair(N, precision, les, scheme, seed):
    """One step of `scheme` through sles_step (a) and sles_step_noalloc (b)
    from the same state; returns both domains."""
    from sles import fused

    a = _domain(N, precision, les, seed)
    b = _domain(N, precision, les, seed)
    aux_a = aux_b = None
    if scheme == fused.SCHEME_ETD2:
        aux_a = torch.zeros_like(a.U_hat)
        fused.nonlinear(a, aux_a)  # F_{n-1} = F(u_0)
        aux_b = aux_a.clone()
    for d in (a, b):
        d.U_hat1.copy_(d.U_hat)
    ddt = scheme in (fused.SCHEME_RK4, fused.SCHEME_RK3)
    prm = fused.params(a, dt_device=ddt, cfl=0.5)
    _call("sles_step", a, scheme, prm, aux_a, work=fused.work(a))
    _call("sles_step_noalloc", b, scheme, prm, aux_b)
    torch.cuda.synchronize()
    return a, b, aux_a, aux_b


CASES = [
    # (first N, second N, precision, LES model, scheme)
    (64, 32, "fp64", "Smagorinsky", 0),
    (96, 48, "fp64", "Vreman", 0),
    (96, 64, "fp32", "Smagorinsky", 0),
    (64, 24, "fp32", None, 0),
    (64, 32, "fp64", "Vreman", 2),
    (48, 16, "fp32", "Smagorinsky", 1),
]


@pytest.mark.parametrize("n1,n2,precision,les,scheme", CASES)
def test_noalloc_step_matches_step(n1, n2, precision, les, scheme):
    for N, seed in ((n1, 1), (n2, 2)):
        print(f"[no-alloc] N={N} {precision} les={les} scheme={scheme}", flush=True)
        a, b, xa, xb = _step_pair(N, precision, les, scheme, seed)
        for name in ("U_hat", "U_hat0", "U_hat1", "_U", "curl"):
            assert torch.equal(getattr(a, name), getattr(b, name)), (N, name)
        if xa is not None:
            assert torch.equal(xa, xb), (N, "aux")
        if scheme != 2:
            assert torch.equal(a._scalars[:2], b._scalars[:2]), N
        assert torch.isfinite(b.U_hat).all()


def test_noalloc_rejects_grids_above_the_scratch():
    from sles import _lib, fused

    d = _domain(128, "fp64", "Smagorinsky", 3)  # 146 MiB of workspace > 64 MiB
    before = d.U_hat.clone()
    prm = fused.params(d, dt=1e-3)
    with pytest.raises(RuntimeError, match="scratch"):
        _call("sles_step_noalloc", d, fused.SCHEME_RK4, prm, None)
    torch.cuda.synchronize()
    assert torch.equal(d.U_hat, before)
    assert int(_lib.lib.sles_step_work_bytes(d.plan.handle, d.sgs_model)) > (64 << 20)
