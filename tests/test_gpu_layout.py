"""The padded workspace layout (SLES_TUNE_PAD): kx planes of the fused
step's workspace -- and of the RK-state copies that the device driver steps
on -- start every N (N/2+1) values + pad 128-byte lines.  Only addresses
change, so every result must equal the compact layout's bit for bit: the
single step (workspace padded, caller's arrays in place), the driver loop
(state copied into the padded layout and back), the batched ensemble driver
and the physical output.  A driver call that takes no step leaves U_hat0,
U_hat1 and aux as they were."""

import ctypes
import math

import numpy as np
import pytest
import torch

from helpers import random_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _domain(N, precision, les, method, pad, seed=3):
    from sles.initial import initial_velocity
    from sles.initialize_domain import InitializeDomain

    kw = dict(LES_ON=1, LES_model=les, Cs=0.17) if les else {}
    d = InitializeDomain(N=N, nu=0.005, method=method, precision=precision, fused=True, tuned=False, **kw)
    d.plan.set_tuning(pad=pad)
    u = random_field(N, seed, k_cut=max(2.0, N / 6))
    initial_velocity(d, u[0], u[1], u[2])
    d.U_hat1.copy_(d.U_hat)
    return d


def _state(device):
    st = torch.zeros(8, dtype=torch.float64, device=device)
    st[2] = math.nan  # t
    st[6] = math.inf  # no pauses
    st[7] = math.inf  # no final time
    return st


def _drive(d, scheme, aux, state, n):
    from sles import _lib, fused

    prm = fused.params(d, cfl=0.5)
    w = fused.work(d)
    _lib.call("sles_step_drive", d.plan.handle, scheme, ctypes.byref(prm), 0.3, int(n), d.U_hat.data_ptr(),
              d.U_hat0.data_ptr(), d.U_hat1.data_ptr(), aux.data_ptr() if aux is not None else None,
              d._U.data_ptr(), d.curl.data_ptr(), state.data_ptr(), w.data_ptr(), d.stream)


def _same(a, b, tag, aux_a=None, aux_b=None):
    for name in ("U_hat", "U_hat0", "U_hat1", "_U", "curl"):
        assert torch.equal(getattr(a, name), getattr(b, name)), (tag, name)
    if aux_a is not None:
        assert torch.equal(aux_a, aux_b), (tag, "aux")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("les", [None, "Smagorinsky", "Vreman"])
@pytest.mark.parametrize("N,pad", [(16, 1), (24, 5), (32, 17), (64, 3)])
def test_drive_rk4_padded_equals_compact(N, pad, les, precision):
    from sles import fused

    a = _domain(N, precision, les, "RK4", 0)
    b = _domain(N, precision, les, "RK4", pad)
    assert b.plan.get_tuning()["pad"] == pad
    sa, sb = _state(a.device), _state(b.device)
    for k, n in enumerate((2, 1, 3)):
        print(f"[layout] drive RK4 N={N} pad={pad} {les} {precision} call {k}", flush=True)
        _drive(a, fused.SCHEME_RK4, None, sa, n)
        _drive(b, fused.SCHEME_RK4, None, sb, n)
        torch.cuda.synchronize()
        _same(a, b, k)
        assert torch.equal(sa, sb)
    assert sb[4].item() == 6  # steps taken


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("method", ["RK3", "ETD2"])
def test_drive_rk3_etd2_padded_equals_compact(method, precision):
    from sles import fused

    N, pad = 24, 9
    a = _domain(N, precision, "Smagorinsky", method, 0)
    b = _domain(N, precision, "Smagorinsky", method, pad)
    sa, sb = _state(a.device), _state(b.device)
    if method == "RK3":
        for k in range(3):
            print(f"[layout] drive RK3 N={N} pad={pad} {precision} call {k}", flush=True)
            _drive(a, fused.SCHEME_RK3, None, sa, 2)
            _drive(b, fused.SCHEME_RK3, None, sb, 2)
            torch.cuda.synchronize()
            _same(a, b, k)
            assert torch.equal(sa, sb)
        return
    # ETD2: an RK4 step that stores F(u_0) in aux, then ETD2 steps reading it
    print(f"[layout] drive ETD2 N={N} pad={pad} {precision}", flush=True)
    _drive(a, fused.SCHEME_RK4, a.F0, sa, 1)
    _drive(b, fused.SCHEME_RK4, b.F0, sb, 1)
    torch.cuda.synchronize()
    _same(a, b, "rk4 start", a.F0, b.F0)
    for k in range(3):
        _drive(a, fused.SCHEME_ETD2, a.F0, sa, 2)
        _drive(b, fused.SCHEME_ETD2, b.F0, sb, 2)
        torch.cuda.synchronize()
        _same(a, b, k, a.F0, b.F0)
        assert torch.equal(sa, sb)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_drive_padded_without_a_step_keeps_rk_arrays(precision):
    """A call whose state is already DONE takes no step: U_hat is copied out
    unchanged and U_hat0 / U_hat1 / aux keep the caller's values (the state
    copies in the workspace hold nothing of this call)."""
    from sles import fused

    d = _domain(32, precision, "Smagorinsky", "RK4", 17)
    st = _state(d.device)
    _drive(d, fused.SCHEME_RK4, None, st, 1)  # fills the workspace copies
    d.U_hat0.fill_(1.5)
    d.U_hat1.fill_(-2.5)
    aux = torch.full_like(d.U_hat, 0.25)
    before = d.U_hat.clone()
    st[3] = 1.0  # status DONE
    _drive(d, fused.SCHEME_RK4, aux, st, 2)
    torch.cuda.synchronize()
    assert torch.equal(d.U_hat, before)
    assert bool((d.U_hat0 == 1.5).all()) and bool((d.U_hat1 == -2.5).all())
    assert bool((aux == 0.25).all())
    assert st[4].item() == 1 and st[3].item() == 1.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("les", [None, "Vreman"])
def test_single_step_with_padded_workspace(les, precision):
    """sles_step keeps the caller's arrays in the reference layout and pads
    only its workspace: every scheme, host and device dt, the RHS forms and
    the adjoint's stage-state recomputation equal the compact layout's."""
    from sles import _lib, fused

    N = 24
    a = _domain(N, precision, les, "RK4", 0)
    b = _domain(N, precision, les, "RK4", 13)
    for k, scheme in enumerate((fused.SCHEME_RK4, fused.SCHEME_RK3, fused.SCHEME_RK4)):
        fused.step(a, scheme, dt_device=(k == 2), dt=2e-3)
        fused.step(b, scheme, dt_device=(k == 2), dt=2e-3)
        torch.cuda.synchronize()
        _same(a, b, ("step", k))
    for fn in (fused.rhs, fused.nonlinear):
        oa, ob = torch.empty_like(a.U_hat), torch.empty_like(b.U_hat)
        fn(a, oa)
        fn(b, ob)
        torch.cuda.synchronize()
        assert torch.equal(oa, ob), fn.__name__
    # ETD2 through sles_step (aux read and written)
    fa, fb = torch.zeros_like(a.U_hat), torch.zeros_like(b.U_hat)
    for scheme in (fused.SCHEME_RK4, fused.SCHEME_ETD2, fused.SCHEME_ETD2):
        fused.step(a, scheme, aux=fa, dt=2e-3)
        fused.step(b, scheme, aux=fb, dt=2e-3)
    torch.cuda.synchronize()
    _same(a, b, "etd2", fa, fb)
    # the adjoint's RK4 stage states (written in the caller's layout)
    prm = fused.params(a, dt=2e-3)
    sts = [torch.empty((9,) + tuple(a.U_hat.shape[1:]), dtype=a.cdtype, device=a.device) for _ in range(2)]
    for d, st in zip((a, b), sts):
        _lib.call("sles_rk4_stage_states", d.plan.handle, ctypes.byref(prm), d.U_hat.data_ptr(),
                  d.U_hat0.data_ptr(), d.U_hat1.data_ptr(), st.data_ptr(), fused.work(d).data_ptr(), d.stream)
    torch.cuda.synchronize()
    assert torch.equal(sts[0], sts[1])
    assert torch.equal(a.U_hat, b.U_hat)


def test_work_bytes_follow_the_padding():
    from sles import _lib

    d = _domain(32, "fp64", "Smagorinsky", "RK4", 0)
    nc = 32 * 32 * 17
    w0 = int(_lib.lib.sles_step_work_bytes(d.plan.handle, 1))
    assert w0 == 9 * nc * 16 + 8 * (32 * 32 + 64)
    d.plan.set_tuning(pad=2)
    plane = 32 * 17 + 2 * 128 // 16
    w2 = int(_lib.lib.sles_step_work_bytes(d.plan.handle, 1))
    fields = 9 * 32 * plane * 16 + 8 * (32 * 32 + 64)
    assert w2 == (fields + 255) // 256 * 256 + 256 + 12 * 32 * plane * 16
    with pytest.raises(RuntimeError, match="PAD"):
        d.plan.set_tuning(pad=65)
    assert d.plan.get_tuning()["pad"] == 2


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_drive_padded_equals_compact(precision):
    """sles_step_drive_batch with each member's state copies in its own
    workspace: the same states as the compact layout, member by member."""
    from sles.ensemble import Ensemble

    groups = ((16, 3), (32, 2))
    a = Ensemble(groups, precision=precision, seed=7, batched=True, batch=8)
    b = Ensemble(groups, precision=precision, seed=7, batched=True, batch=8)
    for g, pad in zip(b.groups, (3, 17)):
        for d in g.members:
            d.plan.set_tuning(pad=pad)
        g.plan.set_tuning(pad=pad)
        g.realloc_work()
    for k in range(3):
        print(f"[layout] batched drive {precision} step {k}", flush=True)
        a.step_eager(1)
        b.step_eager(1)
        torch.cuda.synchronize()
        for da, db in zip(a.members, b.members):
            assert torch.equal(da.U_hat, db.U_hat), (k, da.N)
            assert torch.equal(da.U_hat0, db.U_hat0), (k, da.N)
            assert torch.equal(da.U_hat1, db.U_hat1), (k, da.N)
        assert torch.equal(a.states(), b.states())
    assert bool((b.states()[:, 4] == 3).all())
    assert np.isfinite(b.members[-1].U_hat.abs().max().item())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("les", [None, "Smagorinsky"])
@pytest.mark.parametrize("N", [64, 96, 128])
def test_narrow_tiles_drive_bitwise(N, les, precision):
    """SLES_TUNE_NARROW (four strided tiles per CU instead of one) moves
    columns between blocks but not the arithmetic on them: the driver loop
    over several calls gives the same states bit for bit."""
    from sles import fused

    a = _domain(N, precision, les, "RK4", 5)
    b = _domain(N, precision, les, "RK4", 5)
    b.plan.set_tuning(narrow=1)
    assert b.plan.get_tuning()["narrow"] == 1 and a.plan.get_tuning()["narrow"] == 0
    sa, sb = _state(a.device), _state(b.device)
    for k, n in enumerate((3, 1)):
        print(f"[layout] narrow tiles N={N} {les} {precision} call {k}", flush=True)
        _drive(a, fused.SCHEME_RK4, None, sa, n)
        _drive(b, fused.SCHEME_RK4, None, sb, n)
        torch.cuda.synchronize()
        _same(a, b, k)
        assert torch.equal(sa, sb)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("pad", [0, 5])
def test_drive_stopping_mid_call_equals_shorter_call(pad, precision):
    """RK4 runs each step's last x-forward together with the next step's
    first x-inverse inside a driver call.  A call whose final time stops it
    after 2 of 4 steps must leave the state, U and curl exactly as a call of
    those 2 steps does: the steps it skips hand on the x-inverse of the
    unchanged U_hat, and the call's last step writes U and curl from it."""
    from sles import fused

    a = _domain(32, precision, "Smagorinsky", "RK4", pad)
    b = _domain(32, precision, "Smagorinsky", "RK4", pad)
    sa, sb = _state(a.device), _state(b.device)
    print(f"[layout] stop mid-call pad={pad} {precision}", flush=True)
    _drive(a, fused.SCHEME_RK4, None, sa, 2)
    torch.cuda.synchronize()
    sb[7] = sa[2].item()  # final time: reached by the second step
    _drive(b, fused.SCHEME_RK4, None, sb, 4)
    torch.cuda.synchronize()
    _same(a, b, "stop")
    assert sb[4].item() == 2 and sb[3].item() == 1.0  # steps taken, status DONE
    assert sb[2].item() == sa[2].item()  # t
