"""Ensembles of uneven grids: the batched launches (one launch per pass over
all members of a grid size) give the same states, bit for bit, as one
sles_step_drive call per member, and every replay of the captured graph gives
the same states as the same steps run eagerly (replays interleaved with eager
work and synchronisation)."""

import math

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _same(a, b, tag):
    for da, db in zip(a.members, b.members):
        assert torch.equal(da.U_hat, db.U_hat), (tag, da.N)
    assert torch.equal(a.states(), b.states()), tag


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("les_model", ["Smagorinsky", "Vreman", None])
def test_batched_matches_per_member(precision, les_model):
    """Member offsets of every kernel of the step (x/y passes, z kernel,
    x-forward + update, both CFL reduction levels) and the per-member SGS
    coefficient: groups of 3, 2 and 2 members."""
    from sles.ensemble import Ensemble

    groups = ((16, 3), (24, 2), (96, 2))
    a = Ensemble(groups, precision=precision, seed=11, les_model=les_model, batched=True, batch=8)
    assert [len(g.members) for g in a.groups] == [3, 2, 2]
    b = Ensemble(groups, precision=precision, seed=11, les_model=les_model, batched=False)
    for k in range(3):
        a.step_eager(1)
        b.step_eager(1)
        torch.cuda.synchronize()
        print(f"[ensemble] {precision} {les_model} batched vs per-member, step {k + 1}", flush=True)
        _same(a, b, k)
    st = a.states()
    assert (st[:, 4] == 3).all()  # steps
    assert (st[:, 3] == 0).all()  # status RUNNING
    # members differ (own seeds, own Cs, own dt)
    assert len(set(st[:, 5].tolist())) == len(a.members)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_eager(precision):
    from sles.ensemble import Ensemble

    groups = ((16, 2), (24, 1), (32, 2))
    a = Ensemble(groups, precision=precision, seed=3)
    b = Ensemble(groups, precision=precision, seed=3)
    for da, db in zip(a.members, b.members):
        assert torch.equal(da.U_hat, db.U_hat)
        assert da.Cs == db.Cs
    b.capture(steps_per_graph=2)
    for rep in range(3):
        b.graph.replay()
        torch.cuda.synchronize()
        a.step_eager(2)
        torch.cuda.synchronize()
        _ = [torch.zeros(1024, device="cuda").sum().item() for _ in range(2)]  # unrelated eager work
        _same(a, b, rep)
    assert (b.states()[:, 4] == 6).all()


def test_members_have_distinct_seeds_and_cs():
    from sles.ensemble import Ensemble

    e = Ensemble(((16, 3),), precision="fp32", seed=5)
    cs = [d.Cs for d in e.members]
    assert len(set(cs)) == 3 and all(0.10 <= c <= 0.20 for c in cs)
    assert not torch.equal(e.members[0].U_hat, e.members[1].U_hat)
    g = e.groups[0]
    # members' domains are views of the group's packed state
    assert e.members[1].U_hat.data_ptr() == g.U_hat[1].data_ptr()
    assert torch.allclose(g.coeff.cpu(), torch.tensor([d.sgs_coeff() for d in e.members], dtype=torch.float64))
    st = e.states()
    assert st.shape == (3, 8) and all(math.isnan(t) for t in st[:, 2].tolist())

