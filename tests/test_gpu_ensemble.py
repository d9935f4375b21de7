"""Ensemble of uneven grids under a captured HIP graph: every replay gives
the same states as the same steps run eagerly (replays interleaved with
eager work and synchronisation)."""

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_eager(precision):
    from sles.ensemble import Ensemble

    groups = ((16, 2), (24, 1), (32, 2))
    a = Ensemble(groups, precision=precision, seed=3)
    b = Ensemble(groups, precision=precision, seed=3)
    for da, db in zip(a.members, b.members):
        assert torch.equal(da.U_hat, db.U_hat)
        assert da.Cs == db.Cs
    b.capture(steps_per_graph=1)  # runs the first (host-dt) step and one eager warm step
    a.step_eager(1)  # the same two steps on the eager twin
    torch.cuda.synchronize()
    for rep in range(3):
        b.graph.replay()
        torch.cuda.synchronize()
        a.step_eager(1)
        torch.cuda.synchronize()
        _ = [torch.zeros(1024, device="cuda").sum().item() for _ in range(2)]  # unrelated eager work
        for da, db in zip(a.members, b.members):
            assert torch.equal(da.U_hat, db.U_hat), (rep, da.N)
            assert torch.equal(da._scalars[:2], db._scalars[:2])


def test_members_have_distinct_seeds_and_cs():
    from sles.ensemble import Ensemble

    e = Ensemble(((16, 3),), precision="fp32", seed=5)
    cs = [d.Cs for d in e.members]
    assert len(set(cs)) == 3 and all(0.10 <= c <= 0.20 for c in cs)
    assert not torch.equal(e.members[0].U_hat, e.members[1].U_hat)
