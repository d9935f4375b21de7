"""Host logic of the ensemble: how members of one grid size are split into
batched launches."""


def test_batch_split():
    from sles.ensemble import batch_size

    # default: about 2^20 grid points per launch
    assert batch_size(64, 8) == 4 and batch_size(96, 8) == 1 and batch_size(128, 8) == 1
    assert batch_size(32, 8) == 8 and batch_size(16, 3) == 3 and batch_size(256, 2) == 1
    # explicit sizes, capped by the member count; dicts by grid size
    assert batch_size(64, 8, 3) == 3 and batch_size(64, 2, 8) == 2
    assert batch_size(64, 8, {64: 8}) == 8 and batch_size(96, 8, {64: 8}) == 1
    assert batch_size(64, 8, 0) == 1
