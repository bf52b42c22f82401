"""Launch selection of the autotuner (host logic only)."""
from forcepho_amd import autotune


def L(K, W, P, G):
    return dict(pix_per_lane=K, waves_per_group=W, persistent=P, groups_per_cu=G, cull=1)


def test_pick_fastest_when_clearly_faster():
    table = [(70.0, L(8, 4, 1, 8)), (76.0, L(8, 1, 0, 8)), (90.0, L(4, 1, 0, 8))]
    assert autotune.pick(table) == L(8, 4, 1, 8)


def test_pick_prefers_grid_launch_within_tie():
    # persistent 0.6% faster: a tie, so the grid launch wins
    table = [(76.00, L(8, 4, 1, 8)), (76.43, L(8, 1, 0, 8)), (76.31, L(8, 2, 0, 8))]
    assert autotune.pick(table) == L(8, 1, 0, 8)


def test_pick_fewer_waves_then_fewer_groups():
    table = [(50.0, L(4, 2, 1, 16)), (50.2, L(4, 2, 1, 8)), (50.1, L(4, 4, 1, 4))]
    assert autotune.pick(table) == L(4, 2, 1, 8)
    assert autotune.pick(table, tie=0.0) == L(4, 2, 1, 16)


def test_candidates_cover_grid_and_persistent():
    c = autotune.candidates("batch")
    assert any(x["persistent"] == 0 for x in c) and any(x["persistent"] == 1 for x in c)
    assert all(x["pix_per_lane"] in (1, 2, 4, 8) for x in c)
    assert all(x["persistent"] == 0 for x in autotune.candidates("patch"))
