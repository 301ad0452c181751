"""The chain from the sorted hits to the score pass's class summary as a
hipGraph (layout key "graph"): captured per slice, its executable graph
updated in place between batches (or built again when the chain's shape
changes), launched, and replayed many times. Every replay equals the direct
run; eager work on the same context between replays, batches of other sizes
in between, and a buffer that must grow during a capture (the chain then
runs directly) change nothing.
"""
import numpy as np
import pytest

from test_gpu_dropin import _score_batch
from test_gpu_search import EXTENDED, workload_dict
from conftest import mutate, rand_seq

pytestmark = pytest.mark.gpu


def _search(dmr, ctx, w, preset="sensitive"):
    r = ctx.blastp(w["q_letters"], w["q_offsets"], dmr.SearchConfig(sensitivity=preset))
    return r.format(w["q_ids"], w["db_ids"], EXTENDED)


def test_graph_replays_equal_direct_runs(dmr, synth):
    big = synth.make("blastp-sensitive-1024")
    sizes = {"all": list(range(0, 1024)), "half": list(range(0, 1024, 2)), "few": list(range(3, 1024, 97))}
    ws = {k: workload_dict(big.subset_queries(v)) for k, v in sizes.items()}
    direct = dmr.Context(0)
    direct.set_tuning("graph", 0)
    direct.db_load(ws["all"]["db_letters"], ws["all"]["db_offsets"])
    want = {k: _search(dmr, direct, w) for k, w in ws.items()}
    want["default"] = _search(dmr, direct, ws["half"], "default")
    direct.close()

    ctx = dmr.Context(0)
    ctx.set_tuning("graph", 1)
    ctx.db_load(ws["all"]["db_letters"], ws["all"]["db_offsets"])
    # The first search runs directly (it sizes the chain's buffers), the rest
    # replay the graph; smaller batches update it, a changed chain rebuilds it.
    order = ["all", "all", "all", "half", "all", "few", "half", "all"]
    rng = np.random.default_rng(3)
    for k, name in enumerate(order):
        assert _search(dmr, ctx, ws[name]) == want[name], (k, name)
        if k % 2:  # eager work on the context between replays
            q = rand_seq(rng, 200)
            cases = [(q, mutate(rng, q, 0.4, 0.02), -20, 20), (q, q.copy(), -12, 12)]
            got = _score_batch(ctx, cases)
            assert got[1] is not None and got[1] > 0
    assert _search(dmr, ctx, ws["half"], "default") == want["default"]
    launches, builds = ctx.get_tuning("graph_launches"), ctx.get_tuning("graph_builds")
    assert launches >= len(order) - 1 and 1 <= builds < launches, (launches, builds)
    ctx.close()


def test_graph_buffer_growth_falls_back(dmr, synth):
    """A batch with more hits than any earlier one runs directly (its buffers
    grow), and the next replay of it uses the graph."""
    big = synth.make("blastp-sensitive-1024")
    small = workload_dict(big.subset_queries(list(range(0, 1024, 8))))
    full = workload_dict(big)
    direct = dmr.Context(0)
    direct.set_tuning("graph", 0)
    direct.db_load(full["db_letters"], full["db_offsets"])
    want_small, want_full = _search(dmr, direct, small), _search(dmr, direct, full)
    direct.close()
    ctx = dmr.Context(0)
    ctx.set_tuning("graph", 1)
    ctx.db_load(full["db_letters"], full["db_offsets"])
    assert _search(dmr, ctx, small) == want_small
    assert _search(dmr, ctx, small) == want_small
    n0 = ctx.get_tuning("graph_launches")
    assert n0 >= 1
    assert _search(dmr, ctx, full) == want_full  # more hits: direct
    assert ctx.get_tuning("graph_launches") == n0
    assert _search(dmr, ctx, full) == want_full  # now replayed
    assert ctx.get_tuning("graph_launches") == n0 + 1
    ctx.close()
