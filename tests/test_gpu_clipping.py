"""DP window clipping (src/search.rs:287-354): a band whose span times the
query length passes the 16,000,000-cell budget is computed on a query
window of max(16e6 / row_cells, 2 * band + 16) rows around the band's
anchor, with the band's diagonals shifted by the window start.

The workload makes such a band: a 4,000+ letter query and a subject built
from mutated copies of one repeat unit, so the segments of every copy pair
chain into one merged band thousands of diagonals wide. One copy pair is
identical and sits in the middle of both, so the anchor (the band's best
segment) is there and the window starts past the query's first letter.
"""
import numpy as np
import pytest

from test_gpu_search import EXTENDED

pytestmark = pytest.mark.gpu


def clipping_workload(synth, seed=77, copies=40, unit_len=90, linker=30, core_len=250,
                      identity=(0.6, 0.8)):
    """Query and subject: `copies` mutated copies of one repeat unit, each
    followed by a random linker (period < 2 * band, so the bands of the copy
    pairs' diagonals overlap and merge); in the middle of both, the same
    core_len letters, whose segment is the band's unique best."""
    rng = np.random.default_rng(seed)
    res = lambda n: synth._residues(rng, n)
    unit = res(unit_len)
    core = res(core_len)

    def copy(k, mid):
        body = core if k == mid else synth.mutate(rng, unit, rng.uniform(*identity), 0.0)
        return np.concatenate([body, res(linker)])

    mid = copies // 2
    query = np.concatenate([copy(k, mid) for k in range(copies)])
    subject = np.concatenate([copy(k, mid) for k in range(copies)])
    background = [res(int(n)) for n in rng.integers(80, 600, size=40)]
    db = background[:20] + [subject] + background[20:]
    db_off = np.zeros(len(db) + 1, np.uint64)
    np.cumsum([len(s) for s in db], out=db_off[1:])
    q_off = np.array([0, len(query)], np.uint64)
    return dict(db_letters=np.concatenate(db).astype(np.uint8), db_offsets=db_off,
                q_letters=query.astype(np.uint8), q_offsets=q_off,
                db_ids=[f"c{k:03d}" for k in range(len(db))], q_ids=["repeat_query"])


@pytest.mark.parametrize("sens", ["default", "sensitive"])
def test_clipped_window_matches_oracle(dmr, ctx, orc, synth, sens):
    w = clipping_workload(synth)
    qlen = int(w["q_offsets"][-1])
    ref = orc.search(w["db_letters"], w["db_offsets"], w["q_letters"], w["q_offsets"], sensitivity=sens,
                     threads=8)
    c = ref.counters()
    # The case is what it claims: a clipped DP, and no tied band maxima whose
    # order the reference's unstable sorts would leave open.
    assert c["clipped_dps"] >= 1
    assert c["anchor_tie_events"] == 0 and c["band_tie_events"] == 0
    h = ref.hsps()
    assert len(ref) >= 1
    top = int(np.argmax(h["score"]))
    # The window starts inside the query and ends before its last letter.
    assert 0 < h["q_begin"][top] and h["q_end"][top] < qlen
    ctx.db_load(w["db_letters"], w["db_offsets"])
    got = ctx.blastp(w["q_letters"], w["q_offsets"], dmr.SearchConfig(sensitivity=sens))
    assert got.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)


@pytest.mark.parametrize("copies", [16, 26, 44, 60])
def test_wide_merged_bands_match_oracle(dmr, ctx, orc, synth, copies):
    """Merged bands 2,000-15,000 diagonals wide (the block-per-DP kernel in
    the score pass and, with the score pass's best, in the traceback),
    clipped or not, against the oracle."""
    w = clipping_workload(synth, copies=copies, seed=31 + copies)
    ref = orc.search(w["db_letters"], w["db_offsets"], w["q_letters"], w["q_offsets"], sensitivity="default",
                     threads=8)
    c = ref.counters()
    assert c["anchor_tie_events"] == 0 and c["band_tie_events"] == 0 and len(ref) >= 1
    ctx.db_load(w["db_letters"], w["db_offsets"])
    got = ctx.blastp(w["q_letters"], w["q_offsets"], dmr.SearchConfig(sensitivity="default"))
    prof = ctx.class_profile()
    assert prof["rowexact"]["jobs"] >= 1  # the class of bands wider than 2048
    assert got.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)
