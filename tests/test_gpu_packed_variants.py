"""The packed int16 score pass in both layouts against the oracle.

Each width class runs either with one lane per DP pair (the whole band in a
lane's registers) or split over four lanes (quarter bands, with the
horizontal-gap correction across lanes). DMR_SW16_SPLIT selects the layout
per class when a context is created (bit c = class c); both extremes are
checked on drop-in batches (one DP per lane) and on a batched search (two DPs
of the same query per lane).
"""
import numpy as np
import pytest

from conftest import mutate, rand_seq
from test_gpu_dropin import _score_batch
from test_gpu_search import EXTENDED, run_both, workload_dict

pytestmark = pytest.mark.gpu

MASKS = [0x00, 0x1F]


@pytest.fixture(params=MASKS, ids=["full", "split"])
def layout_ctx(request, dmr, monkeypatch):
    monkeypatch.setenv("DMR_SW16_SPLIT", str(request.param))
    c = dmr.Context(0)
    yield c
    c.close()


def test_band_widths_both_layouts(layout_ctx, orc):
    rng = np.random.default_rng(77)
    cases = []
    for _ in range(400):
        band = int(rng.choice([12, 16, 30, 40, 64]))
        n = int(rng.integers(1, 900 if rng.random() < 0.2 else 300))
        q = rand_seq(rng, n, canonical_only=rng.random() < 0.9)
        r = rng.random()
        s = q.copy() if r < 0.2 else (mutate(rng, q, float(rng.uniform(0.1, 0.6)), 0.06) if r < 0.7
                                      else rand_seq(rng, int(rng.integers(1, 700))))
        if len(s) == 0:
            s = rand_seq(rng, 2)
        c = int(rng.integers(-len(s) - band, n + band))
        cases.append((q, s, c - band, c + band))
    got = _score_batch(layout_ctx, cases)
    for (q, s, d0, d1), g in zip(cases, got):
        ref = orc.extend_span(q, s, d0, d1)
        assert g == (ref["score"] if ref else None), (len(q), len(s), d0, d1)


def test_batched_search_both_layouts(dmr, layout_ctx, orc, synth):
    """24 queries of the 1024-query --sensitive workload: many DPs per query,
    so lanes carry two DPs."""
    w = synth.make("blastp-sensitive-1024").subset_queries(list(range(5, 1024, 43)))
    w = workload_dict(w)
    res, ref = run_both(dmr, layout_ctx, orc, w, "sensitive")
    assert len(ref) > 0
    assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)
