"""The packed 16-bit score pass in both layouts and both arithmetics against
the oracle, and the merged-band classes (16 or 8 lanes per pair, per-DP walls).

Each width class runs either with one lane per DP pair (the whole band in a
lane's registers) or split over two or four lanes (part bands, with the
horizontal-gap correction across lanes). DMR_SW16_SPLIT selects the layout
per class when a context is created (bit c = class c). The pairs compute in
f16 (exact for integers up to 2048; DPs scoring 1900 or more are re-scored in
int32) or, with DMR_SW16_F16=0, in saturating int16 (re-scored from 32000).
Every combination is checked on drop-in batches (one DP per lane), on scores
around the f16 limit, and on a batched search (two DPs of the same query per
lane).
"""
import numpy as np
import pytest

from conftest import mutate, rand_seq
from test_gpu_dropin import _score_batch
from test_gpu_search import EXTENDED, run_both, workload_dict

pytestmark = pytest.mark.gpu

LAYOUTS = [(0x00, "1"), (0x1F, "1"), (0x00, "0"), (0x1F, "0")]


@pytest.fixture(params=LAYOUTS, ids=["full-f16", "split-f16", "full-i16", "split-i16"])
def layout_ctx(request, dmr, monkeypatch):
    mask, f16 = request.param
    monkeypatch.setenv("DMR_SW16_SPLIT", str(mask))
    monkeypatch.setenv("DMR_SW16_F16", f16)
    c = dmr.Context(0)
    got = c.score_layout()
    assert got == (mask, f16 == "1"), got
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


def test_scores_around_f16_limit(layout_ctx, orc):
    """Self and near-self alignments scoring 1500..2550 in every packed class:
    below 1900 the f16 pass stands, from 1900 on it is re-scored exactly; the
    largest stay far enough past 2048 in the row frames that f16 rounds."""
    rng = np.random.default_rng(5)
    hi = np.array([17, 4, 8, 18, 14, 13], np.uint8)  # W C H Y P F: self-scores 11, 9, 8, 7, 7, 6
    cases = []
    for n in range(200, 320, 4):
        q = hi[rng.integers(0, len(hi), size=n)]
        for band in (12, 16, 30, 40, 64):
            s = q.copy() if rng.random() < 0.5 else mutate(rng, q, 0.03, 0.01)
            c = int(rng.integers(-3, 4))
            cases.append((q, s, c - band, c + band))
    got = _score_batch(layout_ctx, cases)
    ref = [orc.extend_span(*c) for c in cases]
    ref = [r["score"] if r else None for r in ref]
    scores = [r for r in ref if r is not None]
    assert min(scores) < 1850 and max(scores) > 2200
    assert sum(1850 <= r <= 1950 for r in scores) >= 10
    for c, g, r in zip(cases, got, ref):
        assert g == r, (len(c[0]), len(c[1]), c[2], c[3], r)


def test_merged_band_widths(layout_ctx, orc):
    """Bands of widths 62..256 other than 81 and 129 (merged bands) run the
    split kernels with a wall at each DP's own width (classes m193/m256);
    61, 81, 129 and 257+ stay on their own kernels. Every width near a class
    edge is included."""
    rng = np.random.default_rng(91)
    edges = [60, 61, 62, 63, 64, 65, 80, 81, 82, 100, 127, 128, 129, 130, 131, 192, 193, 194, 195, 255,
             256, 257, 258]
    cases = []
    for k in range(420):
        w = edges[k] if k < len(edges) else int(rng.integers(62, 257))
        n = int(rng.integers(1, 1000 if rng.random() < 0.2 else 400))
        q = rand_seq(rng, n, canonical_only=rng.random() < 0.9)
        r = rng.random()
        s = q.copy() if r < 0.2 else (mutate(rng, q, float(rng.uniform(0.1, 0.6)), 0.08) if r < 0.7
                                      else rand_seq(rng, int(rng.integers(1, 900))))
        if len(s) == 0:
            s = rand_seq(rng, 2)
        d0 = int(rng.integers(-len(s) - w, n + 1))
        cases.append((q, s, d0, d0 + w - 1))
    assert sum(61 < c[3] - c[2] + 1 <= 128 for c in cases) > 50
    assert sum(129 < c[3] - c[2] + 1 <= 193 for c in cases) > 50
    assert sum(193 < c[3] - c[2] + 1 <= 256 for c in cases) > 50
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
    prof = layout_ctx.class_profile()
    assert prof["m193"]["jobs"] > 0 and prof["m256"]["jobs"] > 0
    assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)
