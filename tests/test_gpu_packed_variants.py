"""The packed 16-bit score pass in both layouts and both arithmetics against
the oracle, the merged-band classes (per-DP walls), and every compiled
packed kernel.

A packed DP runs with the query or the subject on the rows ("transpose": 0
query, 1 the one with fewer rows, 2 subject; lanes pair two DPs with the same
row sequence), which changes no score.

Each band_for width class runs either with one lane per DP pair (the whole
band in a lane's registers) or split over two or more lanes (part bands,
with the horizontal-gap correction across lanes); dmr_ctx_set_tuning selects
the kernel per class ("lanes_<class>", "threads_<class>") and the merged-band
kernels in use ("merged_mask"). The pairs compute in f16 (exact for integers
up to 2048; DPs scoring 1900 or more are re-scored in int32) or, with "f16"
0, in saturating int16 (re-scored from 32000). Every combination is checked
on drop-in batches (one DP per lane), on scores around the f16 limit, and on
a batched search (two DPs of the same query per lane); every compiled kernel
on the widths of its class.
"""
import numpy as np
import pytest

from conftest import mutate, rand_seq
from test_gpu_dropin import _score_batch
from test_gpu_search import EXTENDED, run_both, workload_dict

pytestmark = pytest.mark.gpu

EXACT = ["w25", "w33", "w61", "w81", "w129"]
LAYOUTS = [(False, True), (True, True), (False, False), (True, False)]


def first_split(dmr, name):
    return next(v for v in dmr.sw16_variants() if v["cls"] == name and v["lanes"] > 1)


@pytest.fixture(params=LAYOUTS, ids=["full-f16", "split-f16", "full-i16", "split-i16"])
def layout_ctx(request, dmr):
    split, f16 = request.param
    c = dmr.Context(0)
    c.load_tuning(None)
    for name in EXACT:
        c.set_tuning(f"lanes_{name}", first_split(dmr, name)["lanes"] if split else 1)
    c.set_tuning("f16", int(f16))
    mask = sum(1 << k for k, n in enumerate(EXACT) if split)
    got = c.score_layout()
    assert got == (mask, f16), got
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


def merged_cases(seed=91, count=640):
    """Bands of widths 62..512 (merged bands), with every width near a
    merged-class edge and the band_for widths between them."""
    rng = np.random.default_rng(seed)
    edges = [60, 61, 62, 63, 64, 65, 80, 81, 82, 96, 97, 98, 100, 127, 128, 129, 130, 131, 160, 161, 162,
             192, 193, 194, 195, 224, 225, 226, 255, 256, 257, 258, 320, 321, 383, 384, 385, 386, 448, 449,
             511, 512, 513, 514]
    cases = []
    for k in range(count):
        w = edges[k] if k < len(edges) else int(rng.integers(62, 513))
        n = int(rng.integers(1, 1000 if rng.random() < 0.2 else 400))
        q = rand_seq(rng, n, canonical_only=rng.random() < 0.9)
        r = rng.random()
        s = q.copy() if r < 0.2 else (mutate(rng, q, float(rng.uniform(0.1, 0.6)), 0.08) if r < 0.7
                                      else rand_seq(rng, int(rng.integers(1, 900))))
        if len(s) == 0:
            s = rand_seq(rng, 2)
        d0 = int(rng.integers(-len(s) - w, n + 1))
        cases.append((q, s, d0, d0 + w - 1))
    return cases


@pytest.mark.parametrize("f16,mask", [(1, 0xA8), (0, 0xA8), (1, 0xFF), (0, 0xFF)],
                         ids=["f16-m193-m257", "i16-m193-m257", "f16-all", "i16-all"])
def test_fused_launch_band_widths(dmr, orc, f16, mask):
    """The fused launch (merged bands with walls, widths 81 and 61 split over
    two lanes) on drop-in DPs of every width it takes and of the others, with
    three and with all eight merged-band classes in use."""
    cases = merged_cases(seed=93, count=300)
    rng = np.random.default_rng(94)
    for _ in range(200):
        band = int(rng.choice([30, 40]))
        n = int(rng.integers(1, 700))
        q = rand_seq(rng, n)
        s = q.copy() if rng.random() < 0.3 else mutate(rng, q, float(rng.uniform(0.1, 0.6)), 0.06)
        if len(s) == 0:
            s = rand_seq(rng, 2)
        c0 = int(rng.integers(-len(s) - band, n + band))
        cases.append((q, s, c0 - band, c0 + band))
    ref = [orc.extend_span(*c) for c in cases]
    ref = [r["score"] if r else None for r in ref]
    c = dmr.Context(0)
    c.load_tuning(None)
    c.set_tuning("f16", f16)
    c.set_tuning("fuse", 1)
    c.set_tuning("merged_mask", mask)
    got = _score_batch(c, cases)
    prof = c.class_profile()
    assert prof["w81"]["ms"] > 0 and prof["m257"]["ms"] > 0
    used = (["m97", "m129", "m161", "m193", "m225", "m257", "m385", "m513"] if mask == 0xFF
            else ["m193", "m257", "m513"])
    assert all(prof[k]["ms"] > 0 for k in used)
    for cs, g, r in zip(cases, got, ref):
        assert g == r, (len(cs[0]), len(cs[1]), cs[2], cs[3])
    c.close()


def test_merged_band_widths(layout_ctx, orc):
    """Bands of widths 62..512 other than 81 and 129 (merged bands) run the
    split kernels with a wall at each DP's own width (m193 and m257, with
    wider bands on the int32 wave kernels; m193, m257 and m513; then all
    eight merged kernels); 61, 81, 129 and 513+ stay on their own kernels."""
    cases = merged_cases()
    assert sum(61 < c[3] - c[2] + 1 <= 128 for c in cases) > 50
    assert sum(129 < c[3] - c[2] + 1 <= 193 for c in cases) > 50
    assert sum(193 < c[3] - c[2] + 1 <= 256 for c in cases) > 50
    assert sum(256 < c[3] - c[2] + 1 <= 384 for c in cases) > 50
    assert sum(384 < c[3] - c[2] + 1 <= 512 for c in cases) > 50
    ref = [orc.extend_span(*c) for c in cases]
    ref = [r["score"] if r else None for r in ref]
    for mask in (0x28, 0xA8, 0xFF):
        layout_ctx.set_tuning("merged_mask", mask)
        got = _score_batch(layout_ctx, cases)
        for c, g, r in zip(cases, got, ref):
            assert g == r, (mask, len(c[0]), len(c[1]), c[2], c[3])


def test_batched_search_both_layouts(dmr, layout_ctx, orc, synth):
    """24 queries of the 1024-query --sensitive workload: many DPs per query,
    so lanes carry two DPs."""
    w = synth.make("blastp-sensitive-1024").subset_queries(list(range(5, 1024, 43)))
    w = workload_dict(w)
    res, ref = run_both(dmr, layout_ctx, orc, w, "sensitive")
    assert len(ref) > 0
    prof = layout_ctx.class_profile()
    assert prof["m193"]["jobs"] > 0 and prof["m257"]["jobs"] > 0
    assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)


@pytest.mark.parametrize("transpose,fuse", [(0, 0), (1, 0), (2, 0), (1, 1), (2, 1)])
def test_batched_search_orientations(dmr, orc, synth, transpose, fuse):
    """The 24-query subset with every packed DP's query on the rows (lanes
    pair DPs of one query) and with every packed DP's subject on the rows
    (lanes pair DPs of one subject, from different queries); with fuse, the
    merged-band, width-81 and width-61 classes in one launch."""
    w = synth.make("blastp-sensitive-1024").subset_queries(list(range(5, 1024, 43)))
    w = workload_dict(w)
    c = dmr.Context(0)
    c.load_tuning(None)
    c.set_tuning("transpose", transpose)
    c.set_tuning("fuse", fuse)
    assert c.get_tuning("transpose") == transpose and c.get_tuning("fuse") == fuse
    res, ref = run_both(dmr, c, orc, w, "sensitive")
    assert len(ref) > 0
    assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)
    c.close()


@pytest.fixture(scope="module")
def variant_cases(orc):
    """Per packed class, DPs of its widths with the oracle's scores."""
    rng = np.random.default_rng(404)
    merged = merged_cases(seed=405, count=900)
    out = {}
    bounds = {"m97": (62, 97), "m129": (98, 129), "m161": (130, 161), "m193": (162, 193), "m225": (194, 225),
              "m257": (226, 256), "m385": (257, 384), "m513": (385, 512)}
    for name, (lo, hi) in bounds.items():
        out[name] = [c for c in merged if lo <= c[3] - c[2] + 1 <= hi and c[3] - c[2] + 1 not in (81, 129)]
    for name in EXACT:
        w = int(name[1:])
        cases = []
        for _ in range(150):
            n = int(rng.integers(1, 900 if rng.random() < 0.2 else 300))
            q = rand_seq(rng, n, canonical_only=rng.random() < 0.9)
            s = q.copy() if rng.random() < 0.3 else mutate(rng, q, float(rng.uniform(0.1, 0.6)), 0.06)
            if len(s) == 0:
                s = rand_seq(rng, 2)
            d0 = int(rng.integers(-len(s) - w, n + 1))
            cases.append((q, s, d0, d0 + w - 1))
        out[name] = cases
    for name, cases in out.items():
        assert len(cases) >= 40, name
        out[name] = (cases, [(r["score"] if r else None) for r in (orc.extend_span(*c) for c in cases)])
    return out


@pytest.mark.parametrize("f16,transpose", [(1, 1), (0, 1), (1, 0), (1, 2)],
                         ids=["f16", "i16", "f16-query-rows", "f16-subject-rows"])
def test_every_compiled_kernel(dmr, variant_cases, f16, transpose):
    """Each compiled packed kernel (class, lanes per DP pair, threads), pinned
    for its class with every merged kernel in use, scores the widths of its
    class as the oracle does, with either sequence on the rows."""
    c = dmr.Context(0)
    c.load_tuning(None)
    c.set_tuning("f16", f16)
    c.set_tuning("transpose", transpose)
    c.set_tuning("merged_mask", 0xFF)
    seen = 0
    for v in dmr.sw16_variants():
        c.set_tuning(f"lanes_{v['cls']}", v["lanes"])
        c.set_tuning(f"threads_{v['cls']}", v["threads"])
        cases, ref = variant_cases[v["cls"]]
        got = _score_batch(c, cases)
        assert c.get_tuning(f"lanes_{v['cls']}") == v["lanes"]
        assert c.class_profile()[v["cls"]]["ms"] > 0, v  # the class's kernel ran
        for cs, g, r in zip(cases, got, ref):
            assert g == r, (v, len(cs[0]), len(cs[1]), cs[2], cs[3])
        seen += 1
    assert seen == len(dmr.sw16_variants()) >= 28
    c.close()
