"""GPU blastp against the oracle: byte-identical outfmt-6 text (default
fields and the extended ones: raw score, CIGAR, BTOP, aligned strings).

The max_expand cap counts query seeds over the whole batch
(src/search.rs:109-131), so both sides search the same query set in one call.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXTENDED = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue "
            "bitscore score nident positive gaps ppos qlen slen qcovhsp scovhsp cigar btop "
            "qseq sseq qframe qstrand")


def run_both(dmr, ctx, orc, w, sens, **kw):
    ctx.db_load(w["db_letters"], w["db_offsets"])
    cfg = dmr.SearchConfig(sensitivity=sens, **kw)
    res = ctx.blastp(w["q_letters"], w["q_offsets"], cfg)
    ref = orc.search(w["db_letters"], w["db_offsets"], w["q_letters"], w["q_offsets"],
                     sensitivity=sens, threads=8, **{k: v for k, v in kw.items()
                                                     if k in ("evalue", "max_target_seqs", "max_hsps")})
    return res, ref


def workload_dict(w):
    return dict(db_letters=w.db_letters, db_offsets=w.db_offsets, q_letters=w.q_letters,
                q_offsets=w.q_offsets, db_ids=w.db_ids, q_ids=w.q_ids)


@pytest.mark.parametrize("sens", ["default", "sensitive", "very-sensitive", "fast"])
def test_tiny_workload(dmr, ctx, orc, synth, sens):
    w = workload_dict(synth.make("tiny"))
    res, ref = run_both(dmr, ctx, orc, w, sens)
    got = res.format(w["q_ids"], w["db_ids"], EXTENDED)
    exp = ref.format(w["q_ids"], w["db_ids"], EXTENDED)
    assert len(ref) > 0
    assert got == exp


@pytest.mark.parametrize("kw", [dict(max_hsps=3), dict(max_target_seqs=2), dict(evalue=10.0)])
def test_tiny_options(dmr, ctx, orc, synth, kw):
    w = workload_dict(synth.make("tiny"))
    res, ref = run_both(dmr, ctx, orc, w, "sensitive", **kw)
    assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)


@pytest.mark.parametrize("sens", ["default", "fast", "sensitive"])
def test_reference_e2e_data(dmr, ctx, orc, e2e, sens):
    res, ref = run_both(dmr, ctx, orc, e2e, sens)
    got = res.format(e2e["q_ids"], e2e["db_ids"], EXTENDED)
    assert got == ref.format(e2e["q_ids"], e2e["db_ids"], EXTENDED)
    assert len(res) == {"default": 4, "fast": 1, "sensitive": 7}[sens]
    # qtitle / stitle: the full FASTA headers (src/output.rs:201-202).
    f = "qseqid qtitle sseqid stitle qlen slen"
    titled = res.format(e2e["q_ids"], e2e["db_ids"], f, e2e["q_headers"], e2e["db_headers"])
    assert titled == ref.format(e2e["q_ids"], e2e["db_ids"], f, e2e["q_headers"], e2e["db_headers"])
    assert all(len(l.split("\t")[3]) > len(l.split("\t")[2]) for l in titled.splitlines())


def test_sensitive_batch_subset(dmr, ctx, orc, synth):
    """48 queries of the 1024-query --sensitive workload against the 2.26 Maa DB."""
    w = synth.make("blastp-sensitive-1024").subset_queries(list(range(0, 1024, 21)))
    w = workload_dict(w)
    res, ref = run_both(dmr, ctx, orc, w, "sensitive")
    assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)
    st = res.stats()
    c = ref.counters()
    assert st["hits"] == c["hits"]
    assert st["jobs"] == c["bands"]


@pytest.mark.parametrize("sens", ["default", "sensitive"])
def test_blastx_reads(dmr, ctx, orc, synth, sens):
    """blastx: six frames per read, --min-orf masking, nucleotide coordinates
    and frames in the output (src/workflow.rs:112-127, src/translate.rs:54-137),
    against the oracle on reads from both strands."""
    w = synth.make_reads("blastx-tiny")
    ctx.db_load(w.db_letters, w.db_offsets)
    res = ctx.blastx(w.reads, dmr.SearchConfig(sensitivity=sens))
    nt = np.frombuffer("".join(w.reads).encode(), dtype=np.uint8)
    off = np.zeros(len(w.reads) + 1, dtype=np.uint64)
    np.cumsum([len(r) for r in w.reads], out=off[1:])
    ref = orc.search(w.db_letters, w.db_offsets, nt, off, sensitivity=sens, threads=8, translated=True)
    assert len(ref) > 0
    exp = ref.format(w.read_ids, w.db_ids, EXTENDED)
    assert "\t-" in exp and "\t+" in exp  # hits on both strands
    assert res.format(w.read_ids, w.db_ids, EXTENDED) == exp


def test_results_keep_their_ops(dmr, synth):
    """A result's alignment ops live in a pinned block that the result holds:
    later searches on the same context, results being destroyed (their blocks
    reused) and the context itself going away leave every live result's text
    and ops unchanged."""
    w = workload_dict(synth.make("tiny"))
    c = dmr.Context(0)
    c.db_load(w["db_letters"], w["db_offsets"])
    run = lambda sens: c.blastp(w["q_letters"], w["q_offsets"], dmr.SearchConfig(sensitivity=sens))
    fmt = lambda r: r.format(w["q_ids"], w["db_ids"], EXTENDED)
    a, b = run("default"), run("sensitive")
    assert len(a) > 0 and len(b) > 0
    text_b = fmt(b)
    ops_b = [list(b.ops(i)) for i in range(len(b))]
    del a  # its block goes back to the pool; the next search takes it
    c2 = run("default")
    text_c = fmt(c2)
    assert fmt(b) == text_b
    c.close()
    assert fmt(b) == text_b and fmt(c2) == text_c
    assert [list(b.ops(i)) for i in range(len(b))] == ops_b
    assert all(len(o) > 0 for o in ops_b)


def test_full_sensitive_batch_matches_oracle(dmr, ctx, synth):
    """The whole 1024-query --sensitive workload the benchmark times, as one
    batch (the cap counts every query of it, src/search.rs:109-131), equals
    the oracle's search of the same batch, stored by
    tests/golden/make_search_golden.py (12 default fields, score, CIGAR)."""
    import gzip
    import pathlib
    g = pathlib.Path(__file__).resolve().parent / "golden" / "search" / "blastp-sensitive-1024.tsv.gz"
    exp = gzip.open(g, "rb").read().decode()
    w = synth.make("blastp-sensitive-1024")
    ctx.db_load(w.db_letters, w.db_offsets)
    ctx.queries_load(w.q_letters, w.q_offsets)
    res = ctx.search_loaded(dmr.SearchConfig(sensitivity="sensitive"))
    fields = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue bitscore "
              "score cigar")
    got = res.format(w.q_ids, w.db_ids, fields)
    assert got.count("\n") == exp.count("\n") == 4735
    assert got == exp


@pytest.mark.parametrize("slice_hits", [1, 300])
def test_sliced_search_matches_oracle(dmr, orc, synth, slice_hits):
    """A search run in slices of targets (each slice's hits under slice_hits;
    1 = one target per slice) prints what the oracle prints for the whole
    batch: the cap decisions are the batch's, slices only bound memory."""
    w = workload_dict(synth.make("tiny"))
    c = dmr.Context(0)
    c.set_tuning("slice_hits", slice_hits)
    c.db_load(w["db_letters"], w["db_offsets"])
    for sens in ["sensitive", "default"]:
        res = c.blastp(w["q_letters"], w["q_offsets"], dmr.SearchConfig(sensitivity=sens))
        assert c.get_tuning("last_slices") >= 2
        ref = orc.search(w["db_letters"], w["db_offsets"], w["q_letters"], w["q_offsets"], sensitivity=sens,
                         threads=8)
        assert res.format(w["q_ids"], w["db_ids"], EXTENDED) == ref.format(w["q_ids"], w["db_ids"], EXTENDED)
        assert res.stats()["hits"] == ref.counters()["hits"]
    c.close()


def test_sliced_full_batch_and_profile(dmr, synth):
    """The 1024-query batch in about ten slices equals the stored oracle text;
    with profiling on, the DP cells and class jobs summed over the slices
    equal those of the same search in one slice."""
    import gzip
    import pathlib
    g = pathlib.Path(__file__).resolve().parent / "golden" / "search" / "blastp-sensitive-1024.tsv.gz"
    exp = gzip.open(g, "rb").read().decode()
    fields = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue bitscore "
              "score cigar")
    w = synth.make("blastp-sensitive-1024")
    c = dmr.Context(0)
    c.db_load(w.db_letters, w.db_offsets)
    c.queries_load(w.q_letters, w.q_offsets)
    cfg = dmr.SearchConfig(sensitivity="sensitive")
    c.search_loaded(cfg)
    assert c.get_tuning("last_slices") == 1
    cells1 = c.last_dp_cells()
    prof1 = c.class_profile()
    c.set_tuning("slice_hits", 300000)
    c.set_tuning("profile", 1)
    res = c.search_loaded(cfg)
    assert c.get_tuning("last_slices") >= 8
    assert res.format(w.q_ids, w.db_ids, fields) == exp
    assert c.last_dp_cells() == cells1
    prof = c.class_profile()
    assert {k: (v["jobs"], v["cells"]) for k, v in prof.items()} == {k: (v["jobs"], v["cells"])
                                                                     for k, v in prof1.items()}
    c.close()
