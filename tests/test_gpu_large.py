"""The 60 M letter workloads (200,000 proteins) on the GPU against the
oracle's search of the same batches, stored as fixtures by
tests/golden/make_search_golden.py (the oracle takes minutes at this size).

--very-sensitive (16 shapes of weight 7, diamond9, one index chunk): 64 of
the 4,096 queries, searched as one batch on both sides (the max_expand cap
counts the batch, src/search.rs:109-131).
"""
import gzip
import pathlib

import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden" / "search"
FIELDS = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue bitscore "
          "score cigar")


def test_very_sensitive_subset_matches_oracle(dmr, ctx):
    import make_search_golden as mg
    w, sens = mg.workload("blastp-very-sensitive-4096-q64")
    exp = gzip.open(GOLD / "blastp-very-sensitive-4096-q64.tsv.gz", "rb").read().decode()
    ctx.db_load(w.db_letters, w.db_offsets)
    got = ctx.blastp(w.q_letters, w.q_offsets, dmr.SearchConfig(sensitivity=sens)).format(w.q_ids, w.db_ids,
                                                                                          FIELDS)
    assert exp.count("\n") > 50
    assert got == exp


def test_blastx_subset_matches_oracle(dmr, ctx):
    """blastx: 64 of the 2,048 reads (150-3,000 bases, reverse-translated DB
    proteins with sequencing errors, both strands) against the 60 M letter
    DB; frames translated and masked on the device; nucleotide coordinates,
    frames and strands as the oracle prints them."""
    import make_search_golden as mg
    w, idx, sens = mg.read_workload("blastx-2048-r64")
    exp = gzip.open(GOLD / "blastx-2048-r64.tsv.gz", "rb").read().decode()
    ctx.db_load(w.db_letters, w.db_offsets)
    res = ctx.blastx([w.reads[i] for i in idx], dmr.SearchConfig(sensitivity=sens))
    got = res.format([w.read_ids[i] for i in idx], w.db_ids, mg.READ_FIELDS)
    assert exp.count("\n") > 30 and "\t-" in exp and "\t+" in exp
    assert got == exp
