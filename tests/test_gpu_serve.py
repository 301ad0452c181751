"""The serving loop (dmr_server_*): requests of mixed presets and options
submitted at once and from two threads; every request's outfmt-6 text
equals dmr_blastp's for that query alone (the cap counts per request), and
a few equal the oracle's. A workload where the batch-wide cap
(src/search.rs:109-131) would drop seeds shows that the server keeps the
per-request semantics.
"""
import threading

import numpy as np
import pytest

from conftest import mutate, rand_seq
from test_gpu_search import EXTENDED

pytestmark = pytest.mark.gpu

PRESETS = ["default"] * 5 + ["sensitive"] * 4 + ["very-sensitive"]


def one(w, i):
    return w.q_letters[w.q_offsets[i]:w.q_offsets[i + 1]]


def test_requests_alone_equal_in_the_mix(dmr, synth, orc):
    w = synth.make("tiny")
    ctx = dmr.Context(0)
    ctx.db_load(w.db_letters, w.db_offsets)
    srv = dmr.Server(ctx, workers=2, hbm_budget=256 << 20)
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(72):
        qi = int(rng.integers(0, w.n_q))
        cfg = dmr.SearchConfig(sensitivity=PRESETS[k % len(PRESETS)])
        if k % 9 == 4:
            cfg.max_target_seqs = 3
        if k % 11 == 7:
            cfg.evalue = 10.0
        reqs.append((qi, cfg))
    tickets = [srv.submit(one(w, qi), cfg) for qi, cfg in reqs]
    texts, lat = [], []
    for t in tickets:
        r, ms = srv.wait(t)
        texts.append(r.format(["q"], w.db_ids, EXTENDED))
        lat.append(ms)
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["batches"] < len(reqs) and st["mean_batch"] > 1
    assert all(x > 0 for x in lat)
    for (qi, cfg), text in zip(reqs, texts):
        alone = ctx.blastp(one(w, qi), np.array([0, len(one(w, qi))], np.uint64), cfg)
        assert text == alone.format(["q"], w.db_ids, EXTENDED)
    # A few against the oracle, each query searched alone.
    for qi, cfg in reqs[:6]:
        q = one(w, qi)
        ref = orc.search(w.db_letters, w.db_offsets, q, np.array([0, len(q)], np.uint64),
                         sensitivity=cfg.sensitivity, threads=8, evalue=cfg.evalue,
                         max_target_seqs=cfg.max_target_seqs)
        assert texts[reqs.index((qi, cfg))] == ref.format(["q"], w.db_ids, EXTENDED)
    srv.close()
    ctx.close()


def test_cap_counts_each_request(dmr, orc):
    """Thirty requests of one query against a DB holding 300 near copies of
    it, in one index chunk: as one batch the cap drops the query's seeds
    (30 x ~220 entries per key > 4096), alone it keeps them; the server gives
    every request the alone result."""
    rng = np.random.default_rng(9)
    p = rand_seq(rng, 120)
    db = [mutate(rng, p, sub=0.03, indel=0.0) for _ in range(300)] + [rand_seq(rng, int(n)) for n in
                                                                      rng.integers(50, 400, size=100)]
    off = np.zeros(len(db) + 1, np.uint64)
    np.cumsum([len(s) for s in db], out=off[1:])
    letters = np.concatenate(db).astype(np.uint8)
    ids = [f"d{k}" for k in range(len(db))]
    ctx = dmr.Context(0)
    ctx.db_load(letters, off)
    cfg = dmr.SearchConfig(sensitivity="default", index_chunks=1)
    alone = ctx.blastp(p, np.array([0, len(p)], np.uint64), cfg).format(["q"], ids)
    batch = ctx.blastp(np.tile(p, 30), np.arange(31, dtype=np.uint64) * len(p), cfg)
    assert batch.stats()["hits"] < 30 * ctx.blastp(p, np.array([0, len(p)], np.uint64), cfg).stats()["hits"]
    ref = orc.search(letters, off, p, np.array([0, len(p)], np.uint64), sensitivity="default", threads=8,
                     index_chunks=1)
    assert alone == ref.format(["q"], ids) and alone.count("\n") > 20
    srv = dmr.Server(ctx, workers=2, hbm_budget=256 << 20, presets="default")
    out = {}

    def client(k0):
        ts = [srv.submit(p, cfg) for _ in range(15)]
        out[k0] = [srv.wait(t)[0].format(["q"], ids) for t in ts]

    th = [threading.Thread(target=client, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert srv.stats()["mean_batch"] > 1
    assert all(x == alone for k in out for x in out[k]) and len(out) == 2
    with pytest.raises(dmr.DmrError, match="unknown or already collected ticket"):
        srv.wait(10**9, 0)
    srv.close()
    ctx.close()


def test_shared_database_concurrent_searches(dmr, synth):
    """Contexts sharing one resident database (dmr_ctx_share_db) search their
    own query batches from concurrent host threads; every result equals the
    same batch searched on a context with its own copy of the database. The
    owner builds only the --sensitive index, so the sharers build the other
    two presets' indexes themselves, concurrently (once each, under the
    database's lock). After the owner loads another database the sharers
    keep searching theirs."""
    import threading
    w = synth.make("tiny")
    ref = dmr.Context(0)
    ref.db_load(w.db_letters, w.db_offsets)
    owner = dmr.Context(0)
    owner.db_load(w.db_letters, w.db_offsets)
    owner.build_index("sensitive")
    subs = [w.subset_queries(list(range(k, w.n_q, 3))) for k in range(3)]
    cfgs = [dmr.SearchConfig(sensitivity=s) for s in ("sensitive", "default", "very-sensitive")]
    alone = [[ref.blastp(s.q_letters, s.q_offsets, c).format(s.q_ids, w.db_ids, EXTENDED) for c in cfgs]
             for s in subs]
    ref.close()
    sharers = []
    for s in subs:
        c = dmr.Context(0)
        c.share_db(owner)
        c.queries_load(s.q_letters, s.q_offsets)
        sharers.append(c)
    got = [[None] * len(cfgs) for _ in subs]
    errors = []

    def run(k):
        order = list(range(len(cfgs)))
        if k % 2:
            order.reverse()
        try:
            for rep in range(3):
                for ci in order:
                    t = sharers[k].search_loaded(cfgs[ci]).format(subs[k].q_ids, w.db_ids, EXTENDED)
                    if got[k][ci] is None:
                        got[k][ci] = t
                    elif got[k][ci] != t:
                        errors.append((k, ci, rep))
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=run, args=(k,)) for k in range(len(subs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors
    assert got == alone
    assert all(any(t for t in row) for row in alone)
    owner.db_load(w.db_letters[:1000], np.array([0, 500, 1000], np.uint64))
    for k, c in enumerate(sharers):
        assert c.search_loaded(cfgs[0]).format(subs[k].q_ids, w.db_ids, EXTENDED) == alone[k][0]
        c.close()
    owner.close()
