#!/usr/bin/env python3
"""Benchmark of the MI355X search core.

Workload (one step = one batched search): blastp --sensitive, 1024 synthetic
queries (log-normal length, median 300, clipped 50..1000, 0-10 planted
homologs each) against a synthetic DB of 7,860 proteins (~2.26 M letters),
BLOSUM62 11/1, evalue 1e-3, k 25, max-hsps 1. The DB, its resident seed
index and the query batch are in HBM before timing; a step runs the whole
search (seeds, join, X-drop, bands, banded DP, selection, traceback) and
renders the outfmt-6 text.

The timed steps run twice: one batch at a time (per-class kernel times,
GCUPS and the roofline come from these), then with --streams batches in
flight (default 2: host threads with contexts sharing the resident database
and index), which gives the headline throughput.

Prints one JSON line on stdout; progress goes to stderr.
  value        queries/s (whole job, one GPU, --streams batches in flight)
  gcups        banded-DP cells / score-kernel time
  roofline     the score kernel against the packed 16-bit VALU peak
  cpu_baseline the oracle (CPU restatement of the reference path, all host
               threads up to 16) on a bounded query sample
  plain_form   the per-call drop-in form on a small sample
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
METRIC = "queries/s + gapped GCUPS at --sensitive, bit-exact outfmt6; p50/p99 latency"
PACKED16_VALU_PEAK_TOPS = 78.6   # 256 CU x 64 lanes x 2 (packed int16/f16) x 2.4 GHz
OPS_PER_CELL = 11             # E 3, F 3, H 1 add + 3 max, best 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load(name, path, pkg=False):
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[str(path.parent)] if pkg else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def hbm_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` measured by rocprofv3 PMC passes on this
    workload (profiles/hbm_traffic.json), or None when not measured."""
    try:
        d = json.loads((ROOT / "profiles" / "hbm_traffic.json").read_text())
    except (OSError, ValueError):
        return None
    if d.get("workload") != workload:
        return None
    return d.get("kernels", {}).get(kernel)


def as_fasta(letters, offsets, ids, width=60):
    """FASTA text of encoded sequences (NCBI-order symbols, 60 per line)."""
    sym = np.frombuffer(b"ARNDCQEGHILKMFPSTWYVBJZX*", np.uint8)[letters].tobytes()
    off = offsets.astype(np.int64)
    parts = []
    for k, rid in enumerate(ids):
        s = sym[off[k]:off[k + 1]]
        parts.append(b">" + rid.encode() + b" synthetic\n" +
                     b"".join(s[i:i + width] + b"\n" for i in range(0, len(s), width)))
    return b"".join(parts)


def fasta_ingress(dmr, w):
    """dmr_fasta_parse + encode (host) on the workload's queries and database
    written as FASTA; not part of the timed step."""
    res = {}
    for name, letters, offsets, ids in (("queries", w.q_letters, w.q_offsets, w.q_ids),
                                        ("database", w.db_letters, w.db_offsets, w.db_ids)):
        text = as_fasta(letters, offsets, ids)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            f = dmr.parse_fasta(text)
            times.append(time.perf_counter() - t0)
        same = (np.array_equal(f.letters, letters) and np.array_equal(f.offsets, offsets)
                and f.ids == list(ids))
        ms = float(np.median(times)) * 1e3
        res[name] = {"bytes": len(text), "ms": round(ms, 3), "mb_per_s": round(len(text) / ms / 1e3, 1),
                     "matches_workload": bool(same)}
    log(f"[bench] FASTA ingress: {res}")
    return res


HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def stage_traffic(stage: str, workload: str):
    """Measured HBM bytes of a search stage (rocprofv3 FETCH_SIZE/WRITE_SIZE,
    profiles/hbm_traffic.json), or None when not measured on this workload."""
    try:
        d = json.loads((ROOT / "profiles" / "hbm_traffic.json").read_text())
    except (OSError, ValueError):
        return None
    return d.get("stages", {}).get(stage) if d.get("workload") == workload else None


def stage_rooflines(stats, nshapes, q_letters, workload=None):
    """Seeds -> bands stages against the 8 TB/s HBM peak, from algorithmic
    bytes: seed enumeration reads each query letter once per shape and writes
    12 B per seed (8 B tagged key, 4 B letter index), the sort reads and
    writes each seed once (24 B); the join reads 12 B per query seed and
    writes 8 B per hit plus 4 B of reference letter index read per hit; the
    hit sort reads and writes 8 B per hit (16 B); the ungapped stage reads
    8 B per hit and writes 16 B per segment (residue reads excluded: a lower
    bound); the bands read 24 B per segment and write a 32-byte job. `traffic`
    is the stage's measured HBM bytes (rocprofv3 FETCH_SIZE/WRITE_SIZE) where
    profiled on this workload."""
    E, H, S, J = stats["query_seeds"], stats["hits"], stats["segments"], stats["jobs"]
    rows = {
        "seed_and_sort": (q_letters * nshapes + E * 12 + E * 24, stats["ms_seed"], "seed"),
        "join": (E * 12 + H * 12, stats["ms_join"], "join"),
        "sort_hits": (H * 16, stats["ms_sort_hits"], "sort_hits"),
        "ungapped": (H * 8 + S * 16, stats["ms_ungapped"], "ungapped"),
        "bands": (S * 24 + J * 32, stats["ms_bands"], "bands"),
    }
    out = {}
    for k, (b, ms, pmc) in rows.items():
        gbs = b / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        out[k] = {"bytes": int(b), "ms": round(ms, 4), "achieved_gbs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4),
                  "traffic": stage_traffic(pmc, workload) if workload else None}
    return out


def fixture(name):
    import gzip
    g = ROOT / "tests" / "golden" / "search" / f"{name}.tsv.gz"
    return gzip.open(g, "rb").read().decode() if g.exists() else None


FIX_FIELDS = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue bitscore "
              "score cigar")


def leg_very_sensitive(dmr, synth, steps):
    """--very-sensitive, 4096 queries vs 200,000 proteins (60 M letters),
    one batch (sliced by seed hits)."""
    t = time.time()
    w = synth.make("blastp-very-sensitive-4096")
    gen_s = time.time() - t
    ctx = dmr.Context(0)
    ctx.db_load(w.db_letters, w.db_offsets)
    index_ms = ctx.build_index("very-sensitive")
    ctx.queries_load(w.q_letters, w.q_offsets)
    cfg = dmr.SearchConfig(sensitivity="very-sensitive")
    qids, sids = dmr.ids_text(w.q_ids), dmr.ids_text(w.db_ids)
    r = ctx.search_loaded(cfg)  # warmup
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        r = ctx.search_loaded(cfg)
        text = r.format(qids, sids)
        times.append(time.perf_counter() - t0)
        log(f"[bench] very-sensitive step: {times[-1] * 1e3:.0f} ms")
    st = r.stats()
    slices = int(ctx.get_tuning("last_slices"))
    # DP cells of every slice (profiling adds work, so its own run).
    ctx.set_tuning("profile", 1)
    rp = ctx.search_loaded(cfg)
    cells = ctx.last_dp_cells()
    sw_ms = rp.stats()["ms_sw"]
    ctx.set_tuning("profile", 0)
    # Parity: 64 of the queries as their own batch against the oracle's text.
    exp = fixture("blastp-very-sensitive-4096-q64")
    sub = w.subset_queries(list(range(0, 4096, 64)))
    got = ctx.blastp(sub.q_letters, sub.q_offsets, cfg).format(sub.q_ids, w.db_ids, FIX_FIELDS)
    ctx.close()
    ms = float(np.mean(times)) * 1e3
    return {"workload": f"{w.name}: blastp --very-sensitive, {w.n_q} queries vs {w.n_db} proteins / "
                        f"{int(w.db_offsets[-1])} letters, one batch",
            "value": round(w.n_q / (ms / 1e3), 2), "unit": "queries/s", "ms_per_step": round(ms, 1),
            "steps": steps, "index_build_ms": round(index_ms, 1), "generate_s": round(gen_s, 1),
            "slices": slices, "hsps": text.count("\n"), "dp_cells": int(cells),
            "gcups": round(cells / (sw_ms / 1e3) / 1e9, 1) if sw_ms > 0 else None,
            "stage_ms": {k[3:]: round(v, 2) for k, v in st.items() if k.startswith("ms_")},
            "counts": {k: int(v) for k, v in st.items() if not k.startswith("ms_")},
            "stage_rooflines": stage_rooflines(st, 16, int(w.q_offsets[-1])),
            "bit_exact_subset": {"outfmt6_vs_oracle": got == exp, "lines": exp.count("\n") if exp else 0,
                                 "sample": "queries 0, 64, ... (64) as one batch"}}


def leg_blastx(dmr, synth, steps):
    """blastx: 2048 reads of 150-3000 bases vs the 60 M letter DB, --default;
    frames translated and masked on the device."""
    t = time.time()
    w = synth.make_reads("blastx-2048")
    gen_s = time.time() - t
    ctx = dmr.Context(0)
    ctx.db_load(w.db_letters, w.db_offsets)
    index_ms = ctx.build_index("default")
    cfg = dmr.SearchConfig(sensitivity="default")
    rids, sids = dmr.ids_text(w.read_ids), dmr.ids_text(w.db_ids)
    # A step is about 10 ms: a few warmups (the first calls also size the
    # read and frame buffers) and at least 5 timed steps.
    for _ in range(3):
        r = ctx.blastx(w.reads, cfg)
    steps = max(steps, 5)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        r = ctx.blastx(w.reads, cfg)
        text = r.format(rids, sids)
        times.append(time.perf_counter() - t0)
    st = r.stats()
    import importlib
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    mg = importlib.import_module("make_search_golden")
    exp = fixture("blastx-2048-r64")
    _, idx, _ = mg.READ_SUBSETS["blastx-2048-r64"]
    got = ctx.blastx([w.reads[i] for i in idx], cfg).format([w.read_ids[i] for i in idx], w.db_ids,
                                                            mg.READ_FIELDS)
    ctx.close()
    ms = float(np.mean(times)) * 1e3
    nt = sum(len(x) for x in w.reads)
    return {"workload": f"{w.name}: blastx --default, {len(w.reads)} reads ({nt} bases) vs {w.n_db} proteins / "
                        f"{int(w.db_offsets[-1])} letters",
            "value": round(len(w.reads) / (ms / 1e3), 2), "unit": "reads/s", "ms_per_step": round(ms, 2),
            "steps": steps, "index_build_ms": round(index_ms, 1), "generate_s": round(gen_s, 1),
            "translate_ms": round(st["ms_translate"], 3), "hsps": text.count("\n"),
            "note": "the step includes the reads' upload (host buffers in) and the device translation",
            "stage_ms": {k[3:]: round(v, 2) for k, v in st.items() if k.startswith("ms_")},
            "counts": {k: int(v) for k, v in st.items() if not k.startswith("ms_")},
            "bit_exact_subset": {"outfmt6_vs_oracle": got == exp, "lines": exp.count("\n") if exp else 0,
                                 "sample": "reads 0, 32, ... (64) as one batch"}}


def leg_serving(dmr, synth, duration_s=10.0):
    """Serving loop: 1,000,000 proteins (~300 M letters) resident with the
    default, sensitive and very-sensitive indexes; requests of the pooled
    queries (lengths log-normal median 350, 30..5000), presets 50/40/10.
    First a burst (all submitted at once) gives the saturated throughput;
    then a Poisson stream of 200 queries/s (at most half the burst's rate)
    gives per-request latency. The burst's rate depends on how its requests
    fall into batches (it varied from 370 to 780 queries/s between runs), so
    the stream's rate is fixed rather than taken from it."""
    import threading
    t = time.time()
    w = synth.make_serving("serve-1m")
    gen_s = time.time() - t
    ctx = dmr.Context(0)
    t0 = time.perf_counter()
    ctx.db_load(w.db_letters, w.db_offsets)
    load_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    srv = dmr.Server(ctx, workers=2, hbm_budget=32 << 30, presets="default,sensitive,very-sensitive")
    index_s = time.perf_counter() - t0
    log(f"[bench] serving: DB loaded in {load_s:.1f} s, three indexes in {index_s:.1f} s")
    cfgs = {p: dmr.SearchConfig(sensitivity=p) for p in ("default", "sensitive", "very-sensitive")}

    def run(stream):
        lat = [None] * len(stream)
        tickets = [None] * len(stream)
        ready = threading.Semaphore(0)

        def collect():
            for k in range(len(stream)):
                ready.acquire()
                _, lat[k] = srv.wait(tickets[k])

        th = threading.Thread(target=collect)
        th.start()
        start = time.perf_counter()
        for k, (at, qi, preset) in enumerate(stream):
            d = start + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[k] = srv.submit(w.pool[qi], cfgs[preset])
            ready.release()
        th.join()
        return lat, time.perf_counter() - start

    # Warm every preset's kernels and buffers.
    run(w.stream(64, 1e9, seed=1))
    burst = w.stream(1024, 1e9, seed=2)
    lat_b, wall_b = run(burst)
    capacity = len(burst) / wall_b
    qps = max(10.0, min(200.0, round(capacity * 0.5)))
    n = int(min(4000, max(200, qps * duration_s)))
    stream = w.stream(n, qps, seed=3)
    lat, wall = run(stream)
    st = srv.stats()
    # A sample of served requests equals the same queries searched alone.
    sample = w.stream(24, 1e9, seed=4)
    served = []
    for _, qi, preset in sample:
        served.append((qi, preset, srv.submit(w.pool[qi], cfgs[preset])))
    same = True
    sids = dmr.ids_text(w.db_ids)
    for qi, preset, tk in served:
        res, _ = srv.wait(tk)
        q = w.pool[qi]
        alone = ctx.blastp(q, np.array([0, len(q)], np.uint64), cfgs[preset])
        same &= res.format(["q"], sids) == alone.format(["q"], sids)
    srv.close()
    ctx.close()
    lat = np.array(lat)
    return {"workload": f"{w.name}: {w.n_db} proteins / {int(w.db_offsets[-1])} letters resident "
                        "(default + sensitive + very-sensitive indexes), pooled queries of log-normal "
                        "length (median 350, 30..5000), presets 50/40/10, Poisson arrivals",
            "qps_offered": qps, "requests": n, "value": round(n / wall, 2), "unit": "queries/s",
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2),
                           "mean": round(float(lat.mean()), 2), "of": "arrival to result, per request"},
            "saturated": {"queries_per_s": round(capacity, 1), "requests": len(burst),
                          "latency_p50_ms": round(float(np.percentile(lat_b, 50)), 1)},
            "server": st, "db_load_s": round(load_s, 1), "index_build_s": round(index_s, 1),
            "generate_s": round(gen_s, 1),
            "served_equals_alone_on_sample": bool(same), "sample": len(sample)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="blastp-sensitive-1024")
    ap.add_argument("--cpu-sample", type=int, default=96, help="queries timed on the CPU oracle")
    ap.add_argument("--plain-sample", type=int, default=4, help="queries timed on the plain form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--legs", default="very-sensitive,blastx,serving",
                    help="other configurations to run after the headline one ('' for none)")
    ap.add_argument("--leg-steps", type=int, default=2)
    ap.add_argument("--streams", type=int, default=2,
                    help="batches in flight for the headline throughput: host threads, each with a context "
                         "sharing the resident database and index (1 = one batch at a time)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="pin a score-pass setting (dmr_ctx_set_tuning) over the tuning file, e.g. merged_mask=63")
    a = ap.parse_args()
    pins = {}
    for t in a.tune:
        k, _, v = t.partition("=")
        pins[k] = float(v)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported "
              "(the multi-GPU split is a design only, not run).", file=sys.stderr)
        sys.exit(2)

    dmr = load("diamond_rust_amd", ROOT / "diamond-rust_amd" / "__init__.py", pkg=True)
    synth = load("diamond_rust_amd.synth", ROOT / "diamond-rust_amd" / "synth.py")
    t = time.time()
    w = synth.make(a.workload)
    sens = w.meta["sensitivity"]
    log(f"[bench] workload {w.name}: {w.n_q} queries, {w.n_db} DB proteins, "
        f"{int(w.db_offsets[-1])} letters ({time.time() - t:.1f}s to generate)")
    # Ids as the newline-separated text dmr_result_format reads (as a serving
    # loop holds them), joined once.
    qids, sids = dmr.ids_text(w.q_ids), dmr.ids_text(w.db_ids)
    cfg = dmr.SearchConfig(sensitivity=sens)

    ctx = dmr.Context(0)
    for k in sorted(pins, key=lambda k: (not k.startswith("lanes_"), k)):
        ctx.set_tuning(k, pins[k])
    ctx.db_load(w.db_letters, w.db_offsets)
    index_ms = ctx.build_index(sens)
    log(f"[bench] resident seed index built in {index_ms:.1f} ms")
    ctx.queries_load(w.q_letters, w.q_offsets)

    for k in range(a.warmup):
        r = ctx.search_loaded(cfg)
        txt = r.format(qids, sids)
        log(f"[bench] warmup {k}: {r.stats()['ms_total']:.1f} ms, {len(r)} HSPs")
    steps, sw_ms, stats = [], [], None
    split = {"search_call": [], "format": []}
    text = None
    for k in range(a.steps):
        t0 = time.perf_counter()
        r = ctx.search_loaded(cfg)
        t1 = time.perf_counter()
        text = r.format(qids, sids)
        dt = time.perf_counter() - t0
        split["search_call"].append((t1 - t0) * 1e3)
        split["format"].append((time.perf_counter() - t1) * 1e3)
        steps.append(dt)
        stats = r.stats()
        sw_ms.append(stats["ms_sw"])
        log(f"[bench] step {k}: {dt * 1e3:.1f} ms")
    try:
        cells = ctx.last_dp_cells()
        prof = ctx.class_profile()
    except dmr.DmrError:  # a workload searched in slices: count its cells in a profiled run of its own
        ctx.set_tuning("profile", 1)
        ctx.search_loaded(cfg)
        cells = ctx.last_dp_cells()
        prof = ctx.class_profile()
        ctx.set_tuning("profile", 0)
    ms1 = float(np.mean(steps)) * 1e3
    qps1 = w.n_q / (ms1 / 1e3)
    lat1 = [float(np.percentile(np.array(steps) * 1e3, q)) for q in (50, 99)]
    # Headline throughput: the same steps with `streams` batches in flight.
    # Each extra context shares the first one's resident database and index
    # (dmr_ctx_share_db) and holds its own copy of the query batch; host
    # threads run whole steps (search + outfmt-6), so one batch's host work,
    # short kernels and traceback overlap another's score pass. Every step's
    # text must equal the one-at-a-time text.
    ms, qps, lat, same_text = ms1, qps1, lat1, True
    if a.streams > 1:
        import threading
        ctxs = [ctx]
        for _ in range(a.streams - 1):
            c = dmr.Context(0)
            for k in sorted(pins, key=lambda k: (not k.startswith("lanes_"), k)):
                c.set_tuning(k, pins[k])
            c.share_db(ctx)
            c.queries_load(w.q_letters, w.q_offsets)
            for _ in range(a.warmup):
                same_text &= c.search_loaded(cfg).format(qids, sids) == text
            ctxs.append(c)
        times, errors = [], []

        def worker(c, n):
            try:
                for _ in range(n):
                    t0 = time.perf_counter()
                    txt = c.search_loaded(cfg).format(qids, sids)
                    times.append(time.perf_counter() - t0)
                    if txt != text:
                        errors.append("text differs")
            except Exception as e:  # reported after the join
                errors.append(f"{type(e).__name__}: {e}")

        share = [a.steps // a.streams + (1 if i < a.steps % a.streams else 0) for i in range(a.streams)]
        th = [threading.Thread(target=worker, args=(c, n)) for c, n in zip(ctxs, share)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
        if errors:
            raise RuntimeError(f"concurrent steps failed: {errors[0]}")
        same_text &= not errors
        ms = wall * 1e3 / a.steps
        qps = w.n_q / (ms / 1e3)
        lat = [float(np.percentile(np.array(times) * 1e3, q)) for q in (50, 99)]
        for c in ctxs[1:]:
            c.close()
        log(f"[bench] {a.streams} batches in flight: {qps:.1f} queries/s, {ms:.2f} ms per step, "
            f"latency p50 {lat[0]:.1f} ms")
    sw = float(np.mean(sw_ms))
    gcups = cells / (sw / 1e3) / 1e9
    # Roofline of the dominant score kernel (the class with the most DP
    # cells): algorithmic cells of its jobs x 11 16-bit ops / its launch time.
    # The classes run concurrently on separate streams, so a launch's time
    # includes sharing the chip; the score pass as a whole is reported too.
    dom = max(prof, key=lambda k: prof[k]["cells"])
    dcells, dms = prof[dom]["cells"], max(prof[dom]["ms"], 1e-9)
    achieved = dcells * OPS_PER_CELL / (dms / 1e3) / 1e12
    kname = prof[dom]["kernel"]
    traffic = hbm_traffic(kname, w.name)
    p50, p99 = lat
    log(f"[bench] one batch at a time: {qps1:.1f} queries/s, {ms1:.1f} ms/step, SW {sw:.2f} ms for {cells} "
        f"cells = {gcups:.1f} GCUPS")

    out = {
        "metric": METRIC, "value": round(qps, 2), "unit": "queries/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f16-held exact integers (< 2048; int32 re-run from 1900)" if "F16" in kname else "int16")
                 if kname.startswith("k_sw16") else "int32",
        "data": "synthetic (seeded generator diamond-rust_amd/synth.py)",
        "config": {"workload": f"{w.name}: blastp --sensitive, {w.n_q} queries vs {w.n_db} "
                               f"proteins / {int(w.db_offsets[-1])} letters",
                   "global_batch": w.n_q,
                   "parallelism": "single GPU" + (f", {a.streams} batches in flight (contexts sharing the "
                                                  "resident database and index)" if a.streams > 1 else ""),
                   **({"tuning_pins": pins} if pins else {})},
        "gcups": round(gcups, 2),
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3),
                       "of": f"one {w.n_q}-query batch, start to text, {a.streams} in flight"},
        "one_batch_at_a_time": {"value": round(qps1, 2), "ms_per_step": round(ms1, 3),
                                "latency_ms": {"p50": round(lat1[0], 3), "p99": round(lat1[1], 3)},
                                "note": "the per-class times, GCUPS and roofline come from these steps"},
        "in_flight_text_equal": same_text,
        "index_build_ms": round(index_ms, 2),
        "stage_ms": {k[3:]: round(v, 3) for k, v in stats.items() if k.startswith("ms_")},
        "step_split_ms": {k: round(float(np.mean(v)), 3) for k, v in split.items()},
        "counts": {k: int(v) for k, v in stats.items() if not k.startswith("ms_")},
        "dp_cells": int(cells),
        "hsps": text.count("\n"),
        "roofline": {"bound": "valu",
                     "kernel": kname,
                     "achieved": round(achieved, 3), "peak": PACKED16_VALU_PEAK_TOPS,
                     "unit": "Tops/s (16-bit lanes)", "frac": round(achieved / PACKED16_VALU_PEAK_TOPS, 4),
                     "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                     "copy-calibrated; profiles/hbm_traffic.json)",
                     "gcups": round(dcells / (dms / 1e3) / 1e9, 2),
                     "basis": f"{OPS_PER_CELL} 16-bit ops per reference DP cell x cells of the "
                              "launch / launch time (HIP events on the stream it runs on)"},
        "score_pass_classes": prof,
        "roofline_score_pass": {
            "bound": "valu", "achieved": round(cells * OPS_PER_CELL / (sw / 1e3) / 1e12, 3),
            "peak": PACKED16_VALU_PEAK_TOPS, "unit": "Tops/s (16-bit lanes)",
            "frac": round(cells * OPS_PER_CELL / (sw / 1e3) / 1e12 / PACKED16_VALU_PEAK_TOPS, 4),
            "basis": "all banded-DP cells x 11 16-bit ops / score-pass wall time (all classes, "
                     "concurrent streams joined)"},
    }

    # The timed batch's text against the oracle's search of the same whole
    # batch, stored as a fixture (tests/golden/make_search_golden.py): its
    # first 12 columns are the default fields the step renders.
    gold = ROOT / "tests" / "golden" / "search" / f"{w.name}.tsv.gz"
    if gold.exists():
        import gzip
        exp = "".join("\t".join(l.split("\t")[:12]) + "\n"
                      for l in gzip.open(gold, "rb").read().decode().splitlines())
        out["bit_exact_full_batch"] = {"outfmt6_vs_oracle": text == exp, "lines": exp.count("\n"),
                                       "fixture": str(gold.relative_to(ROOT))}
        log(f"[bench] timed batch == oracle fixture: {text == exp}")

    out["stage_rooflines"] = stage_rooflines(stats, 2 if sens in ("default", "fast", "faster") else 16,
                                             int(w.q_offsets[-1]),
                                             w.name)
    out["fasta_ingress"] = fasta_ingress(dmr, w)

    if not a.no_plain and a.plain_sample > 0:
        sub = w.subset_queries(list(range(a.plain_sample)))
        t0 = time.perf_counter()
        rp = ctx.blastp_plain(sub.q_letters, sub.q_offsets, cfg)
        tp = time.perf_counter() - t0
        rb = ctx.blastp(sub.q_letters, sub.q_offsets, cfg)
        same = rp.format(sub.q_ids, sids) == rb.format(sub.q_ids, sids)
        out["plain_form"] = {"value": round(sub.n_q / tp, 3), "unit": "queries/s",
                             "sample": f"first {sub.n_q} queries, one launch per reference call",
                             "bit_exact_vs_batched": same,
                             "batched_over_plain": round(qps1 / (sub.n_q / tp), 1)}
        log(f"[bench] plain form: {sub.n_q / tp:.2f} queries/s, bit-exact vs batched: {same}")
        ctx.queries_load(w.q_letters, w.q_offsets)

    if not a.no_cpu and a.cpu_sample > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle
        pyoracle.lib()
        sub = w.subset_queries(list(range(a.cpu_sample)))
        threads = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        ref = pyoracle.search(sub.db_letters, sub.db_offsets, sub.q_letters, sub.q_offsets,
                              sensitivity=sens, threads=threads)
        tc = time.perf_counter() - t0
        exp = ref.format(sub.q_ids, w.db_ids)
        got = ctx.blastp(sub.q_letters, sub.q_offsets, cfg).format(sub.q_ids, sids)
        out["cpu_baseline"] = {
            "value": round(sub.n_q / tc, 3), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {sub.n_q} of the {w.n_q} queries, same DB, full search::run "
                      f"restatement ({tc:.1f} s)"}
        out["bit_exact"] = {"outfmt6_gpu_vs_cpu_on_sample": got == exp, "lines": exp.count("\n")}
        log(f"[bench] CPU oracle: {sub.n_q / tc:.2f} queries/s on {threads} threads; "
            f"GPU == CPU outfmt6: {got == exp}")
    ctx.close()
    legs = {}
    runners = {"very-sensitive": lambda: leg_very_sensitive(dmr, synth, a.leg_steps),
               "blastx": lambda: leg_blastx(dmr, synth, a.leg_steps),
               "serving": lambda: leg_serving(dmr, synth)}
    for name in [x for x in a.legs.split(",") if x]:
        t0 = time.time()
        try:
            legs[name] = runners[name]()
        except Exception as e:  # a failing leg is reported, the headline line still prints
            legs[name] = {"error": f"{type(e).__name__}: {e}"}
        legs[name]["wall_s"] = round(time.time() - t0, 1)
        log(f"[bench] leg {name}: {json.dumps(legs[name])[:400]}")
    if legs:
        out["legs"] = legs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
