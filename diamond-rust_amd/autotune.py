"""Autotuner of the score-pass kernels: picks, on the device, the packed
kernel of every width class (lanes per DP pair and threads per block among
the compiled ones), the merged-band width buckets in use, the arithmetic
(f16 or int16 pairs), which sequence runs along a packed DP's rows, the
stream plan and whether the chain before the score pass runs as one
hipGraph, for a search preset on a workload,
and writes the choice with every trial's timing to the tuning file that the
library loads at context creation (tuning_<arch>.json beside libdmr.so).

None of these settings changes a result: every trial's outfmt-6 text must
equal the first one's, or the tuner stops.

Search: coordinate descent from the built-in defaults. Each step tries the
values of one setting with the others held, and keeps a value only when its
median search time beats the best so far by more than --min-gain (noise
guard): f16; the graph; the fused launch, alone and with every merged-band bucket in
use; the merged-band kernels in use (one bucket bound added or removed at a time, until
no change helps); the row sequence
(transpose_<class>) and the kernel of each class that has DPs in the
workload; the stream plan; the merged buckets once more. A trial's time is
the median of --reps searches, or, with --in-flight N, the wall time per
batch with N contexts sharing the database searching at once.

Usage (on a GPU box; the workloads are the seeded synthetic ones):
  python diamond-rust_amd/autotune.py                       # every preset below
  python diamond-rust_amd/autotune.py --preset sensitive --reps 7
"""
from __future__ import annotations

import argparse
import datetime
import importlib.util
import json
import os
import pathlib
import statistics
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent

# preset -> (synthetic workload, queries used: the first n of its batch)
PRESETS = {
    "sensitive": ("blastp-sensitive-1024", 1024),
    "default": ("blastp-sensitive-1024", 1024),
    "very-sensitive": ("blastp-very-sensitive-4096", 512),
}
# Batches in flight per trial, by preset (default 1: a search alone). With two
# in flight the trials of the headline preset differ by less than their
# spread (about 1 %), so it too is tuned on a search alone; --in-flight
# measures the other way.
IN_FLIGHT = {}
MERGED = ["m97", "m129", "m161", "m193", "m225", "m257", "m385", "m513"]
MERGED_ALWAYS = MERGED.index("m257")  # always in use


def _load(name, path, pkg=False):
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[str(path.parent)] if pkg else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print("[autotune]", *a, flush=True)


class Tuner:
    """ctxs[0] is the context tuned; with several contexts (sharing its
    database), a trial's time is the wall time per batch with one batch in
    flight on each (host threads), the way a serving loop or the benchmark's
    headline runs them; with one, the time of a search alone."""

    def __init__(self, dmr, ctxs, cfg, ids, reps, min_gain):
        self.dmr, self.ctxs, self.cfg, self.ids = dmr, ctxs, cfg, ids
        self.ctx = ctxs[0]
        self.reps, self.min_gain = reps, min_gain
        self.trials = []
        self.ref_text = None

    def apply(self, settings):
        # Lanes before threads: a lanes_ key resets its class's block size.
        for c in self.ctxs:
            for k in sorted(settings, key=lambda k: (not k.startswith("lanes_"), k)):
                c.set_tuning(k, settings[k])

    def _in_flight(self):
        import threading
        errors = []

        def work(c):
            try:
                for _ in range(2):
                    if c.search_loaded(self.cfg).format(self.ids[0], self.ids[1]) != self.ref_text:
                        errors.append("text differs")
            except Exception as e:  # reported after the join
                errors.append(f"{type(e).__name__}: {e}")

        th = [threading.Thread(target=work, args=(c,)) for c in self.ctxs]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errors:
            raise SystemExit(f"autotune: concurrent searches failed: {errors[0]}")
        return (time.perf_counter() - t0) * 1e3 / (2 * len(self.ctxs))

    def measure(self, settings, changed):
        self.apply(settings)
        r = self.ctx.search_loaded(self.cfg)  # warm-up (and the result check)
        text = r.format(self.ids[0], self.ids[1])
        if self.ref_text is None:
            self.ref_text = text
        elif text != self.ref_text:
            raise SystemExit(f"autotune: the output changed under {changed}; no setting may change a result")
        for c in self.ctxs[1:]:
            c.search_loaded(self.cfg)  # warm-up
        times, sw = [], []
        for _ in range(self.reps):
            if len(self.ctxs) > 1:
                times.append(self._in_flight())
                r = self.ctx.search_loaded(self.cfg)
            else:
                t0 = time.perf_counter()
                r = self.ctx.search_loaded(self.cfg)
                times.append((time.perf_counter() - t0) * 1e3)
            sw.append(r.stats()["ms_sw"])
        ms = statistics.median(times)
        self.trials.append(dict(changed=changed, ms=round(ms, 3), score_pass_ms=round(statistics.median(sw), 3)))
        log(f"{json.dumps(changed)} {ms:.2f} ms (score pass {statistics.median(sw):.2f} ms)")
        return ms

    def run(self, variants):
        cur = self.ctx.layout()
        best = self.measure(cur, {"defaults": True})
        default_ms = best

        def consider(cand, changed):
            nonlocal cur, best
            ms = self.measure(cand, changed)
            won = ms < best * (1 - self.min_gain)
            self.trials[-1]["accepted"] = won
            if won:
                cur, best = cand, ms
            self.apply(cur)
            return won

        consider({**cur, "f16": 1 - cur["f16"]}, {"f16": 1 - cur["f16"]})
        consider({**cur, "graph": 1 - cur["graph"]}, {"graph": 1 - cur["graph"]})
        consider({**cur, "fuse": 1 - cur["fuse"]}, {"fuse": 1 - cur["fuse"]})
        # Every merged bucket pays off only inside the fused launch (apart,
        # each is a launch of its own, with its own tail), which one step of
        # either setting alone does not show.
        every = (1 << len(MERGED)) - 1
        if (cur["fuse"], cur["merged_mask"]) != (1, every):
            consider({**cur, "fuse": 1, "merged_mask": every}, {"fuse": 1, "merged_mask": every})

        def merged_pass():
            changed = True
            while changed:
                changed = False
                for i in range(len(MERGED)):
                    if i == MERGED_ALWAYS:
                        continue
                    m = cur["merged_mask"] ^ (1 << i)
                    if consider({**cur, "merged_mask": m}, {"merged_mask": m}):
                        changed = True

        merged_pass()
        # Classes with DPs under the current buckets.
        self.ctx.set_tuning("profile", 1)
        self.ctx.search_loaded(self.cfg)
        prof = self.ctx.class_profile()
        self.ctx.set_tuning("profile", 0)
        for name in self.dmr.PACKED_CLASSES:
            if prof[name]["jobs"] == 0:
                continue
            key = f"transpose_{name}"
            for t in range(3):
                if t != cur[key]:
                    consider({**cur, key: t}, {key: t})
            for v in variants:
                if v["cls"] != name or (v["lanes"], v["threads"]) == (cur[f"lanes_{name}"], cur[f"threads_{name}"]):
                    continue
                ch = {f"lanes_{name}": v["lanes"], f"threads_{name}": v["threads"]}
                consider({**cur, **ch}, ch)
        for p in range(3):
            if p != cur["score_plan"]:
                consider({**cur, "score_plan": p}, {"score_plan": p})
        merged_pass()
        return cur, best, default_ms


def tune_preset(dmr, synth, preset, reps, min_gain, queries=None, in_flight=1):
    wl, nq = PRESETS[preset]
    nq = queries or nq
    log(f"preset {preset}: workload {wl}, {nq} queries, {in_flight} in flight")
    w = synth.make(wl)
    if nq < w.n_q:
        w = w.subset_queries(list(range(nq)))
    ctxs = [dmr.Context(0)]
    ctxs[0].load_tuning(None)  # trials start from the built-in defaults
    ctxs[0].db_load(w.db_letters, w.db_offsets)
    ctxs[0].queries_load(w.q_letters, w.q_offsets)
    for _ in range(in_flight - 1):
        c = dmr.Context(0)
        c.load_tuning(None)
        c.share_db(ctxs[0])
        c.queries_load(w.q_letters, w.q_offsets)
        ctxs.append(c)
    cfg = dmr.SearchConfig(sensitivity=preset)
    t = Tuner(dmr, ctxs, cfg, (dmr.ids_text(w.q_ids), dmr.ids_text(w.db_ids)), reps, min_gain)
    layout, ms, default_ms = t.run(dmr.sw16_variants())
    for c in reversed(ctxs):
        c.close()
    log(f"preset {preset}: {default_ms:.2f} ms with the defaults -> {ms:.2f} ms")
    return dict(settings=layout, workload=wl, queries=nq, in_flight=in_flight, reps=reps, min_gain=min_gain,
                ms=round(ms, 3), default_ms=round(default_ms, 3), trials=t.trials,
                tuned_at=datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--preset", action="append", choices=sorted(PRESETS), help="repeatable; default: all")
    ap.add_argument("--queries", type=int, default=None, help="queries of the workload used (default per preset)")
    ap.add_argument("--reps", type=int, default=5, help="timed searches per trial (median)")
    ap.add_argument("--min-gain", type=float, default=0.01, help="fraction a trial must gain to be kept")
    ap.add_argument("--in-flight", type=int, default=None,
                    help="batches in flight per trial (contexts sharing the database; default per preset)")
    ap.add_argument("--arch", default=os.environ.get("DMR_OFFLOAD_ARCH", "gfx950"))
    ap.add_argument("--out", default=None, help="tuning file (default: tuning_<arch>.json beside libdmr.so)")
    a = ap.parse_args(argv)
    dmr = _load("diamond_rust_amd", HERE / "__init__.py", pkg=True)
    synth = _load("diamond_rust_amd.synth", HERE / "synth.py")
    out = pathlib.Path(a.out) if a.out else HERE / f"tuning_{a.arch}.json"
    doc = json.loads(out.read_text()) if out.exists() else {}
    doc.setdefault("presets", {})
    doc["arch"] = a.arch
    doc["library"] = dmr.lib().dmr_version().decode()
    doc["variants"] = [[v["cls"], v["width"], v["lanes"], v["threads"]] for v in dmr.sw16_variants()]
    for p in a.preset or list(PRESETS):
        doc["presets"][p] = tune_preset(dmr, synth, p, a.reps, a.min_gain, a.queries,
                                        a.in_flight or IN_FLIGHT.get(p, 1))
        tmp = out.with_suffix(".json.tmp")
        tmp.write_text(json.dumps(doc, indent=1) + "\n")
        os.replace(tmp, out)
        log(f"wrote {out}")
    # The file must load: a context created now takes it.
    ctx = dmr.Context(0)
    ctx.load_tuning(str(out))
    assert ctx.get_tuning("tuned_presets") == len(doc["presets"])
    ctx.close()


if __name__ == "__main__":
    main()
