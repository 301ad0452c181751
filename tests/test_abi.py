"""CPU checks of the C ABI: libdmr.so loads, exports every function that
include/dmr.h declares, and its host-only helpers agree with the oracle."""
import ctypes
import math
import pathlib
import re

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "dmr.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(dmr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_header(dmr):
    lib = ctypes.CDLL(str(dmr.LIB_PATH))
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(dmr.EXPORTS) == names


def test_band_for_matches_oracle(dmr, orc):
    for n in list(range(0, 700)) + [5000, 100000]:
        assert dmr.band_for(n) == orc.band_for(n)


def test_host_evalues_bit_exact(dmr, orc):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        raw = int(rng.integers(-5, 3000))
        q = int(rng.integers(1, 6000))
        s = int(rng.integers(1, 6000))
        db = int(rng.integers(1, 10 ** 10))
        e, b = dmr.evalue(raw, q, s, db)
        eo = orc.evalue(raw, q, s, db)
        assert (e == eo) or (math.isnan(e) and math.isnan(eo)), (raw, q, s, db, e, eo)
        assert b == orc.bit_score(raw)


def test_unsupported_gap_costs_rejected(dmr):
    import pytest
    with pytest.raises(dmr.DmrError) as ei:
        dmr.evalue(50, 100, 100, 1000, gap_open=42, gap_extend=7)
    assert ei.value.code == dmr.DMR_E_ARG


def test_product_tables_match_reference(dmr, orc, synth):
    import json
    g = json.loads((ROOT / "tests" / "golden" / "blosum62.json").read_text())
    # The generator's canonical block and the product's ALP parameters.
    assert synth._B62.astype(int).tolist() == [row[:20] for row in g["scores"][:20]]
    row = [p for p in g["params"] if p[0] == 11 and p[1] == 1][0]
    e, b = dmr.evalue(100, 300, 300, 10 ** 6)
    assert abs(b - (row[2] * 100 - math.log(row[3])) / math.log(2)) < 1e-12


def test_committed_tuning_file_names_compiled_kernels(dmr):
    """The autotuner's file beside the library (loaded by every context):
    every preset's settings name known keys and compiled kernels, and the
    tuned time is no worse than the defaults' (dmr_sw16_variants is host
    code: no GPU needed)."""
    import json
    path = dmr.LIB_PATH.parent / "tuning_gfx950.json"
    doc = json.loads(path.read_text())
    compiled = {(v["cls"], v["lanes"], v["threads"]) for v in dmr.sw16_variants()}
    assert len(compiled) >= 28 and doc["arch"] == "gfx950"
    keys = {f"{k}_{c}" for c in dmr.PACKED_CLASSES for k in ("lanes", "threads", "transpose")} | {
        "merged_mask", "f16", "score_plan", "graph", "fuse"}
    for name, entry in doc["presets"].items():
        s = entry["settings"]
        assert set(s) <= keys, name
        for c in dmr.PACKED_CLASSES:
            assert (c, s[f"lanes_{c}"], s[f"threads_{c}"]) in compiled, (name, c)
        assert 0 <= s["merged_mask"] < 256 and s["score_plan"] in (0, 1, 2)
        assert all(s.get(f"transpose_{c}", 0) in (0, 1, 2) for c in dmr.PACKED_CLASSES)
        assert entry["ms"] <= entry["default_ms"] and len(entry["trials"]) > 10
