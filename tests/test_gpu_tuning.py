"""Score-pass tuning: the tuning file's settings per preset, keys pinned over
them, refusal of keys and files that name no compiled kernel, and the
committed file that a context loads by itself. No setting changes a result.
"""
import json

import pytest

from test_gpu_search import EXTENDED, workload_dict

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(synth):
    return workload_dict(synth.make("blastp-sensitive-1024").subset_queries(list(range(0, 1024, 32))))


def search(dmr, ctx, w, preset):
    r = ctx.blastp(w["q_letters"], w["q_offsets"], dmr.SearchConfig(sensitivity=preset))
    return r.format(w["q_ids"], w["db_ids"], EXTENDED)


def test_file_per_preset_and_pinned_keys(dmr, small, tmp_path):
    ctx = dmr.Context(0)
    ctx.load_tuning(None)
    assert ctx.get_tuning("tuned_presets") == 0
    ctx.db_load(small["db_letters"], small["db_offsets"])
    base = {p: search(dmr, ctx, small, p) for p in ("default", "sensitive")}
    f = tmp_path / "t.json"
    f.write_text(json.dumps({"arch": "gfx950", "presets": {
        "*": {"settings": {"merged_mask": 63, "lanes_m97": 4}},
        "sensitive": {"settings": {"lanes_w129": 8, "f16": 0, "score_plan": 2}, "ms": 1.0, "trials": [1, "x"]},
    }}))
    ctx.load_tuning(f)
    assert ctx.get_tuning("tuned_presets") == 2
    assert search(dmr, ctx, small, "sensitive") == base["sensitive"]
    lay = ctx.layout()
    assert (lay["lanes_w129"], lay["threads_w129"], lay["f16"], lay["score_plan"]) == (8, 512, 0, 2)
    assert (lay["merged_mask"], lay["lanes_m97"]) == (63, 4)
    assert search(dmr, ctx, small, "default") == base["default"]
    lay = ctx.layout()
    assert (lay["lanes_w129"], lay["f16"], lay["merged_mask"], lay["lanes_m97"]) == (4, 1, 63, 4)
    # Pinned keys win over every preset's entry.
    ctx.set_tuning("lanes_m97", 8)
    ctx.set_tuning("f16", 1)
    assert search(dmr, ctx, small, "sensitive") == base["sensitive"]
    lay = ctx.layout()
    assert (lay["lanes_m97"], lay["f16"], lay["lanes_w129"]) == (8, 1, 8)
    ctx.close()


def test_refusals_change_nothing(dmr, tmp_path):
    ctx = dmr.Context(0)
    ctx.load_tuning(None)
    before = ctx.layout()
    for key, value, msg in [("lanes_w129", 3, "no compiled kernel"), ("threads_w25", 512, "no compiled kernel"),
                            ("lanes_m97", 1, "no compiled kernel"), ("merged_mask", 256, "8-bit"),
                            ("score_plan", 3, "score_plan"), ("lanes_w130", 4, "unknown tuning key"),
                            ("split_mask", 1, "unknown tuning key")]:
        with pytest.raises(dmr.DmrError, match=msg):
            ctx.set_tuning(key, value)
        assert ctx.layout() == before
    ctx.set_tuning("lanes_w129", 4)
    ctx.set_tuning("threads_w129", 512)
    with pytest.raises(dmr.DmrError, match="768 threads|no compiled kernel"):
        ctx.set_tuning("lanes_w129", 8) or ctx.set_tuning("threads_w129", 768)
    good = tmp_path / "good.json"
    good.write_text(json.dumps({"presets": {"default": {"settings": {"lanes_w61": 2}}}}))
    ctx.load_tuning(good)
    assert ctx.get_tuning("tuned_presets") == 1
    for text, msg in [("{", "malformed|unexpected end"), ('{"presets": {"x": {"settings": {"lanes_w61": 7}}}}',
                                                          "no compiled kernel"),
                      ('{"presets": {"x": {"settings": {"bogus": 1}}}}', "unknown tuning key"),
                      ('{"presets": {"x": {}}}', "settings"), ('{"presets": 3}', "not an object"),
                      ('{"presets": {}} x', "trailing")]:
        bad = tmp_path / "bad.json"
        bad.write_text(text)
        with pytest.raises(dmr.DmrError, match=msg):
            ctx.load_tuning(bad)
        assert ctx.get_tuning("tuned_presets") == 1
    with pytest.raises(dmr.DmrError, match="cannot open"):
        ctx.load_tuning(tmp_path / "missing.json")
    ctx.close()


def test_committed_tuning_file_loads(dmr):
    """tuning_gfx950.json beside the library (the autotuner's output) is
    loaded by every new context and names only compiled kernels."""
    path = dmr.LIB_PATH.parent / "tuning_gfx950.json"
    if not path.exists():
        pytest.skip("no tuning file committed")
    doc = json.loads(path.read_text())
    ctx = dmr.Context(0)
    assert ctx.get_tuning("tuned_presets") == len(doc["presets"]) > 0
    compiled = {(v["cls"], v["lanes"], v["threads"]) for v in dmr.sw16_variants()}
    for name, entry in doc["presets"].items():
        s = entry["settings"]
        for c in dmr.PACKED_CLASSES:
            if f"lanes_{c}" in s:
                assert (c, s[f"lanes_{c}"], s[f"threads_{c}"]) in compiled or s[f"lanes_{c}"] == 1, (name, c)
        assert entry["ms"] <= entry["default_ms"] * 1.05, name
    ctx.close()
