"""The autotuner's host parts without a GPU: the degree-histogram heuristic
(shape key, lanes per vertex, row tiers) and the tuning table's text format
(load, save, reject)."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import dynlp


def test_heuristic_from_degree_histogram(c1_inputs):
    key, t = dynlp.autotune_heuristic(c1_inputs)
    deg = np.diff(c1_inputs.rowPtr)
    assert key["logNodes"] == 13  # 2^12 < 5000 <= 2^13
    assert key["meanDegree"] == round(c1_inputs.numEdges / 5000)
    assert 2 ** (key["logMaxDegree"] - 1) < deg.max() + 1 <= 2 ** key["logMaxDegree"]
    # 90th-percentile row length 7: its class [4, 8) ends at 7, which four
    # lanes cover in two passes
    assert np.percentile(deg, 90) <= 7
    assert t["propGroupLanes"] == 4 and t["waveRowMin"] == 64 and t["blockRowMin"] == -1
    assert t["propBlocks"] == 512  # the grid floor for small graphs
    deg7 = L.generate(40000, 7, 3, 9)  # 90th percentile in [8, 16): eight lanes
    assert 8 <= np.percentile(np.diff(deg7.rowPtr), 90) < 16
    assert dynlp.autotune_heuristic(deg7)[1]["propGroupLanes"] == 8
    dense = L.generate(3000, 60, 3, 9)  # 90th percentile in [64, 128): 32 lanes
    assert dynlp.autotune_heuristic(dense)[1]["propGroupLanes"] == 32


def test_heuristic_turns_the_workgroup_tier_on_for_hubs():
    inp = L.generate(4000, 5, 3, 1)
    # a hub: vertex 0 joined to 1,500 vertices (symmetric CSR rebuilt here)
    rows = [set(inp.colInd[inp.rowPtr[u]:inp.rowPtr[u + 1]].tolist()) for u in range(4000)]
    for v in range(1, 1501):
        rows[0].add(v)
        rows[v].add(0)
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    inp.rowPtr = rp
    inp.colInd = np.concatenate([sorted(r) for r in rows]).astype(np.int32)
    inp.values = np.full(inp.colInd.shape[0], 0.5, np.float32)
    key, t = dynlp.autotune_heuristic(inp)
    assert t["blockRowMin"] == 1024 and key["logMaxDegree"] == 11


def test_tuning_table_round_trip(tmp_path):
    text = ("# a table\n"
            "context 19 7 5 : 8 0 1024 1017 0 3 1024 64 -1 512 ms=412.500\n"
            "multi 8 24 : 1 100 8 1 16384 ms=19000.000\n")
    src = tmp_path / "in.tune"
    src.write_text(text)
    dynlp.tuning_table_clear()
    try:
        assert dynlp.tuning_table_load(str(src)) == 2
        assert dynlp.tuning_table_save(str(tmp_path / "out.tune")) == 2
        out = [ln for ln in (tmp_path / "out.tune").read_text().splitlines()
               if not ln.startswith("#")]
        assert out == text.splitlines()[1:]
        bad = tmp_path / "bad.tune"
        bad.write_text("context 1 2 : 3\n")
        with pytest.raises(L.DynLPError):
            dynlp.tuning_table_load(str(bad))
        with pytest.raises(L.DynLPError):
            dynlp.tuning_table_load(str(tmp_path / "missing.tune"))
    finally:
        dynlp.tuning_table_clear()
