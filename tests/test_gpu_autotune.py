"""The autotuner on the GPU, through the C-ABI: timed trials change only the
time (every candidate gives the oracle's labels and counts), the choice goes
into the process's tuning table, and a context or runner created later for
the same shape starts from it."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import dynlp
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def test_context_autotune_keeps_results_and_fills_the_table(gpu_available, c1_inputs, tmp_path):
    inp = c1_inputs
    dynlp.tuning_table_clear()
    try:
        with L.Context(inp, L.make_config(inp)) as ctx:
            key, src = ctx.shape_key()
            assert src == "heuristic"
            chosen, ms = ctx.autotune(trialBatches=6)
            assert ms > 0 and ctx.shape_key()[1] == "autotune"
            assert chosen["propGroupLanes"] in (2, 4, 8) and chosen["localShape"] in (1, 3)
            ctx.run_all()
            orc = Oracle(inp)
            orc.run_all()
            assert np.array_equal(bits(ctx.labels()), bits(orc.labels()))
        with L.Context(inp, L.make_config(inp)) as ctx2:
            assert ctx2.shape_key() == (key, "table")
            t2 = ctx2.tuning()
            assert all(t2[k] == chosen[k] for k in ("propGroupLanes", "localShape",
                                                      "waveRowMin", "blockRowMin"))
        path = str(tmp_path / "t.tune")
        assert dynlp.tuning_table_save(path) == 1
        assert f"context {key['logNodes']} {key['meanDegree']} {key['logMaxDegree']} :" in \
            open(path).read()
    finally:
        dynlp.tuning_table_clear()


def test_multi_autotune_keeps_results_and_fills_the_table(gpu_available):
    graphs = [L.generate(3000 + 700 * g, (3, 5, 7)[g % 3], 4, 60 + g) for g in range(6)]
    cfgs = [L.make_config(g, maxIterations=2000) for g in graphs]
    plain = []
    for g, c in zip(graphs, cfgs):
        with L.Context(g, c) as ctx:
            ctx.run_all()
            plain.append(ctx.labels())
    dynlp.tuning_table_clear()
    try:
        with L.Multi(graphs, cfgs) as m:
            assert m.tuning_source() == "defaults"
            chosen, ms = m.autotune(trialBatches=2)
            assert ms > 0 and m.tuning_source() == "autotune"
            m.run_all()
            m.synchronize()
            for g in range(len(graphs)):
                assert np.array_equal(bits(m.labels(g)), bits(plain[g])), f"graph {g}"
        with L.Multi(graphs, cfgs) as m2:
            assert m2.tuning_source() == "table"
            assert m2.tuning()["localShape"] == chosen["localShape"]
    finally:
        dynlp.tuning_table_clear()
