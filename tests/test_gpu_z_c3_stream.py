"""The HBM-scale stream at its configured size: 50,000,000 vertices,
average degree 7 (350 M CSR entries), 20 batches (90/1/9), seed 42.

* Capped leg (maxIterations 40, so the CPU oracle finishes in seconds per
  batch; from batch 4 on the frontiers hold millions of vertices and the
  oracle would need minutes): the first four batches match the oracle
  exactly, the first three at every step -- the
  CSR with its -1 holes, the affected set, G' in build order, the component
  ids, the initial labels, the frontier sizes -- and then in labels and
  per-batch counts.
* Full leg (maxIterations 100,000, all 20 batches): size-independent
  properties -- labels in [0, 1], ground truth held at its class, every
  vertex that never arrived at its initial 0.5 -- and the same labels and
  per-batch iteration / update counts under two different engine settings.
"""
import os
import time

import numpy as np
import pytest

import labelpropcuda_amd as L
from oracle.oracle import Oracle, set_threads

pytestmark = pytest.mark.gpu
N, DEG, B = 50_000_000, 7, 20


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def c3():
    t0 = time.perf_counter()
    inp = L.generate(N, DEG, B, 42)
    print(f"C3 inputs generated in {time.perf_counter() - t0:.0f} s", flush=True)
    assert inp.numNodes == N and inp.numEdges == 2 * N * DEG // 2 and inp.numBatches == B
    return inp


def check_step_trace(tr, orc, b):
    rp, ci, vv = orc.csr()
    nnz = int(tr.rowPtr[-1])
    assert np.array_equal(tr.rowPtr, rp), f"batch {b}: rowPtr"
    assert np.array_equal(tr.colInd[:nnz], ci), f"batch {b}: colInd with holes"
    assert np.array_equal(bits(tr.values[:nnz]), bits(vv)), f"batch {b}: values"
    del rp, ci, vv
    assert np.array_equal(tr.affected, orc.affected()), f"batch {b}: affected set"
    ss, sd, sw = orc.subgraph()
    ns = tr.t.numSubEdges
    assert ns == ss.shape[0], f"batch {b}: |G'|"
    assert np.array_equal(tr.subSrc[:ns], ss) and np.array_equal(tr.subDst[:ns], sd), \
        f"batch {b}: G' edges"
    assert np.array_equal(bits(tr.subW[:ns]), bits(sw)), f"batch {b}: G' weights"
    comp, nc = orc.component_ids()
    assert tr.t.numComponents == nc, f"batch {b}: numComponents"
    assert np.array_equal(tr.compId, comp), f"batch {b}: component ids"
    assert np.array_equal(bits(tr.initLabels), bits(orc.init_labels())), \
        f"batch {b}: initial labels"
    fs = orc.frontier_sizes()
    nf = min(len(fs), tr.t.numFrontierSizes)
    assert np.array_equal(tr.frontier[:nf], fs[:nf]), f"batch {b}: frontier sizes"


@pytest.mark.timeout(900)
def test_c3_capped_stream_matches_oracle(gpu_available, c3):
    cap, traced, checked = 40, 3, 4
    cfg = L.make_config(c3, maxIterations=cap)
    orc = Oracle(c3, maxIterations=cap)
    # the oracle's frontier-parallel propagation (same results as the
    # sequential restatement) on the host cores
    set_threads(max(1, min(16, os.cpu_count() or 1)))
    try:
        run_capped(c3, cfg, orc, cap, traced, checked)
    finally:
        set_threads(1)


def run_capped(c3, cfg, orc, cap, traced, checked):
    with L.Context(c3, cfg) as ctx:
        assert ctx.tau == orc.tau
        tr = ctx.enable_trace(maxFrontierSizes=cap + 1)
        ctx.reset()
        for b in range(checked):
            if b == traced:
                ctx.disable_trace()
            t0 = time.perf_counter()
            ctx.run_batch(b)
            t1 = time.perf_counter()
            orc.run_batch(b)
            print(f"C3 capped batch {b}: GPU {t1 - t0:.1f} s, oracle "
                  f"{time.perf_counter() - t1:.1f} s", flush=True)
            if b < traced:
                check_step_trace(tr, orc, b)
            st, ost = ctx.stats(b), orc.stats()
            assert (st["iterations"], st["updates"], st["frontier0"], st["numMarked"]) == \
                (ost["iterations"], ost["updates"], ost["frontier0"], ost["numMarked"]), \
                f"batch {b}: counts"
            assert st["accuracy"] == ost["accuracy"], f"batch {b}: accuracy"
            assert np.array_equal(bits(ctx.labels()), bits(orc.labels())), f"batch {b}: labels"


def full_run(c3, **tuning):
    with L.Context(c3, L.make_config(c3)) as ctx:
        if tuning:
            ctx.set_tuning(**tuning)
        ctx.reset()
        counts = []
        for b in range(B):
            ctx.run_batch(b)
            st = ctx.stats(b)
            counts.append((st["iterations"], st["updates"], st["numSubEdges"],
                           st["numComponents"]))
            print(f"C3 batch {b}: {st['iterations']} iterations, {st['updates']} updates, "
                  f"{st['timeMs']:.0f} ms", flush=True)
        return ctx.labels(), counts


@pytest.mark.timeout(900)
def test_c3_full_stream_properties(gpu_available, c3):
    lab, counts = full_run(c3)
    assert np.all(np.isfinite(lab)) and lab.min() >= 0.0 and lab.max() <= 1.0
    gt_ids, gt_lab = c3.gtIds, c3.gtLab
    assert np.array_equal(lab[gt_ids], np.where(gt_lab < 0.5, 0.0, 1.0).astype(np.float32)), \
        "ground truth not held at its class"
    arrived = np.zeros(N, bool)
    arrived[c3.ins] = True
    assert np.all(lab[~arrived] == 0.5)
    assert sum(u for _, u, _, _ in counts) > 0
    # another engine setting: 16 lanes per frontier vertex, 2,048 blocks,
    # chunks of 64, the LDS kernel's second work-group shape
    lab2, counts2 = full_run(c3, groupLanes=16, blocks=2048, chunk=64, localShape=2)
    assert counts2 == counts, "per-batch counts differ between engine settings"
    assert np.array_equal(bits(lab2), bits(lab)), "labels differ between engine settings"
