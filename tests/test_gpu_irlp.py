"""IrLP full recompute on the GPU (dlp_irlp, dlp_irlp_validation,
dlp_irlp_baseline and the CLI's --validate) against the oracle, through the
C-ABI.  Labels bit for bit, iteration counts exactly."""
import os
import subprocess

import numpy as np
import pytest

import labelpropcuda_amd as L
import oracle.oracle as O
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
F32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


@pytest.fixture(scope="module")
def final_states():
    out = {}
    for n, d, b, seed in ((5000, 5, 10, 42), (20000, 7, 6, 3), (3000, 40, 5, 9)):
        inp = L.generate(n, d, b, seed)
        out[(n, d)] = (inp, O.irlp_final_state(inp))
    return out


@pytest.mark.parametrize("groupLanes,firstChunk,maxChunk",
                         [(1, 8, 64), (1, 1, 1), (8, 8, 64), (4, 1, 1), (16, 3, 5),
                          (64, 256, 256), (32, 2, 200)])
def test_irlp_matches_oracle(gpu_available, final_states, groupLanes, firstChunk, maxChunk):
    t = L.dynlp.irlp_tuning(groupLanes=groupLanes, firstChunk=firstChunk, maxChunk=maxChunk)
    for key, (_, (rp, ci, vv, st, gt, lab)) in final_states.items():
        exp, ei = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 100000)
        got, gi, ms = L.dynlp.irlp(rp, ci, vv, lab, st, gt, 1e-4, 100000, tuning=t)
        print(f"IrLP {key} G={groupLanes}: {gi} iterations, {ms:.2f} ms", flush=True)
        assert gi == ei, key
        assert np.array_equal(bits(got), bits(exp)), key


@pytest.mark.parametrize("maxIt", [1, 2, 5, 8, 9])
def test_irlp_iteration_cap_and_final_buffer(gpu_available, final_states, maxIt):
    _, (rp, ci, vv, st, gt, lab) = final_states[(5000, 5)]
    exp, ei = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, maxIt)
    got, gi, _ = L.dynlp.irlp(rp, ci, vv, lab, st, gt, 1e-4, maxIt,
                              tuning=L.dynlp.irlp_tuning(firstChunk=4, maxChunk=4))
    assert gi == ei == maxIt
    assert np.array_equal(bits(got), bits(exp))


def test_irlp_random_states_with_holes(gpu_available):
    """Holes (negative columns), deleted neighbours, GT with stale labels,
    isolated vertices and a row longer than 256."""
    from test_oracle import random_graph, random_state
    rng = np.random.default_rng(21)
    for n, p, hub in ((400, 0.02, 300), (2000, 0.004, 0)):
        rowPtr, colInd, values = random_graph(rng, n, p, hub=hub)
        status, gtClass, labels, _, col = random_state(rng, n, colInd)
        labels[gtClass >= 0] = 0.25  # enforcement must overwrite these
        for delta in (1e-4, 0.0):
            exp, ei = O.irlp(rowPtr, col, values, labels, status, gtClass, delta, 500)
            for g in (1, 8):
                got, gi, _ = L.dynlp.irlp(rowPtr, col, values, labels, status, gtClass, delta,
                                          500, tuning=L.dynlp.irlp_tuning(groupLanes=g))
                assert gi == ei
                assert np.array_equal(bits(got), bits(exp))


def test_irlp_validation_host_arrays(gpu_available, final_states):
    _, (rp, ci, vv, st, gt, lab) = final_states[(20000, 7)]
    exp, ei = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 100000)
    got, gi, skipped = L.dynlp.irlp_validation(rp, ci, vv, st, gt, lab, 1e-4, 100000)
    assert not skipped and gi == ei
    assert np.array_equal(bits(got), bits(exp))
    # an empty graph is skipped with the labels untouched (validation.cu:216-219)
    n = 10
    got, gi, skipped = L.dynlp.irlp_validation(np.zeros(n + 1, np.int32), np.zeros(0, np.int32),
                                               np.zeros(0, F32), np.zeros(n, np.int32),
                                               np.full(n, -1, np.int32), np.full(n, 0.5, F32))
    assert skipped and gi == 0 and np.all(got == F32(0.5))


def test_irlp_baseline_matches_replay_and_oracle(gpu_available, final_states):
    for key, (inp, (rp, ci, vv, st, gt, lab)) in final_states.items():
        cfg = L.make_config(inp)
        got, res = L.dynlp.irlp_baseline(inp, cfg)
        exp, ei = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 100000)
        assert res["iterations"] == ei and not res["skipped"], key
        assert res["activeEdges"] == ci.shape[0], key
        assert np.array_equal(bits(got), bits(exp)), key
        keep = (st != 1) & (st != 2)
        acc = np.float32(np.sum((exp[keep] >= 0.5) == (inp.trueLabels[keep] >= 0.5))) / \
            np.float32(keep.sum())
        assert res["accuracy"] == acc, key
        assert res["timeMs"] > 0


def test_cli_validate_writes_irlp_labels(gpu_available, tmp_path):
    exe = os.path.join(ROOT, "labelpropcuda_amd", "bin", "dynLP")
    r = subprocess.run([exe, "--totalNodes", "4000", "--numBatches", "5", "--avgDegree", "5",
                        "--seed", "7", "--validate", "--output", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "IrLP baseline:" in r.stdout and "Speedup vs IrLP:" in r.stdout
    inp = L.generate(4000, 5, 5, 7)
    rp, ci, vv, st, gt, lab = O.irlp_final_state(inp)
    exp, ei = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 100000)
    assert f"IrLP baseline: {ei} iterations" in r.stdout
    rows = np.loadtxt(tmp_path / "irlp_labels.txt")
    assert rows.shape == (4000, 2)
    assert np.array_equal(rows[:, 0].astype(int), np.arange(4000))
    assert np.array_equal(rows[:, 1], np.array([float(f"{x:.6f}") for x in exp]))
    # the DynLP outputs against the oracle: labels file text exactly, the
    # accuracy column of results.csv exactly, times with two decimals
    orc = Oracle(inp)
    accs = []
    for b in range(inp.numBatches):
        orc.run_batch(b)
        accs.append(orc.stats()["accuracy"])
    want = "".join(f"{i} {float(x):.6f}\n" for i, x in enumerate(orc.labels()))
    assert (tmp_path / "dynlp_labels.txt").read_text() == want
    csv = (tmp_path / "results.csv").read_text().splitlines()
    assert csv[0] == "batch,time_ms,accuracy" and len(csv) == 1 + inp.numBatches
    for b, line in enumerate(csv[1:]):
        idx, ms, acc = line.split(",")
        assert int(idx) == b and acc == f"{float(np.float32(accs[b])):.6f}", line
        assert len(ms.split(".")[1]) == 2 and float(ms) >= 0.0, line
    for line in ("  dynlp_labels.txt  \u2014 Final predicted labels",
                 "  results.csv       \u2014 Per-batch timing and accuracy",
                 "  irlp_labels.txt   \u2014 IrLP baseline labels",
                 "  graph/            \u2014 Generated graph in CSR format"):
        assert line in r.stdout.splitlines(), line
    assert "Total DynLP time:" in r.stdout and "Per-batch results:" in r.stdout
