"""Work proportional to the batch: the CSR update that rebuilds only the
rows a batch can change, and the accuracy kept as running totals over the
vertices a batch changes.  Both must leave exactly the state and counts of
the reference's whole-graph passes (dynLP.cu:65-85, :111-167): the oracle's
CSR with its holes, affected set, G', components, labels and accuracy after
every batch, and the same results as the every-row update at scale."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from oracle.oracle import Oracle
from test_gpu_parity import assert_labels_equal, bits, compare_trace

pytestmark = pytest.mark.gpu


def small_batch_stream(n, seed, batches=12):
    rng = np.random.default_rng(seed)
    sizes = [0] + [int(x) for x in rng.integers(20, 600, batches)]
    return L.generate_stream(n, 7, sizes, seed)


@pytest.mark.parametrize("mode", [0, 2])
def test_listed_rows_every_step_matches_oracle(gpu_available, mode):
    """Batches of 20-600 vertices on a 30,000-vertex graph: few enough rows
    change that mode 0 takes the listed-row update too."""
    inp = small_batch_stream(30000, 77)
    with L.Context(inp, L.make_config(inp, maxIterations=3000)) as ctx:
        assert ctx.row_update(mode) is True
        tr = ctx.enable_trace(maxFrontierSizes=1 << 12)
        orc = Oracle(inp, maxIterations=3000)
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            compare_trace(tr, orc, b, inp)
            assert_labels_equal(ctx.labels(), orc.labels(), f"batch {b}: labels")
            st, ost = ctx.stats(b), orc.stats()
            assert (st["iterations"], st["updates"], st["numMarked"]) == \
                (ost["iterations"], ost["updates"], ost["numMarked"]), f"batch {b}"
            assert st["accuracy"] == ost["accuracy"], f"batch {b}: accuracy"


def test_listed_rows_on_the_reference_batches(gpu_available, c1_inputs):
    """Mode 2 forces the listed-row update on the reference's own batch
    stream (10 batches of about 10 % of the graph each): CSR, holes and
    every later step as the oracle, after every batch."""
    inp = c1_inputs
    with L.Context(inp, L.make_config(inp)) as ctx:
        assert ctx.row_update(2) is True
        tr = ctx.enable_trace(maxFrontierSizes=1 << 17)
        orc = Oracle(inp)
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            compare_trace(tr, orc, b, inp)
            assert_labels_equal(ctx.labels(), orc.labels(), f"batch {b}: labels")
            assert ctx.stats(b)["accuracy"] == orc.stats()["accuracy"]


def test_row_modes_agree_at_scale(gpu_available):
    """2,000,000 vertices, batches of 300 to 30,000: the listed-row update
    (mode 2) and the every-row update (mode 1) leave the same CSR after every
    batch and the same labels, counts and accuracy; after a reset and a
    second pass, the same again (the row stamps and running totals restart)."""
    rng = np.random.default_rng(5)
    sizes = [0] + [int(x) for x in np.rint(np.exp(rng.uniform(np.log(300), np.log(30000), 10)))]
    inp = L.generate_stream(2_000_000, 7, sizes, 123)
    cfg = L.make_config(inp, maxIterations=2000)
    with L.Context(inp, cfg) as a, L.Context(inp, cfg) as f:
        assert a.row_update(2) and f.row_update(1)
        ta = a.enable_trace(maxFrontierSizes=16)
        tf = f.enable_trace(maxFrontierSizes=16)
        for rep in range(2):
            a.reset()
            f.reset()
            for b in range(inp.numBatches):
                a.run_batch(b)
                f.run_batch(b)
                sa, sf = ta.snapshot(), tf.snapshot()
                for key in ("rowPtr", "colInd", "affected", "compId"):
                    assert np.array_equal(sa[key], sf[key]), f"pass {rep} batch {b}: {key}"
                assert np.array_equal(bits(sa["values"]), bits(sf["values"]))
                assert np.array_equal(bits(a.labels()), bits(f.labels())), f"batch {b}"
                xa, xf = a.stats(b), f.stats(b)
                for key in ("iterations", "updates", "numMarked", "accuracy", "numTouched"):
                    assert xa[key] == xf[key], f"pass {rep} batch {b}: {key}"
            print(f"pass {rep}: final accuracy {a.stats(inp.numBatches - 1)['accuracy']:.4f}",
                  flush=True)


def asymmetric_inputs(tmp_path, n=4000, keep_reverse=0.7):
    """A graph in which only some edges are stored in both rows: the
    listed-row CSR update and the multi-wave register-mask kernels (which
    read the next frontier off the rows) must not be used on it."""
    rng = np.random.default_rng(8)
    rows = [[] for _ in range(n)]
    seen = set()
    for u, v in rng.integers(0, n, (14000, 2)).tolist():
        if u == v or (u, v) in seen or (v, u) in seen:
            continue
        seen.add((u, v))
        w = float(rng.uniform(0.01, 1.0))
        rows[u].append((v, w))
        if rng.random() < keep_reverse:
            rows[v].append((u, w))
    prefix = str(tmp_path / "asym_")
    ordered = [sorted(r) for r in rows]
    with open(prefix + "RowPtr.txt", "w") as fh:
        fh.write("\n".join(str(x) for x in np.cumsum([0] + [len(r) for r in ordered])) + "\n")
    with open(prefix + "ColInd.txt", "w") as fh:
        fh.write("\n".join(str(v) for r in ordered for v, _ in r) + "\n")
    with open(prefix + "Values.txt", "w") as fh:
        fh.write("\n".join(f"{w:.6f}" for r in ordered for _, w in r) + "\n")
    labels = rng.permutation(np.arange(n) % 2)
    with open(prefix + "TrueLabels.txt", "w") as fh:
        fh.write("\n".join(f"{i} {c}" for i, c in enumerate(labels)) + "\n")
    return L.dynlp.from_graph_files(prefix, 6, 42)


@pytest.mark.parametrize("mode", [0, 2])
def test_asymmetric_graph_matches_oracle(gpu_available, tmp_path, mode):
    """An asymmetric full graph is detected, every mode then rebuilds every
    row, and the engines that need symmetric rows stay off: oracle CSR,
    labels, iterations and accuracy after every batch, with the owner and
    register-mask kernels allowed early (owner mode after 8 iterations)."""
    inp = asymmetric_inputs(tmp_path)
    with L.Context(inp, L.make_config(inp, maxIterations=5000)) as ctx:
        assert ctx.row_update(mode) is False
        ctx.set_tuning(localOwnerAfter=8)
        tr = ctx.enable_trace(maxFrontierSizes=1 << 13)
        orc = Oracle(inp, maxIterations=5000)
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            compare_trace(tr, orc, b, inp)
            assert_labels_equal(ctx.labels(), orc.labels(), f"batch {b}: labels")
            assert ctx.stats(b)["accuracy"] == orc.stats()["accuracy"]


def test_symmetric_generators_are_detected(gpu_available, c1_inputs):
    """The reference generator's graphs are symmetric (the default row
    choice can take the listed-row update on them)."""
    with L.Context(c1_inputs, L.make_config(c1_inputs)) as ctx:
        assert ctx.row_update() is True
    inp = small_batch_stream(5000, 3, batches=3)
    with L.Context(inp, L.make_config(inp)) as ctx:
        assert ctx.row_update() is True
