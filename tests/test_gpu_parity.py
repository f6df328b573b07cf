"""HIP path vs the CPU oracle, through the C-ABI.

Integer and graph state is compared exactly (CSR with its -1 holes, affected
set, G' in build order, CC partition and ids, iteration counts, frontier
sizes); FP32 labels are compared bit for bit, which implies the stated 1e-5
tolerance and the same 0.5-threshold class.
"""
import numpy as np
import pytest

import labelpropcuda_amd as L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
F32 = np.float32


def bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


def assert_labels_equal(gpu, ref, what):
    if not np.array_equal(bits(gpu), bits(ref)):
        d = np.abs(gpu.astype(np.float64) - ref.astype(np.float64))
        i = int(np.argmax(d))
        raise AssertionError(f"{what}: {int(np.sum(bits(gpu) != bits(ref)))} labels differ, "
                             f"max |dF| {d[i]:.3g} at vertex {i}")


def compare_trace(ctx_trace, orc, b, inp):
    snap = ctx_trace.snapshot()
    rp, ci, vv = orc.csr()
    assert np.array_equal(snap["rowPtr"], rp), f"batch {b}: rowPtr"
    assert np.array_equal(snap["colInd"], ci), f"batch {b}: colInd (holes)"
    assert np.array_equal(bits(snap["values"]), bits(vv)), f"batch {b}: values"
    assert np.array_equal(snap["affected"], orc.affected()), f"batch {b}: affected set"
    ss, sd, sw = orc.subgraph()
    assert snap["numSubEdges"] == ss.shape[0], f"batch {b}: |G'|"
    assert np.array_equal(snap["subSrc"], ss) and np.array_equal(snap["subDst"], sd)
    assert np.array_equal(bits(snap["subW"]), bits(sw))
    comp, nc = orc.component_ids()
    assert snap["numComponents"] == nc, f"batch {b}: numComponents"
    assert np.array_equal(snap["compId"], comp), f"batch {b}: component ids"
    assert_labels_equal(snap["initLabels"], orc.init_labels(), f"batch {b}: initial labels")
    fs = orc.frontier_sizes()
    assert snap["iterations"] == orc.stats()["iterations"], f"batch {b}: iterations"
    n = min(len(fs), len(snap["frontierSizes"]))
    assert np.array_equal(snap["frontierSizes"][:n], fs[:n]), f"batch {b}: frontier sizes"


def test_c1_every_step_matches_oracle(gpu_available, c1_inputs):
    inp = c1_inputs
    with L.Context(inp, L.make_config(inp)) as ctx:
        tr = ctx.enable_trace(maxFrontierSizes=1 << 17)
        orc = Oracle(inp)
        assert ctx.tau == orc.tau
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            compare_trace(tr, orc, b, inp)
            assert_labels_equal(ctx.labels(), orc.labels(), f"batch {b}: labels")
            st, ost = ctx.stats(b), orc.stats()
            assert st["iterations"] == ost["iterations"]
            assert st["updates"] == ost["updates"]
            assert st["frontier0"] == ost["frontier0"]
            assert st["numMarked"] == ost["numMarked"]
            assert st["accuracy"] == pytest.approx(ost["accuracy"], abs=0)


def test_c1_run_dynlp_matches_oracle(gpu_available, c1_inputs):
    inp = c1_inputs
    labels, times, acc = L.run_dynlp(inp, L.make_config(inp))
    orc = Oracle(inp)
    oacc = []
    for b in range(inp.numBatches):
        orc.run_batch(b)
        oacc.append(orc.stats()["accuracy"])
    assert_labels_equal(labels, orc.labels(), "runDynLP final labels")
    assert np.array_equal(acc, np.array(oacc, F32))
    assert np.all(times > 0)


@pytest.mark.timeout(1200)
def test_c2_matches_oracle_every_batch(gpu_available):
    """Config 2 (500,000 vertices, degree 7, 10 batches): labels bit-exact
    after every batch, plus the error against the FP64 twin."""
    inp = L.generate(500000, 7, 10, 42)
    orc = Oracle(inp)
    o64 = Oracle(inp, fp64=True)
    with L.Context(inp, L.make_config(inp)) as ctx:
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            o64.run_batch(b)
            gpu = ctx.labels()
            assert_labels_equal(gpu, orc.labels(), f"C2 batch {b}")
            st, ost = ctx.stats(b), orc.stats()
            assert (st["iterations"], st["updates"]) == (ost["iterations"], ost["updates"])
            assert st["accuracy"] == ost["accuracy"], f"C2 batch {b}: accuracy"
            err = np.abs(gpu.astype(np.float64) - o64.labels())
            print(f"C2 batch {b}: iterations {st['iterations']} updates {st['updates']} "
                  f"|F32-F64| max {err.max():.3e} mean {err.mean():.3e}")


@pytest.mark.parametrize("lanes,blocks,chunk", [(4, 256, 3), (16, 2048, 8), (32, 64, 1),
                                               (64, 1024, 16)])
def test_work_shapes_give_identical_results(gpu_available, c1_inputs, lanes, blocks, chunk):
    inp = c1_inputs
    cfg = L.make_config(inp, maxIterations=3000)
    orc = Oracle(inp, maxIterations=3000)
    orc.run_all()
    with L.Context(inp, cfg) as ctx:
        ctx.set_tuning(lanes, blocks, chunk)
        ctx.run_all()
        assert_labels_equal(ctx.labels(), orc.labels(), f"G={lanes} blocks={blocks}")


def run_engine_mix(inp, maxIter, **tuning):
    cfg = L.make_config(inp, maxIterations=maxIter)
    orc = Oracle(inp, maxIterations=maxIter)
    with L.Context(inp, cfg) as ctx:
        ctx.set_tuning(**tuning)
        ctx.set_profiling(True)
        tr = ctx.enable_trace(maxFrontierSizes=1 << 12)
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            snap = tr.snapshot()
            fs = orc.frontier_sizes()
            assert snap["iterations"] == orc.stats()["iterations"], f"batch {b}: iterations"
            n = min(len(fs), len(snap["frontierSizes"]))
            assert np.array_equal(snap["frontierSizes"][:n], fs[:n]), f"batch {b}: frontiers"
            assert_labels_equal(ctx.labels(), orc.labels(), f"batch {b}: labels")
            assert ctx.stats(b)["updates"] == orc.stats()["updates"]
        return ctx.prop_profile()


@pytest.mark.parametrize("threshold,ldsBytes,shape", [
    (-1, 0, 1),          # multi-block kernel only
    (1 << 30, 0, 1),     # LDS-resident kernel whenever the working set fits
    (64, 0, 1),          # small frontiers only: many engine switches
    (1 << 30, 4096, 1),  # 1024 LDS words: core rebuilds, retries, fallbacks
    (200, 16384, 1),
    (1 << 30, 0, 2), (1 << 30, 0, 3), (1 << 30, 0, 4),
    (1 << 30, 0, 5),     # one lane per row up to 8 entries: wave path for the rest
])
def test_engine_mix_gives_identical_results(gpu_available, c1_inputs, threshold, ldsBytes,
                                            shape):
    """The multi-block and the LDS-resident propagation kernels, in any mix
    and work-group shape, leave the oracle's labels, iteration counts and
    frontier sizes."""
    prof = run_engine_mix(c1_inputs, 3000, localThreshold=threshold, localLdsBytes=ldsBytes,
                          localShape=shape)
    if threshold < 0:
        assert prof["localLaunches"] == 0
    else:
        assert prof["localLaunches"] > 0


@pytest.mark.parametrize("shape", [1, 2, 5])
@pytest.mark.parametrize("owner", [1, -1])
def test_engine_mix_long_rows(gpu_available, shape, owner):
    """Average degree 40: rows longer than 32 take the wave-per-row paths of
    both kernels (and the general row sum beyond 64 entries), in owner mode
    and in queue iterations."""
    inp = L.generate(3000, 40, 6, 7)
    prof = run_engine_mix(inp, 2000, localThreshold=1 << 30, localShape=shape,
                          localOwnerAfter=owner)
    assert prof["localLaunches"] > 0
    assert (prof["localOwnerUpdates"] > 0) == (owner > 0)
    # every row is longer than 16 entries: the single-wave owner kernel hands
    # each core back to the work-group owner kernel
    assert prof["localWaveUpdates"] == 0


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("owner,threshold,ldsBytes", [(1, 1 << 30, 0), (-1, 1 << 30, 0),
                                                      (1, 64, 0), (1, 1 << 30, 8192),
                                                      (5, 1 << 30, 0), (0, 1 << 30, 0)])
def test_owner_mode_and_queue_iterations(gpu_available, c1_inputs, shape, owner, threshold,
                                         ldsBytes):
    """The LDS-resident kernel's owner mode (one thread per core vertex,
    frontier as flags) and its queue iterations leave the oracle's labels,
    iteration counts and frontier sizes for every work-group shape, with
    owner mode from the first iteration of a pass or after some queue
    iterations, frequent engine switches and core rebuilds in a small LDS
    budget."""
    prof = run_engine_mix(c1_inputs, 3000, localThreshold=threshold, localLdsBytes=ldsBytes,
                          localShape=shape, localOwnerAfter=owner)
    assert prof["localLaunches"] > 0
    if owner < 0:
        assert prof["localOwnerUpdates"] == 0
    elif owner == 1:
        assert prof["localOwnerUpdates"] > 0
        # small cores run in the single-wave owner kernel
        assert prof["localWaveUpdates"] > 0
    assert prof["localWaveUpdates"] <= prof["localOwnerUpdates"] <= prof["localUpdates"]


def hub_inputs(tmp_path, hubs=((0, 700), (1, 320))):
    """3,000 vertices at degree ~6 plus hubs (by default vertex 0 adjacent to
    1..700, vertex 1 to 2..320), read back through the graph text files, with
    the reference batch generator over it: once most vertices are active the
    hubs' rows exceed 256 entries."""
    rng = np.random.default_rng(3)
    n = 3000
    pairs = set()
    for u, v in rng.integers(0, n, (9000, 2)).tolist():
        if u != v:
            pairs.add((min(u, v), max(u, v)))
    for h, k in hubs:
        pairs |= {(h, v) for v in range(h + 1, h + 1 + k)}
    rows = [[] for _ in range(n)]
    for u, v in sorted(pairs):
        w = float(rng.uniform(0.01, 1.0))
        rows[u].append((v, w))
        rows[v].append((u, w))
    prefix = str(tmp_path / "hub_")
    rowPtr = np.cumsum([0] + [len(r) for r in rows])
    with open(prefix + "RowPtr.txt", "w") as f:
        f.write("\n".join(str(x) for x in rowPtr) + "\n")
    ordered = [sorted(r) for r in rows]
    with open(prefix + "ColInd.txt", "w") as f:
        f.write("\n".join(str(v) for r in ordered for v, _ in r) + "\n")
    with open(prefix + "Values.txt", "w") as f:
        f.write("\n".join(f"{w:.6f}" for r in ordered for _, w in r) + "\n")
    labels = rng.permutation(np.arange(n) % 2)
    with open(prefix + "TrueLabels.txt", "w") as f:
        f.write("\n".join(f"{i} {c}" for i, c in enumerate(labels)) + "\n")
    return L.dynlp.from_graph_files(prefix, 5, 42)


@pytest.mark.parametrize("lanes,threshold", [(4, -1), (8, -1), (16, -1), (64, -1), (8, 0)])
def test_rows_beyond_64_and_256_entries(gpu_available, tmp_path, lanes, threshold):
    """The multi-block kernel sums rows of more than 64 entries with the whole
    wave (and rows past 256 entries with two positions per slot) while the
    lane groups sum the short rows: oracle labels, iterations and frontiers
    for every group width, with and without the LDS-resident kernel."""
    inp = hub_inputs(tmp_path)
    deg = np.diff(np.asarray(inp.rowPtr))
    assert deg[0] >= 700 and deg[1] >= 320 and np.median(deg) < 16
    run_engine_mix(inp, 3000, groupLanes=lanes, localThreshold=threshold)
    dense = L.generate(2000, 80, 4, 5)  # most rows between 65 and 121 entries
    run_engine_mix(dense, 1000, groupLanes=lanes, localThreshold=threshold)


@pytest.mark.parametrize("waveRowMin,blockRowMin,lanes", [
    (64, 64, 8), (64, 300, 8), (64, 1024, 8), (16, 300, 4), (4, 128, 16), (64, -1, 8),
])
def test_row_tiers_lane_wave_block(gpu_available, tmp_path, waveRowMin, blockRowMin, lanes):
    """Rows are summed by a lane group, by a whole wave (longer than
    waveRowMin) or by a whole work-group (longer than blockRowMin): hubs of
    2,500, 1,200 and 400 entries and the degree-6 rest give the oracle's
    labels, iterations and frontiers under every split of the three tiers."""
    inp = hub_inputs(tmp_path, hubs=((0, 2500), (3, 1200), (7, 400)))
    deg = np.diff(np.asarray(inp.rowPtr))
    assert deg[0] >= 2500 and deg[3] >= 1200 and deg[7] >= 400
    run_engine_mix(inp, 3000, groupLanes=lanes, localThreshold=-1, waveRowMin=waveRowMin,
                   blockRowMin=blockRowMin)


def test_replay_is_deterministic(gpu_available, c1_inputs):
    with L.Context(c1_inputs, L.make_config(c1_inputs, maxIterations=5000)) as ctx:
        ctx.run_all()
        a = ctx.labels()
        s1 = [ctx.stats(b) for b in range(10)]
        ctx.run_all()
        b = ctx.labels()
        s2 = [ctx.stats(b) for b in range(10)]
    assert np.array_equal(bits(a), bits(b))
    assert [s["updates"] for s in s1] == [s["updates"] for s in s2]


@pytest.mark.parametrize("tau,maxIter", [(2.0, 100000), (0.0, 100000), (-1.0, 7), (0.3, 1)])
def test_tau_and_iteration_cap_edges(gpu_available, tau, maxIter):
    inp = L.generate(4000, 6, 6, 11)
    cfg = L.make_config(inp, tau=tau, maxIterations=maxIter)
    orc = Oracle(inp, tau=tau, maxIterations=maxIter)
    with L.Context(inp, cfg) as ctx:
        tr = ctx.enable_trace()
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            compare_trace(tr, orc, b, inp)
            assert_labels_equal(ctx.labels(), orc.labels(), f"tau={tau} batch {b}")


def handmade_inputs():
    """A 12-vertex stream exercising the batch-loop branches: GT-only batch
    without edges, a batch with deletions only, a batch whose inserted vertex
    is deleted in the same batch (it stays ACTIVE on the host and returns in
    the next batch's CSR), an empty batch, and a re-labelled vertex."""
    import labelpropcuda_amd.dynlp as D
    N = 12
    edges = [(0, 2, .9), (0, 3, .2), (1, 3, .8), (2, 3, .7), (2, 4, .6), (3, 5, .9), (4, 5, .4),
             (4, 6, .95), (5, 7, .85), (6, 7, .3), (6, 8, .75), (7, 9, .65), (8, 9, .55),
             (8, 10, .9), (9, 11, .8), (10, 11, .6), (1, 11, .7)]
    adj = [[] for _ in range(N)]
    for u, v, w in edges:
        adj[u].append((v, w))
        adj[v].append((u, w))
    rowPtr, colInd, values = [0], [], []
    for u in range(N):
        for v, w in sorted(adj[u]):
            colInd.append(v)
            values.append(w)
        rowPtr.append(len(colInd))
    batches = [
        dict(ins=[0, 1], gt=[(0, 0.0), (1, 1.0)], dels=[]),
        dict(ins=[2, 3, 4, 5], gt=[], dels=[]),
        dict(ins=[6, 7, 8], gt=[(8, 1.0)], dels=[2, 7]),   # 7 inserted+deleted (same batch)
        dict(ins=[], gt=[], dels=[4]),                      # deletions only
        dict(ins=[], gt=[], dels=[]),                       # empty
        dict(ins=[9, 10, 11], gt=[(10, 0.0)], dels=[5]),
    ]
    batchOf = {}
    for b, bt in enumerate(batches):
        for v in bt["ins"]:
            batchOf[v] = b
    per_batch_edges = [[] for _ in batches]
    for u in range(N):
        for v, w in sorted(adj[u]):
            if u < v:
                per_batch_edges[max(batchOf[u], batchOf[v])].append((u, v, w))
    cat = lambda key, f: np.array([x for bt in batches for x in f(bt)], dtype=key)  # noqa: E731
    off = lambda f: np.cumsum([0] + [len(f(bt)) for bt in batches]).astype(np.int64)  # noqa
    return D.Inputs(
        rowPtr=np.array(rowPtr, np.int32), colInd=np.array(colInd, np.int32),
        values=np.array(values, F32),
        trueLabels=np.array([0, 1, 0, 1, 0, 1, 0, 1, 1, 0, 0, 1], F32),
        insOff=off(lambda bt: bt["ins"]), ins=cat(np.int32, lambda bt: bt["ins"]),
        gtOff=off(lambda bt: bt["gt"]), gtIds=cat(np.int32, lambda bt: [g[0] for g in bt["gt"]]),
        gtLab=cat(F32, lambda bt: [g[1] for g in bt["gt"]]),
        delOff=off(lambda bt: bt["dels"]), dels=cat(np.int32, lambda bt: bt["dels"]),
        eOff=np.cumsum([0] + [len(e) for e in per_batch_edges]).astype(np.int64),
        eSrc=np.array([e[0] for pe in per_batch_edges for e in pe], np.int32),
        eDst=np.array([e[1] for pe in per_batch_edges for e in pe], np.int32),
        eW=np.array([e[2] for pe in per_batch_edges for e in pe], F32))


@pytest.mark.parametrize("tau", [0.5, 0.0])
def test_handmade_branches(gpu_available, tau):
    inp = handmade_inputs()
    cfg = L.make_config(inp, tau=tau)
    orc = Oracle(inp, tau=tau)
    with L.Context(inp, cfg) as ctx:
        tr = ctx.enable_trace()
        ctx.reset()
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            compare_trace(tr, orc, b, inp)
            assert_labels_equal(ctx.labels(), orc.labels(), f"handmade batch {b}")
            assert ctx.stats(b)["accuracy"] == orc.stats()["accuracy"]


# ---------------------------------------------------------------------------
# step API (reference launch* signatures) on device memory
# ---------------------------------------------------------------------------
from labelpropcuda_amd import _lib  # noqa: E402
import ctypes as C  # noqa: E402
import oracle.oracle as O  # noqa: E402

_alive = []  # device arrays stay allocated until the test ends


@pytest.fixture()
def torch_dev():
    """Device memory for the raw-pointer step API comes from the library's
    own plumbing entry points (dlp_device_malloc / dlp_copy_*)."""
    yield True
    for a in _alive:
        a.free()
    _alive.clear()


def dev(a):
    d = _lib.DeviceArray(a)
    _alive.append(d)
    return d


def ptr(t):
    return t.ptr


def host(t):
    return t.numpy()


def test_step_mark_deletions(gpu_available, torch_dev, lib, c1_inputs):
    inp = c1_inputs
    N = inp.numNodes
    o = Oracle(inp)
    for b in range(5):
        o.run_batch(b)
    rp, _, _ = o.csr()
    _, hs, _ = o.status()
    from refsim import ref_active_csr
    batchMap = np.full(N, -1, np.int32)
    for b in range(inp.numBatches):
        batchMap[inp.ins[inp.insOff[b]:inp.insOff[b + 1]]] = b
    rp, ci, vv = ref_active_csr(inp.rowPtr, inp.colInd, inp.values, hs, batchMap, 4)
    dels = inp.batch(4)["dels"]
    aff0 = np.zeros(N, np.int32)
    exp = O.mark_deletions(rp, ci, hs, aff0, dels, N)
    t_rp, t_ci, t_st, t_aff, t_del = dev(rp), dev(ci), dev(hs), dev(aff0), dev(dels)
    _lib.check(lib.dlp_mark_deletions(ptr(t_rp), ptr(t_ci), ptr(t_st), ptr(t_aff), ptr(t_del),
                                      len(dels), N, None))
    assert np.array_equal(host(t_ci), exp[0])
    assert np.array_equal(host(t_st), exp[1])
    assert np.array_equal(host(t_aff), exp[2])


def test_step_sparsify(gpu_available, torch_dev, lib, c1_inputs):
    inp = c1_inputs
    N = inp.numNodes
    bt = inp.batch(6)
    isNew = np.zeros(N, np.int32)
    isNew[bt["ins"]] = 1
    tau = np.float32(0.5016883)
    exp = O.sparsify(bt["eSrc"], bt["eDst"], bt["eW"], len(bt["ins"]), isNew, tau,
                     np.zeros(N, np.int32))
    m = len(bt["eSrc"])
    t = {k: dev(v) for k, v in dict(s=bt["eSrc"], d=bt["eDst"], w=bt["eW"], ins=bt["ins"],
                                   new=isNew, aff=np.zeros(N, np.int32),
                                   sc=np.zeros(2 * m, np.int32), sv=np.zeros(2 * m, F32),
                                   ss=np.zeros(2 * m, np.int32)).items()}
    nsn, nse = C.c_int(), C.c_int()
    _lib.check(lib.dlp_build_sparsified_subgraph(
        None, None, None, ptr(t["aff"]), None, ptr(t["s"]), ptr(t["d"]), ptr(t["w"]), m,
        ptr(t["ins"]), len(bt["ins"]), ptr(t["new"]), C.c_float(tau), None, ptr(t["sc"]),
        ptr(t["sv"]), ptr(t["ss"]), C.byref(nsn), C.byref(nse), None))
    assert nsn.value == len(bt["ins"]) and nse.value == len(exp[0])
    n = nse.value
    assert np.array_equal(host(t["ss"])[:n], exp[0])
    assert np.array_equal(host(t["sc"])[:n], exp[1])
    assert np.array_equal(bits(host(t["sv"])[:n]), bits(exp[2]))
    assert np.array_equal(host(t["aff"]), exp[3])


@pytest.mark.parametrize("seed", [0, 1])
def test_step_connected_components(gpu_available, torch_dev, lib, seed):
    rng = np.random.default_rng(seed)
    N = 200000
    inserted = rng.choice(N, 30000, replace=False).astype(np.int32)
    k = 40000
    a, b = rng.choice(inserted, k), rng.choice(inserted, k)
    src = np.concatenate([a, b]).astype(np.int32)
    dst = np.concatenate([b, a]).astype(np.int32)
    exp, nce = O.connected_components(src, dst, inserted, N)
    comp = dev(np.full(N, 5, np.int32))
    nc = C.c_int()
    _lib.check(lib.dlp_connected_components(ptr(dev(src)), ptr(dev(dst)), 2 * k,
                                            ptr(dev(inserted)), len(inserted), ptr(comp),
                                            C.byref(nc), N, None))
    assert nc.value == nce
    assert np.array_equal(host(comp), exp)


def test_step_connected_components_rejects_foreign_endpoint(gpu_available, torch_dev, lib):
    N = 100
    ins = dev(np.array([1, 2, 3], np.int32))
    comp = dev(np.zeros(N, np.int32))
    nc = C.c_int()
    rc = lib.dlp_connected_components(ptr(dev(np.array([1], np.int32))),
                                      ptr(dev(np.array([50], np.int32))), 1, ptr(ins), 3,
                                      ptr(comp), C.byref(nc), N, None)
    assert rc == _lib.DLP_ERR_INVALID


def test_step_label_initialization(gpu_available, torch_dev, lib, c1_inputs):
    inp = c1_inputs
    o = Oracle(inp)
    for b in range(7):
        o.run_batch(b)
    rp, ci, vv = o.csr()
    comp, nc = o.component_ids()
    _, _, gt = o.status()
    ins = inp.batch(6)["ins"]
    base = np.full(inp.numNodes, 0.25, F32)
    exp = O.label_initialization(rp, ci, vv, base, comp, nc, ins, gt)
    lab = dev(base)
    _lib.check(lib.dlp_label_initialization(ptr(dev(rp)), ptr(dev(ci)), ptr(dev(vv)), ptr(lab),
                                            None, ptr(dev(comp)), nc, ptr(dev(ins)), len(ins),
                                            ptr(dev(gt)), inp.numNodes, None))
    assert_labels_equal(host(lab), exp, "step label init")


@pytest.mark.parametrize("cap", [1, 2, 9, 100000])
def test_step_iterative_propagation(gpu_available, torch_dev, lib, c1_inputs, cap):
    inp = c1_inputs
    o = Oracle(inp)
    for b in range(6):
        o.run_batch(b)
    rp, ci, vv = o.csr()
    init, aff = o.init_labels(), o.affected()
    d, _, g = o.status()
    exp = O.propagate(rp, ci, vv, init, d, g, aff, 1e-4, cap)
    lab, buf, t_aff = dev(init), dev(np.zeros_like(init)), dev(aff)
    it = C.c_int()
    _lib.check(lib.dlp_iterative_propagation(ptr(dev(rp)), ptr(dev(ci)), ptr(dev(vv)), ptr(lab),
                                             ptr(buf), ptr(dev(d)), ptr(dev(g)), ptr(t_aff),
                                             inp.numNodes, C.c_float(1e-4), cap, C.byref(it),
                                             None))
    assert it.value == exp[3]
    assert_labels_equal(host(lab), exp[0], "step propagation labels")
    assert_labels_equal(host(buf), exp[1], "step propagation buffer")
    assert np.array_equal(host(t_aff), exp[2])
