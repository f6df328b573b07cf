"""The C-ABI boundary without a GPU: the library loads, exports every entry
point the headers declare, validates inputs before touching a device, and the
C++ drop-in header compiles against it."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dynlp", "dynlp_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dlp_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.SIGNATURES), "ctypes binding lags the header"


def test_version_and_device_count(lib):
    assert b"gfx950" in lib.dlp_version()
    assert L.device_count() >= 0


def test_context_rejects_invalid_inputs_before_device_use(c1_inputs):
    import copy
    bad = copy.deepcopy(c1_inputs)
    bad.ins = bad.ins.copy()
    bad.ins[bad.insOff[2]] = bad.ins[bad.insOff[2] + 1]  # duplicate inserted id
    cfg = L.make_config(bad)
    with pytest.raises(L.DynLPError) as e:
        L.Context(bad, cfg)
    assert e.value.code == _lib.DLP_ERR_INVALID and "twice" in str(e.value)

    bad = copy.deepcopy(c1_inputs)
    bad.eDst = bad.eDst.copy()
    bad.eDst[5] = bad.numNodes + 3
    with pytest.raises(L.DynLPError) as e:
        L.Context(bad, L.make_config(bad))
    assert "out of range" in str(e.value)

    bad = copy.deepcopy(c1_inputs)
    bad.eW = bad.eW.copy()
    bad.eW[7] = np.nan
    with pytest.raises(L.DynLPError) as e:
        L.Context(bad, L.make_config(bad))
    assert "finite" in str(e.value)

    cfg = L.make_config(c1_inputs, numBatches=99)
    with pytest.raises(L.DynLPError) as e:
        L.Context(c1_inputs, cfg)
    assert e.value.code == _lib.DLP_ERR_INVALID


def test_context_without_gpu_reports_no_device(c1_inputs):
    if L.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(L.DynLPError) as e:
        L.Context(c1_inputs, L.make_config(c1_inputs))
    assert e.value.code == _lib.DLP_ERR_NODEVICE


def test_step_api_early_returns_need_no_device(lib):
    nsn, nse = C.c_int(-1), C.c_int(-1)
    rc = lib.dlp_build_sparsified_subgraph(None, None, None, None, None, None, None, None, 0,
                                           None, 17, None, C.c_float(0.5), None, None, None,
                                           None, C.byref(nsn), C.byref(nse), None)
    assert rc == 0 and nsn.value == 17 and nse.value == 0
    nc = C.c_int(-1)
    rc = lib.dlp_connected_components(None, None, 5, None, 0, None, C.byref(nc), 10, None)
    assert rc == 0 and nc.value == 0
    assert lib.dlp_mark_deletions(None, None, None, None, None, 0, 10, None) == 0
    assert lib.dlp_label_initialization(None, None, None, None, None, None, 0, None, 0, None, 10,
                                        None) == 0
    rc = lib.dlp_iterative_propagation(None, None, None, None, None, None, None, None, 10,
                                       C.c_float(1e-4), 5, None, None)
    assert rc == _lib.DLP_ERR_INVALID


def test_tuning_validation_needs_context(lib):
    t = _lib.Tuning(7, 0, 0)
    assert lib.dlp_context_set_tuning(None, C.byref(t)) == _lib.DLP_ERR_INVALID


def test_cxx_dropin_header_compiles_and_links(tmp_path):
    src = tmp_path / "dropin.cpp"
    src.write_text(r'''
#include "dynlp/dynLP.h"
#include <cstdio>
int main(int argc, char** argv) {
  DynLPConfig cfg;
  if (!parseArguments(argc, argv, cfg)) return 2;
  CSRGraph g; std::vector<float> labels; std::vector<BatchData> batches;
  if (!generateSparseGraph(cfg.totalNodes, cfg.avgDegree, cfg.seed, g, labels, "")) return 3;
  if (!generateBatches(g, labels, cfg, batches)) return 4;
  FlatBatches f = flattenBatches(batches);
  dlp_batch_view v = f.view();
  std::printf("%d %d %d %lld\n", g.numNodes, g.numEdges, v.numBatches,
              (long long)v.insertedOffsets[v.numBatches]);
  return 0;
}
''')
    exe = tmp_path / "dropin"
    lib_dir = os.path.join(ROOT, "labelpropcuda_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", lib_dir, "-ldynlp", f"-Wl,-rpath,{lib_dir}"], check=True)
    out = subprocess.run([str(exe), "--totalNodes", "1200", "--numBatches", "4", "--quiet"],
                         check=True, capture_output=True, text=True).stdout
    assert out.strip().splitlines()[-1] == "1200 6000 4 1200"


def test_cli_help_and_flag_errors():
    assert os.path.exists(_lib.CLI_PATH)
    r = subprocess.run([_lib.CLI_PATH, "--help"], capture_output=True, text=True)
    assert r.returncode == 1 and "--totalNodes" in r.stdout
    r = subprocess.run([_lib.CLI_PATH, "--delta", "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "--delta must be positive" in r.stdout
    r = subprocess.run([_lib.CLI_PATH, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown argument" in r.stdout


def header_structs():
    """typedef struct dlp_x {...} dlp_x; names in the C header."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return re.findall(r"typedef struct (dlp_[a-z_]+) \{", text)


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every ctypes mirror of a header struct has the C field names, offsets
    and size (a field added or dropped on one side fails here)."""
    pairs = {"dlp_config": _lib.Config, "dlp_graph_view": _lib.GraphView,
             "dlp_batch_view": _lib.BatchView, "dlp_batch_stats": _lib.BatchStats,
             "dlp_tuning": _lib.Tuning, "dlp_trace": _lib.Trace,
             "dlp_multi_tuning": _lib.MultiTuning, "dlp_serve_request": _lib.ServeRequest,
             "dlp_serve_config": _lib.ServeConfig, "dlp_prop_profile": _lib.PropProfile,
             "dlp_irlp_tuning": _lib.IrlpTuning, "dlp_irlp_result": _lib.IrlpResult,
             "dlp_shape_key": _lib.ShapeKey, "dlp_pool_stats": _lib.PoolStats}
    assert set(header_structs()) == set(pairs), "a header struct has no ctypes mirror"
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "dynlp/dynlp_c.h"',
             'int main(void) {']
    for cname, cls in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    got = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], check=True, capture_output=True,
                              text=True).stdout.splitlines())
    for cname, cls in pairs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname} {f}"]) == getattr(cls, f).offset, f"{cname}.{f}"
