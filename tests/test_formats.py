"""The on-disk text formats: the generator's CSR dump read back (structure
exact, weights as written with 6 decimals), batches regenerated over the
read graph, and the labels / results readers of the C++ API."""
import os
import subprocess

import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import dynlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_csr_dump_round_trip(tmp_path):
    inp = L.generate(3000, 5, 6, 42)
    prefix = str(tmp_path / "graph" / "csr")
    dynlp.write_graph(inp, prefix)
    text = (tmp_path / "graph" / "csrValues.txt").read_text().split()
    assert len(text) == inp.numEdges and all(len(t.split(".")[1]) == 6 for t in text[:50])
    back = dynlp.from_graph_files(prefix, inp.numBatches, 42)
    assert np.array_equal(back.rowPtr, inp.rowPtr)
    assert np.array_equal(back.colInd, inp.colInd)
    assert np.array_equal(back.trueLabels, inp.trueLabels)
    assert np.max(np.abs(back.values - inp.values)) <= 6e-7
    # the batch generator sees the same structure, labels and seed: same
    # vertex lists and edges, weights as stored
    for k in ("insOff", "ins", "gtOff", "gtIds", "gtLab", "delOff", "dels", "eOff", "eSrc",
              "eDst"):
        assert np.array_equal(getattr(back, k), getattr(inp, k)), k
    assert np.max(np.abs(back.eW - inp.eW)) <= 6e-7


def test_csr_reader_rejects_broken_files(tmp_path):
    inp = L.generate(500, 3, 3, 1)
    prefix = str(tmp_path / "csr")
    dynlp.write_graph(inp, prefix)
    col = (tmp_path / "csrColInd.txt")
    lines = col.read_text().splitlines()
    col.write_text("\n".join(lines[:-1]) + "\n")  # one entry short of RowPtr's span
    with pytest.raises(L.DynLPError):
        dynlp.from_graph_files(prefix, 3)
    lines[0] = "500"  # out of range
    col.write_text("\n".join(lines) + "\n")
    with pytest.raises(L.DynLPError):
        dynlp.from_graph_files(prefix, 3)
    with pytest.raises(L.DynLPError):
        dynlp.from_graph_files(str(tmp_path / "missing"), 3)


def test_labels_and_results_readers(tmp_path):
    src = tmp_path / "rw.cpp"
    src.write_text(r'''
#include "dynlp/dynLP.h"
#include <cstdio>
int main(int, char** argv) {
  std::string dir = argv[1];
  std::vector<float> lab = {0.0f, 0.25f, 0.1234567f, 1.0f}, t = {1.5f, 20.25f}, a = {1.0f, 0.5f};
  if (!writeLabelsToFile(dir + "/l.txt", lab.data(), 4)) return 2;
  if (!writeResultsCSV(dir + "/r.csv", t.data(), a.data(), 2)) return 3;
  std::vector<float> lab2, t2, a2;
  if (!readLabelsFile(dir + "/l.txt", lab2) || !readResultsCSV(dir + "/r.csv", t2, a2)) return 4;
  std::printf("%zu %.6f %.6f %zu %.2f %.6f\n", lab2.size(), lab2[1], lab2[2], t2.size(), t2[1],
              a2[1]);
  std::vector<float> bad;
  std::FILE *f = std::fopen((dir + "/b.txt").c_str(), "w");
  std::fprintf(f, "0 0.5\n2 0.5\n");
  std::fclose(f);
  return readLabelsFile(dir + "/b.txt", bad) ? 5 : 0;
}
''')
    exe = tmp_path / "rw"
    lib_dir = os.path.join(ROOT, "labelpropcuda_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", lib_dir, "-ldynlp", f"-Wl,-rpath,{lib_dir}"], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.splitlines()[0] == "4 0.250000 0.123457 2 20.25 0.500000"
    assert "vertex id 2 on line 2" in r.stdout
