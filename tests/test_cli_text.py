"""The drop-in CLI's interface text, byte for byte: --help, every flag
error, the unknown-flag message, the configuration banner and the two
result files' formats (labels "%d %.6f", results.csv "batch,time_ms,accuracy"
with 2 / 6 decimals).  The expected strings are the reference driver's
output (utils.cu:24-200, main.cu:222-295).  No GPU: every case fails or
ends before device work."""
import os
import subprocess

import pytest

from labelpropcuda_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HELP = (
    "Usage: dynLP [options]\n"
    "\n"
    "Options:\n"
    "  --totalNodes N    Total vertices across all batches (default: 50000)\n"
    "  --numBatches N    Number of dynamic batches (default: 10)\n"
    "  --avgDegree  N    Average degree for graph gen (default: 5)\n"
    "  --tau        F    Sparsification threshold (default: auto)\n"
    "  --delta      F    Convergence threshold (default: 0.0001)\n"
    "  --maxIter    N    Max propagation iterations (default: 100000)\n"
    "  --gtFraction F    Ground-truth fraction (default: 0.01)\n"
    "  --seed       N    RNG seed (default: 42)\n"
    "  --output     DIR  Output directory (default: output/)\n"
    "  --validate        Run IrLP baseline comparison\n"
    "  --noValidate      Skip IrLP baseline\n"
    "  --verbose         Print per-batch details\n"
    "  --quiet           Suppress per-batch details\n"
    "  --help, -h        Show this help message\n")


def cli(*args):
    return subprocess.run([_lib.CLI_PATH, *args], capture_output=True, text=True)


@pytest.mark.parametrize("flag", ["--help", "-h"])
def test_help_text(flag):
    r = cli("--seed", "3", flag, "--bogus")
    assert r.returncode == 1
    assert r.stdout == HELP


@pytest.mark.parametrize("args,message", [
    (["--totalNodes", "0"], "Error: --totalNodes must be positive.\n"),
    (["--numBatches", "-2"], "Error: --numBatches must be positive.\n"),
    (["--avgDegree", "x"], "Error: --avgDegree must be positive.\n"),
    (["--delta", "0"], "Error: --delta must be positive.\n"),
    (["--gtFraction", "1"], "Error: --gtFraction must be in (0, 1).\n"),
    (["--gtFraction", "0"], "Error: --gtFraction must be in (0, 1).\n"),
    (["--bogus"], "Error: Unknown argument: --bogus\nRun with --help for usage information.\n"),
    (["--seed", "1", "--tau"],
     "Error: Unknown argument: --tau\nRun with --help for usage information.\n"),
    (["--gtFraction", "0.95"],
     "Error: Batch fractions (GT + unlabeled + deleted) exceed 1.0.\n"),
])
def test_flag_errors(args, message):
    r = cli(*args)
    assert r.returncode == 1
    assert r.stdout == message


def test_result_file_formats(tmp_path):
    src = tmp_path / "w.cpp"
    src.write_text(r'''
#include "dynlp/dynLP.h"
int main(int, char** argv) {
  std::string dir = argv[1];
  std::vector<float> lab = {0.0f, 0.25f, 0.1234567f, 1.0f, 0.5f}, t = {1.5f, 20.256f, 0.004f},
                     a = {1.0f, 0.5f, 0.3333333f};
  DynLPConfig cfg;
  printConfig(cfg);
  cfg.tau = 0.25f;
  cfg.validate = false;
  printConfig(cfg);
  return (writeLabelsToFile(dir + "/sub/l.txt", lab.data(), 5) &&
          writeResultsCSV(dir + "/r.csv", t.data(), a.data(), 3)) ? 0 : 2;
}
''')
    exe = tmp_path / "w"
    lib_dir = os.path.join(ROOT, "labelpropcuda_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", lib_dir, "-ldynlp", f"-Wl,-rpath,{lib_dir}"], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "sub" / "l.txt").read_text() == (
        "0 0.000000\n1 0.250000\n2 0.123457\n3 1.000000\n4 0.500000\n")
    assert (tmp_path / "r.csv").read_text() == (
        "batch,time_ms,accuracy\n0,1.50,1.000000\n1,20.26,0.500000\n2,0.00,0.333333\n")
    rule = "=" * 60
    banner = lambda tau, val: (  # noqa: E731
        f"\n{rule}\n DynLP Configuration\n{rule}\n"
        "  Total nodes:        50000\n"
        "  Number of batches:  10\n"
        "  Average degree:     5\n"
        f"  Tau (sparsify):     {tau}\n"
        "  Delta (converge):   0.0001\n"
        "  Max iterations:     100000\n"
        "  GT fraction:        0.01\n"
        "  Unlabeled fraction: 0.9\n"
        "  Deleted fraction:   0.09\n"
        "  Seed:               42\n"
        "  Output directory:   output/\n"
        f"  Validate (IrLP):    {val}\n"
        "  Verbose:            yes\n"
        f"{rule}\n\n")
    assert r.stdout == banner("auto (avg edge weight)", "yes") + banner("0.250000", "no")
