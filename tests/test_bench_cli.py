"""bench.py's command line without a GPU: the one-GPU guard, and the
hardware-queue count, which must be taken from the command line before
argparse runs (the HIP runtime reads GPU_MAX_HW_QUEUES when it starts).
Each case runs bench.py in a child process, so this process's environment
is left alone."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_py(code, *args):
    return subprocess.run([sys.executable, "-c", code, *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_hw_queues_set_before_the_runtime_starts():
    code = ("import sys, os; sys.argv = ['bench.py'] + sys.argv[1:]; import bench; "
            "print(os.environ.get('GPU_MAX_HW_QUEUES'))")
    assert run_py(code).stdout.strip().splitlines()[-1] == "24"
    assert run_py(code, "--hw-queues", "6").stdout.strip().splitlines()[-1] == "6"
    assert run_py(code, "--hw-queues=12").stdout.strip().splitlines()[-1] == "12"
    env_kept = ("import sys, os; os.environ['GPU_MAX_HW_QUEUES'] = '4'; "
                "sys.argv = ['bench.py', '--hw-queues', '0']; import bench; "
                "print(os.environ.get('GPU_MAX_HW_QUEUES'))")
    assert run_py(env_kept).stdout.strip().splitlines()[-1] == "4"
