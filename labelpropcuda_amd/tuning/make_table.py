"""Writes the committed tuning table of this GPU (<gfx arch>.tune beside this
script), which libdynlp loads when the first context or runner is created.

Autotunes, on the GPU it runs on:
  * the 500,000-vertex, degree-7, 10-batch stream (one context, all batches
    per trial);
  * the batched workload (256 graphs, 2,000-200,000 vertices, degree 3/5/7,
    10 batches each; work-group shape x hand-back frontier, the first
    --multi-trial batch indices per trial).
Graph shapes not in the table start from the degree-histogram heuristic.

    python labelpropcuda_amd/tuning/make_table.py [--out PATH]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main() -> int:
    import labelpropcuda_amd as L
    from labelpropcuda_amd import dynlp
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--multi-trial", type=int, default=3)
    args = ap.parse_args()
    t0 = time.perf_counter()

    def say(msg):
        print(f"[make_table {time.perf_counter() - t0:6.1f} s] {msg}", flush=True)

    dynlp.tuning_table_clear()  # start empty: the table holds timed choices only
    inp = L.generate(500000, 7, 10, 42)
    with L.Context(inp, L.make_config(inp)) as ctx:
        t, ms = ctx.autotune(0)
        say(f"500,000 vertices, degree 7: {ms:.1f} ms per stream with {t}")
    graphs = L.batched_workload(256, 10, 42)
    with L.Multi(graphs) as m:
        t, ms = m.autotune(args.multi_trial)
        say(f"256 graphs: {ms:.1f} ms for {args.multi_trial} batch indices with {t}")
    arch = None
    import subprocess
    try:
        arch = subprocess.run(["rocminfo"], capture_output=True, text=True).stdout
        arch = next(w for w in arch.split() if w.startswith("gfx"))
    except Exception:
        arch = "gfx950"
    out = args.out or os.path.join(HERE, f"{arch}.tune")
    n = dynlp.tuning_table_save(out)
    say(f"wrote {n} entries to {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
