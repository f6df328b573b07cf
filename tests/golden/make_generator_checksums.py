"""Writes tests/golden/generator_checksums.json: sha256 over every array the
input generator produces for a few (N, degree, batches, seed) settings.  Run
after an intentional generator change; the test fails on any other drift
(e.g. a different libstdc++ distribution algorithm)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import labelpropcuda_amd as L  # noqa: E402
from test_generator import digest  # noqa: E402

SETTINGS = [(1000, 5, 5, 42), (5000, 5, 10, 42), (2000, 7, 3, 7)]

if __name__ == "__main__":
    out = {f"{n},{d},{b},{s}": digest(L.generate(n, d, b, s)) for n, d, b, s in SETTINGS}
    with open(os.path.join(HERE, "generator_checksums.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)
