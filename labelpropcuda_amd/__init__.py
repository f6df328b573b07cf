"""MI355X-native DynLP batch update (gfx950 HIP kernels behind a C-ABI).

The product is ``libdynlp.so`` (include/dynlp/dynlp_c.h) and the C++ drop-in
API of include/dynlp/dynLP.h.  This package is the Python plumbing over that
C-ABI used by the tests and the benchmark.
"""
from ._lib import DynLPError, device_count, load  # noqa: F401
from .dynlp import (Context, Inputs, Multi, TraceBuffers, batched_workload,  # noqa: F401
                    generate, generate_stream, make_config, run_dynlp)

__all__ = ["Context", "Inputs", "Multi", "TraceBuffers", "batched_workload", "generate",
           "generate_stream", "make_config", "run_dynlp", "DynLPError", "device_count", "load"]
