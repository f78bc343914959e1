"""MI355X-native PDLP engine: a drop-in for the reference's pdlp::PdlpSolver.

The engine is native (libpdlp_amd.so: gfx950 HIP kernels behind the C ABI of
include/pdlpx.h, plus the C++ API of include/pdlp/*.h).  This package holds
the Python bindings used by tests and benchmarks, the instance generators and
the build script.
"""
from .lp import Lp, box_lp, zipf_lp, block_lp, batch_sizes, batch_family, CONFIGS, BASE_SEED  # noqa: F401
from .solver import Engine, PdlpSolver, PdlpResult, default_options  # noqa: F401

__all__ = ["Lp", "box_lp", "zipf_lp", "block_lp", "batch_sizes", "batch_family", "CONFIGS", "BASE_SEED", "Engine",
           "PdlpSolver", "PdlpResult", "default_options"]
