"""MI355X-native CongFD right-hand side and SSP-RK3 time advance.

The compute path is the HIP library ``libcongfd.so`` (sources in ``csrc/``,
C ABI in ``include/congfd.h``); this package is its Python binding plus the
host-side pieces a driver needs (``Info``-style configuration, scheme
selection, test cases).
"""
from .solver import (  # noqa: F401
    BC, Batch, FluxMethod, InterMethod, Recon, Solver, library_path, load_library, cfd_config,
    BATCH_GRAPH, BATCH_PERSISTENT, BATCH_PER_ZONE, PRECISION_FP32, PRECISION_FP64, PRECISION_FP64_FMA,
)
from . import cases  # noqa: F401
