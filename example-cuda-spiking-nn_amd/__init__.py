"""MI355X-native per-step engine for the conductance-based EIF spiking network of
paulazoo/example-cuda-spiking-nn.

The product is the C++/HIP library ``lib/libburyn_amd.so`` with the C ABI of
``include/buryn_amd.h``; this package is its Python view (ctypes).  There is no Python or CPU
fallback: every call goes to the HIP library, and importing the binding raises if the library
was not built.
"""
from .engine import (  # noqa: F401
    BURYN_FLAGS,
    BatchSimulation,
    MultiSimulation,
    batch_workload_instance,
    LIB_PATH,
    BnError,
    Simulation,
    Window,
    build_library,
    config_dump,
    generate_weight_matrix,
    set_connectivity_cache,
    lib,
    measure_read_bandwidth,
    measure_read_bandwidth_one_launch,
    poisson_table,
)
