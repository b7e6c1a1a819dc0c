"""MI355X (gfx950) rebuild of Ocean's estimation-based SpGEMM.

The product is ``libocean.so`` (hand-written HIP kernels and the C++
orchestrator behind the C ABI of ``include/ocean_capi.h``) and the
``spgemm`` CLI.  ``research_amd.ocean`` is a ctypes mirror of that ABI for
Python callers and tests.
"""
from .ocean import (  # noqa: F401
    Batch,
    CSR,
    Config,
    DeviceCSR,
    DeviceBuffer,
    Mlp,
    OceanError,
    SpGEMM,
    lib,
    library_path,
)
