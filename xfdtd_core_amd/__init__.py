"""MI355X-native FDTD time-stepping hot path with the XFDTD_CORE scenario API.

The product is the C-ABI library ``libxfdtd_core_amd.so`` (HIP kernels for
gfx950 plus the C++20 host that performs the reference's FP64 precompute).
This package locates and loads it (``native``), describes scenarios with the
reference's class names (``spec``) and builds the measured workloads
(``scenarios``).
"""

from . import spec  # noqa: F401
