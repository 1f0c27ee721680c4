"""erf_amd -- MI355X-native compressible dycore step of ERF.

The product is the HIP library erf_amd/liberfx.so (C ABI in include/erfx.h);
this package holds its ctypes binding (lib.py), the configuration mirror
(config.py) and the synthetic problem setups (problems.py).
"""
from . import config  # noqa: F401
