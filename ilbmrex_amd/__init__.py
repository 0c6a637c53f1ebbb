"""MI355X-native per-timestep IB-LBM path of iLBMReX.

The compute path is the HIP library ``_lib/libilbm_hip.so`` (C ABI in
``include/ilbm.h``); this package holds its case description
(``config``), the ctypes binding (``native``) and the reference-style
driver helpers (``driver``).
"""
from . import config  # noqa: F401

__all__ = ["config", "native"]
