"""mulskips_amd -- MI355X-native KMC event engine for MulSKIPS.

The engine itself is HIP C++ behind the C ABI of include/mskg.h
(libmskg.so).  This package holds the host-side mirror of the reference's
driver: start.dat parsing (deck), the engine wrapper (engine) and the
observables used for statistical comparison (observables).
"""
from .engine import Engine, EngineError, find_nn, find_sn  # noqa: F401
