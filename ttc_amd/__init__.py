"""MI355X (gfx950) tensor-transposition generator and runtime.

Drop-in for HPAC/TTC's GPU path: ``ttc_amd.ttc`` (command line and
``generateTransposition``), ``ttc_amd.ttc_util`` (``TTCargs``,
``fuseIndices``) and ``ttc_amd.sql_util`` mirror the reference's modules;
``ttc_amd.runtime`` binds the C ABI of ``lib/libttc_hip.so``
(include/ttc_hip.h).
"""
from . import sql_util, ttc_util  # noqa: F401
from . import ttc  # noqa: F401

__all__ = ["ttc", "ttc_util", "sql_util"]
