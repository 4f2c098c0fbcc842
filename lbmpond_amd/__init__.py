"""lbmpond_amd -- MI355X-native particles-on-demand (PonD) step.

The compute lives in ``libpond.so`` (HIP kernels for gfx950 behind the C ABI of
``include/pond.h``); this package is the Python host mirror of that ABI.
"""
from .pond import (LATTICES, INTERPS, PRECISIONS, QN, LBMParams, Pond, PondError, calcStep, diagnostics, drop,
                   init, lattice_tables, lib, read_arr)

__all__ = ["LATTICES", "INTERPS", "PRECISIONS", "QN", "LBMParams", "Pond", "PondError", "calcStep",
           "diagnostics", "drop", "init", "lattice_tables", "lib", "read_arr"]
