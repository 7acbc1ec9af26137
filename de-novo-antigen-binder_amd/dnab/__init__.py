"""dnab: MI355X-native batched all-atom binder scoring (Python handles).

The product is libdnab.so (HIP kernels for gfx950 behind the C ABI of
include/dnab.h). This package binds it with ctypes for tests, benchmarks and
scripting; it raises on import when the library is not built.
"""
from . import _native
from ._native import DNAB_STATUS_NAN as STATUS_NAN
from ._native import DNAB_STATUS_SATURATED as STATUS_SATURATED
from ._native import DnabError, device_count, visible_gpu
from .api import (AtomEnergyBreakdown, Dataset, DeviceAntigen, DevicePopulation, GpuContext,
                  Mlp, SpatialHashGrid, exact_topk, langevin_params, map_energy, population_features, select_smallest,
                  precision_study, run_allatom)
from .structure import (AMINO_ONE_LETTER, AtomCloud, AtomProtein, PopulationHost, ResidueCloud,
                        aa_from_sequence, parse_peptide, population_from_proteins,
                        protein_from_ca_trace)

__all__ = [
    "AMINO_ONE_LETTER", "AtomCloud", "AtomEnergyBreakdown", "AtomProtein", "Dataset",
    "DeviceAntigen", "Mlp", "exact_topk", "map_energy", "population_features", "select_smallest", "precision_study",
    "DevicePopulation", "DnabError", "GpuContext", "PopulationHost", "ResidueCloud",
    "STATUS_NAN", "STATUS_SATURATED",
    "SpatialHashGrid", "aa_from_sequence", "device_count", "langevin_params", "parse_peptide",
    "population_from_proteins", "protein_from_ca_trace", "run_allatom", "visible_gpu", "_native",
]
