"""haplo_amd -- MI355X-native likelihood engine for HaploADMIXTURE (PLINK BED, haplotype admixture).

Host-side mirror of the reference's driver / data-layout API on top of the C ABI of libhaplo_hip.so.
"""
from ._native import HipError, NativeLibraryError, header_symbols
from .admix import AdmixData2, _admixture_base, _filter_SKFR, run_admixture
from .engine import HipAdmixData, fit_batch
from .select import aic, select_k
from .sim import SimData, init_params, simulate

__all__ = [
    "AdmixData2", "HipAdmixData", "_filter_SKFR", "aic", "fit_batch", "select_k", "HipError", "NativeLibraryError", "SimData", "_admixture_base",
    "header_symbols", "init_params", "run_admixture", "simulate",
]
