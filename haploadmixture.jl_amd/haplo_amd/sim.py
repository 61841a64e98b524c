"""Seeded synthetic PLINK BED data (the haplotype admixture model with Balding-Nichols drift).

``simulate`` wraps ``hk_sim_bed`` (csrc/hk_sim.c): Q_i ~ Dir(alpha_q), ancestral haplotype frequencies
pi_j ~ Dir(1,1,1,1), population frequencies P_kj ~ Dir(pi_j (1-F)/F), two haplotype copies per
individual drawn from their ancestry mix, genotypes = allele-2 counts, i.i.d. missingness per SNP.
``init_params`` wraps ``hk_sim_init``: the random start of ``AdmixData2`` (src/structs.jl:155-199).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native


@dataclass
class SimData:
    I: int
    J: int
    K: int
    bed: np.ndarray          # uint8 payload, 2J * ceil(I/4) bytes, SNP-major
    q_true: np.ndarray | None  # K x I (Fortran order)
    p_true: np.ndarray | None  # K x 4J

    @property
    def bytes_per_snp(self) -> int:
        return (self.I + 3) // 4


def simulate(I: int, J: int, K: int, fst: float = 0.1, miss: float = 0.01, alpha_q: float = 0.5,
             seed: int = 1, truth: bool = True) -> SimData:
    bps = (I + 3) // 4
    bed = np.empty(2 * J * bps, dtype=np.uint8)
    qt = np.zeros((K, I), order="F") if truth else None
    pt = np.zeros((K, 4 * J), order="F") if truth else None
    rc = _native.sim().hk_sim_bed(I, J, K, fst, miss, alpha_q, seed, _native.ptr(bed), _native.ptr(qt),
                                  _native.ptr(pt))
    if rc != 0:
        raise ValueError(f"hk_sim_bed rejected I={I} J={J} K={K} fst={fst} miss={miss}")
    return SimData(I, J, K, bed, qt, pt)


def simulate_world(I: int, J: int, K: int, fst: float, miss: float = 0.01, alpha_q: float = 0.5,
                   world_seed: int = 2024, data_seed: int = 1, truth: bool = True) -> SimData:
    """Dataset from a fixed population structure shared by all datasets of one world_seed (hk_sim_bed_world):
    block j is block j of a fixed panel; population k has the same haplotype frequencies up to the
    dataset's Fst scaling (0 < fst <= 0.15)."""
    bps = (I + 3) // 4
    bed = np.empty(2 * J * bps, dtype=np.uint8)
    qt = np.zeros((K, I), order="F") if truth else None
    pt = np.zeros((K, 4 * J), order="F") if truth else None
    rc = _native.sim().hk_sim_bed_world(I, J, K, fst, miss, alpha_q, world_seed, data_seed, _native.ptr(bed),
                                        _native.ptr(qt), _native.ptr(pt))
    if rc != 0:
        raise ValueError(f"hk_sim_bed_world rejected I={I} J={J} K={K} fst={fst}")
    return SimData(I, J, K, bed, qt, pt)


def init_params(I: int, J: int, K: int, seed: int = 7):
    """Random normalised (q, p) as in AdmixData2's constructor (src/structs.jl:155,178,189-199)."""
    q = np.zeros((K, I), order="F")
    p = np.zeros((K, 4 * J), order="F")
    rc = _native.sim().hk_sim_init(I, J, K, seed, _native.ptr(q), _native.ptr(p))
    if rc != 0:
        raise ValueError("hk_sim_init rejected its arguments")
    return q, p
