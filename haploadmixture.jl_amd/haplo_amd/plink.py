"""Minimal PLINK 1 binary fileset I/O (.bed/.bim/.fam) for the engine.

The reference reads BED files through SnpArrays (``SnpArray(filename)``, src/driver.jl:148) and hands
the raw payload (``g_la.s.data``) to its loops and GPU plugin.  Here ``read_bed`` memory-maps the file,
checks the magic bytes (0x6c 0x1b, SNP-major mode 0x01) and returns the payload view plus its
bytes-per-SNP; ``write_bed`` writes a payload with the magic and optional placeholder .bim/.fam files.
"""
from __future__ import annotations

import os

import numpy as np

MAGIC = bytes([0x6C, 0x1B, 0x01])


def count_lines(path: str) -> int:
    with open(path, "rb") as f:
        return sum(1 for _ in f)


def read_bed(path: str, n_individuals: int | None = None, n_snps: int | None = None):
    """Return (payload uint8 memmap, n_individuals, n_snps, bytes_per_snp).

    Sizes come from the sibling .fam/.bim files when not given."""
    if not path.endswith(".bed"):
        raise ValueError("filename should end with .bed")  # src/driver.jl:58
    stem = path[:-4]
    if n_individuals is None:
        n_individuals = count_lines(stem + ".fam")
    if n_snps is None:
        n_snps = count_lines(stem + ".bim") if os.path.exists(stem + ".bim") else None
    raw = np.memmap(path, dtype=np.uint8, mode="r")
    if bytes(raw[:3]) != MAGIC:
        raise ValueError(f"{path}: not a SNP-major PLINK 1 .bed file")
    bps = (n_individuals + 3) // 4
    payload = raw[3:]
    if n_snps is None:
        n_snps = payload.size // bps
    if payload.size < n_snps * bps:
        raise ValueError(f"{path}: payload has {payload.size} bytes, expected {n_snps * bps}")
    return payload[: n_snps * bps], n_individuals, n_snps, bps


def write_bed(path: str, payload: np.ndarray, n_individuals: int, n_snps: int, write_sidecars: bool = True):
    if not path.endswith(".bed"):
        raise ValueError("filename should end with .bed")
    bps = (n_individuals + 3) // 4
    if payload.size != n_snps * bps:
        raise ValueError("payload size does not match n_snps * ceil(n_individuals / 4)")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.ascontiguousarray(payload, dtype=np.uint8).tobytes())
    if write_sidecars:
        stem = path[:-4]
        with open(stem + ".fam", "w") as f:
            for i in range(n_individuals):
                f.write(f"fam{i} ind{i} 0 0 0 -9\n")
        with open(stem + ".bim", "w") as f:
            for s in range(n_snps):
                f.write(f"1\tsnp{s}\t0\t{s + 1}\tA\tG\n")
