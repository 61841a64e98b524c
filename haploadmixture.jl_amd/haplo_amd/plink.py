"""PLINK 1 binary fileset I/O (.bed/.bim/.fam) for the engine.

The reference reads BED files through SnpArrays (``SnpArray(filename)``, src/driver.jl:148) and hands
the raw payload (``g_la.s.data``) to its loops and GPU plugin; ``SnpArrays.filter`` writes subsets
(src/driver.jl:103,122).  Here:

* ``read_bed`` memory-maps the file, checks the magic bytes (0x6c 0x1b, SNP-major mode 0x01) and returns
  the payload view plus its bytes-per-SNP; ``read_fam`` / ``read_bim`` parse the sidecars;
  ``PlinkFileset`` bundles the three, with ``payload(I, J)`` giving the first I individuals and first 2J
  SNPs in the layout the engine takes (``bytes_per_snp`` of the file; the engine reads the first I
  individuals of each SNP column);
* ``write_bed`` writes a payload with the magic and optional placeholder .bim/.fam files;
* ``filter`` writes the fileset restricted to chosen individuals and SNPs (``SnpArrays.filter``);
* ``write_q`` / ``read_q`` / ``write_p`` / ``read_p``: fitted Q (one line per individual, K columns) and
  haplotype frequencies P (one line per (block, haplotype), K columns), as whitespace-separated text.
"""
from __future__ import annotations

import os

import numpy as np

MAGIC = bytes([0x6C, 0x1B, 0x01])

FAM_DTYPE = [("fid", object), ("iid", object), ("father", object), ("mother", object), ("sex", object),
             ("phenotype", object)]
BIM_DTYPE = [("chrom", object), ("snpid", object), ("cm", np.float64), ("pos", np.int64), ("allele1", object),
             ("allele2", object)]


def count_lines(path: str) -> int:
    with open(path, "rb") as f:
        return sum(1 for _ in f)


def _stem(path: str) -> str:
    if path.endswith(".bed"):
        return path[:-4]
    return path


def read_fam(path: str) -> np.ndarray:
    """.fam records (FID, IID, father, mother, sex, phenotype) as a structured array, one per individual."""
    rows = []
    with open(path) as f:
        for n, line in enumerate(f, 1):
            t = line.split()
            if not t:
                continue
            if len(t) < 6:
                raise ValueError(f"{path}:{n}: a .fam line has 6 fields, found {len(t)}")
            rows.append(tuple(t[:6]))
    return np.array(rows, dtype=FAM_DTYPE)


def read_bim(path: str) -> np.ndarray:
    """.bim records (chromosome, SNP id, cM, bp position, allele 1, allele 2), one per SNP."""
    rows = []
    with open(path) as f:
        for n, line in enumerate(f, 1):
            t = line.split()
            if not t:
                continue
            if len(t) < 6:
                raise ValueError(f"{path}:{n}: a .bim line has 6 fields, found {len(t)}")
            rows.append((t[0], t[1], float(t[2]), int(t[3]), t[4], t[5]))
    return np.array(rows, dtype=BIM_DTYPE)


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
    if raw.size < 3 or bytes(raw[:3]) != MAGIC:
        raise ValueError(f"{path}: not a SNP-major PLINK 1 .bed file")
    bps = (n_individuals + 3) // 4
    payload = raw[3:]
    if n_snps is None:
        n_snps = payload.size // bps
    if payload.size < n_snps * bps:
        raise ValueError(f"{path}: payload has {payload.size} bytes, expected {n_snps * bps}")
    return payload[: n_snps * bps], n_individuals, n_snps, bps


class PlinkFileset:
    """A .bed/.bim/.fam triple (the SnpArray(filename) of the reference, src/driver.jl:148)."""

    def __init__(self, path: str):
        stem = _stem(path)
        self.stem = stem
        self.fam = read_fam(stem + ".fam")
        self.bim = read_bim(stem + ".bim") if os.path.exists(stem + ".bim") else None
        n_snps = None if self.bim is None else len(self.bim)
        self.data, self.n_individuals, self.n_snps, self.bytes_per_snp = read_bed(stem + ".bed", len(self.fam),
                                                                                  n_snps)

    @property
    def shape(self):
        return self.n_individuals, self.n_snps

    def payload(self, I: int | None = None, J: int | None = None):
        """(payload, bytes_per_snp, n_snp) for the first I individuals and the first 2J SNPs."""
        I = self.n_individuals if I is None else int(I)
        J = self.n_snps // 2 if J is None else int(J)
        if not (1 <= I <= self.n_individuals and 1 <= 2 * J <= self.n_snps):
            raise ValueError(f"fileset has {self.n_individuals} individuals and {self.n_snps} SNPs; asked for "
                             f"I={I}, 2J={2 * J}")
        return self.data[: 2 * J * self.bytes_per_snp], self.bytes_per_snp, 2 * J

    def codes(self, snps=None) -> np.ndarray:
        """2-bit codes (n_selected_snps x n_individuals) of the chosen SNPs (all when None)."""
        cols = np.arange(self.n_snps) if snps is None else _as_index(snps, self.n_snps)
        return unpack_codes(self.data, self.n_individuals, cols)


def _as_index(sel, n: int) -> np.ndarray:
    sel = np.asarray(sel)
    if sel.dtype == bool:
        if sel.size != n:
            raise ValueError(f"mask has {sel.size} entries, expected {n}")
        return np.nonzero(sel)[0]
    idx = sel.astype(np.int64).ravel()
    if idx.size and (idx.min() < 0 or idx.max() >= n):
        raise ValueError("index out of range")
    return idx


def unpack_codes(payload, n_individuals: int, snps) -> np.ndarray:
    """Codes (len(snps) x n_individuals) from a SNP-major payload (individual i at bits 2(i%4) of byte i/4)."""
    bps = (n_individuals + 3) // 4
    b = np.asarray(payload).reshape(-1, bps)[np.asarray(snps, dtype=np.int64)]
    out = np.empty((b.shape[0], 4 * bps), dtype=np.uint8)
    for r in range(4):
        out[:, r::4] = (b >> (2 * r)) & 3
    return out[:, :n_individuals]


def pack_codes(codes: np.ndarray) -> np.ndarray:
    """Inverse of unpack_codes: (n_snps x I) codes -> SNP-major payload, padding bits zero."""
    n, I = codes.shape
    bps = (I + 3) // 4
    c = np.zeros((n, 4 * bps), dtype=np.uint8)
    c[:, :I] = codes
    out = c[:, 0::4] | (c[:, 1::4] << 2) | (c[:, 2::4] << 4) | (c[:, 3::4] << 6)
    return np.ascontiguousarray(out, dtype=np.uint8).ravel()


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


def filter(src: str, rows, cols, des: str, chunk_snps: int = 65536) -> str:
    """Write des.bed/.bim/.fam holding the individuals ``rows`` and SNPs ``cols`` of fileset ``src``.

    ``rows``/``cols`` are boolean masks or index arrays (kept in the order given), as in
    ``SnpArrays.filter(src, rowinds, colinds; des)`` (src/driver.jl:103).  Returns ``des + ".bed"``."""
    fs = PlinkFileset(src)
    ri = _as_index(rows, fs.n_individuals)
    ci = _as_index(cols, fs.n_snps)
    des = _stem(des)
    with open(des + ".bed", "wb") as f:
        f.write(MAGIC)
        for c0 in range(0, ci.size, chunk_snps):
            codes = unpack_codes(fs.data, fs.n_individuals, ci[c0:c0 + chunk_snps])[:, ri]
            f.write(pack_codes(codes).tobytes())
    with open(fs.stem + ".fam") as f:
        fam_lines = [ln for ln in f if ln.strip()]
    with open(des + ".fam", "w") as f:
        f.writelines(fam_lines[i] if fam_lines[i].endswith("\n") else fam_lines[i] + "\n" for i in ri)
    if fs.bim is not None:
        with open(fs.stem + ".bim") as f:
            bim_lines = [ln for ln in f if ln.strip()]
        with open(des + ".bim", "w") as f:
            f.writelines(bim_lines[s] if bim_lines[s].endswith("\n") else bim_lines[s] + "\n" for s in ci)
    return des + ".bed"


# ---------------------------------------------------------------------------------------------------------
# fitted parameters
# ---------------------------------------------------------------------------------------------------------

def write_q(path: str, q: np.ndarray, fmt: str = "%.17g"):
    """Q (K x I) as I lines of K values."""
    np.savetxt(path, np.asarray(q, dtype=np.float64).T, fmt=fmt)


def read_q(path: str) -> np.ndarray:
    """Inverse of write_q: returns Q as K x I (Fortran order), ready for ``init_q``."""
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    return np.asfortranarray(a.T)


def write_p(path: str, p: np.ndarray, fmt: str = "%.17g"):
    """P (K x 4J) as 4J lines (block j, haplotype 00/01/10/11 in that order) of K values."""
    np.savetxt(path, np.asarray(p, dtype=np.float64).T, fmt=fmt)


def read_p(path: str) -> np.ndarray:
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if a.shape[0] % 4:
        raise ValueError(f"{path}: {a.shape[0]} lines is not a multiple of 4 (haplotypes per block)")
    return np.asfortranarray(a.T)


def save_fit(prefix: str, q: np.ndarray, p: np.ndarray) -> tuple:
    """Write ``{prefix}.{K}.Q`` and ``{prefix}.{K}.P``; returns the two paths."""
    K = q.shape[0]
    fq, fp = f"{prefix}.{K}.Q", f"{prefix}.{K}.P"
    write_q(fq, q)
    write_p(fp, p)
    return fq, fp
