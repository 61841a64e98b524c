"""Shared test-data helpers (seeded synthetic BED + starting parameters)."""
import numpy as np

from haplo_amd.sim import init_params, simulate


def case(I, J, K, seed=3, miss=0.05, fst=0.1, init_seed=11):
    s = simulate(I, J, K, fst=fst, miss=miss, seed=seed, truth=True)
    q, p = init_params(I, J, K, seed=init_seed)
    return s, q, p


def unpack_codes(bed, I, nsnp):
    """BED codes (nsnp x I) unpacked with numpy straight from the PLINK layout."""
    bps = (I + 3) // 4
    b = bed[: nsnp * bps].reshape(nsnp, bps)
    out = np.empty((nsnp, bps * 4), dtype=np.uint8)
    for r in range(4):
        out[:, r::4] = (b >> (2 * r)) & 3
    return out[:, :I]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
