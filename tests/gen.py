"""Shared test-data helpers (seeded synthetic BED + starting parameters) and the element-wise check."""
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


def rel(a, b, floor=0.0):
    """Largest element-wise error max_i |a_i - b_i| / max(|b_i|, floor).

    floor = 0 checks every element against its own magnitude.  The pass outputs (log-likelihood rows,
    gradients, curvature, EM sums) are sums of same-signed terms, so a relative bound per element is what
    their arithmetic guarantees, down to the smallest entry.  floor = 1 turns the check into an absolute
    one for values of order one or below (parameters on the simplex, log-likelihood traces scaled by
    their own first value are not meant here).  An element whose expected value is 0 (floor 0) must
    match exactly."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch {a.shape} vs {b.shape}")
    if a.size == 0:
        return 0.0
    d = np.abs(a - b)
    den = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(d == 0, 0.0, d / den)
    r = np.where(np.isnan(r), np.inf, r)
    return float(np.max(r))
