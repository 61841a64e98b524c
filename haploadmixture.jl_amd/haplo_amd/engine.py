"""``HipAdmixData``: the device context of the MI355X engine (libhaplo_hip.so through its C ABI).

It plays the role of the reference's ``CuAdmixData`` + ``g_cu`` pair (src/cuda/structs.jl:1-44) and of
the runner functions in src/cuda/runners.jl: ``loglikelihood``, ``em!``/``em_p!``, ``update_q_cuda!``,
``update_p_cuda!``.  Arrays cross the boundary in the reference's column-major layouts (K x I q,
K x 4J p, K x K x n curvature).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native
from ._native import FitOpts, FitResult, HipError, ptr

PRECISIONS = {"fp64": 0, "mixed": 1, "fp32acc": 2}


def _F(a, shape):
    a = np.asfortranarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class HipAdmixData:
    def __init__(self, bed: np.ndarray, I: int, J: int, K: int, bytes_per_snp: int | None = None,
                 n_snp: int | None = None, precision: str = "fp64"):
        self._L = _native.hip()
        self.I, self.J, self.K = int(I), int(J), int(K)
        bps = (self.I + 3) // 4 if bytes_per_snp is None else int(bytes_per_snp)
        bed = np.ascontiguousarray(bed, dtype=np.uint8)
        nsnp = bed.size // bps if n_snp is None else int(n_snp)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        self.precision = precision
        self._bed_keepalive = bed
        h = ctypes.c_void_p()
        rc = self._L.hk_ctx_create(ctypes.byref(h), self.I, self.J, self.K, ptr(bed), bps, nsnp,
                                   PRECISIONS[precision])
        if rc != 0:
            raise HipError(rc, self._L.hk_create_error().decode())
        self._h = h
        del self._bed_keepalive

    @classmethod
    def shared(cls, base: "HipAdmixData", K: int, precision: str | None = None) -> "HipAdmixData":
        """A fit context over base's genotype upload (hk_ctx_create_shared); base must stay open."""
        self = cls.__new__(cls)
        self._L = _native.hip()
        self.I, self.J, self.K = base.I, base.J, int(K)
        self.precision = precision or base.precision
        if self.precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        h = ctypes.c_void_p()
        rc = self._L.hk_ctx_create_shared(ctypes.byref(h), base._h, self.K, PRECISIONS[self.precision])
        if rc != 0:
            raise HipError(rc, self._L.hk_create_error().decode())
        self._h = h
        self._base = base
        return self

    # -- plumbing ---------------------------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise HipError(rc, self._L.hk_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.hk_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- parameters -------------------------------------------------------------------------------
    def set_params(self, q, p):
        q = _F(q, (self.K, self.I))
        p = _F(p, (self.K, 4 * self.J))
        self._check(self._L.hk_set_params(self._h, ptr(q), ptr(p)))

    def get_params(self):
        q = np.zeros((self.K, self.I), order="F")
        p = np.zeros((self.K, 4 * self.J), order="F")
        self._check(self._L.hk_get_params(self._h, ptr(q), ptr(p)))
        return q, p

    # -- reference runner entry points ------------------------------------------------------------
    def loglikelihood(self, q=None, p=None, rows=False):
        """runners.jl:2-16 on (q, p) (or the current parameters)."""
        if q is not None or p is not None:
            self.set_params(q, p)
        ll = ctypes.c_double()
        r = np.zeros(self.I) if rows else None
        self._check(self._L.hk_loglik(self._h, ctypes.byref(ll), ptr(r)))
        return (ll.value, r) if rows else ll.value

    def em(self, fix_q=False):
        """em!(d_cu, g_cu) / em_p! (runners.jl:18-55): unnormalised (q_next, p_next)."""
        qn = None if fix_q else np.zeros((self.K, self.I), order="F")
        pn = np.zeros((self.K, 4 * self.J), order="F")
        self._check(self._L.hk_em(self._h, int(bool(fix_q)), ptr(qn), ptr(pn)))
        return qn, pn

    def update_q(self):
        """update_q_cuda! (runners.jl:58-72): (XtX_q K x K x I, Xtz_q K x I)."""
        XtX = np.zeros((self.K, self.K, self.I), order="F")
        Xtz = np.zeros((self.K, self.I), order="F")
        self._check(self._L.hk_update_q(self._h, ptr(XtX), ptr(Xtz)))
        return XtX, Xtz

    def update_p(self):
        """update_p_cuda! (runners.jl:76-92): (XtX_p K x K x 4J, Xtz_p K x 4J)."""
        XtX = np.zeros((self.K, self.K, 4 * self.J), order="F")
        Xtz = np.zeros((self.K, 4 * self.J), order="F")
        self._check(self._L.hk_update_p(self._h, ptr(XtX), ptr(Xtz)))
        return XtX, Xtz

    def double_missing(self):
        dm = np.zeros(self.I, dtype=np.int64)
        self._check(self._L.hk_double_missing(self._h, ptr(dm)))
        return dm

    def decode(self):
        """BED codes held by the device stream, shape (2J, I)."""
        codes = np.zeros((2 * self.J, self.I), dtype=np.uint8)
        self._check(self._L.hk_decode(self._h, ptr(codes)))
        return codes

    def count_double_missing_file(self, bed: np.ndarray, n_snp: int, bytes_per_snp: int | None = None):
        """count_double_missing!(dm, g) over every SNP pair of the whole file (structs.jl:169 -> utils.jl:1-4),
        as the reference's EM normaliser uses it; returns the new counts."""
        bps = (self.I + 3) // 4 if bytes_per_snp is None else int(bytes_per_snp)
        bed = np.ascontiguousarray(bed, dtype=np.uint8)
        if bed.size < int(n_snp) * bps:
            raise ValueError("payload shorter than n_snp * bytes_per_snp")
        self._check(self._L.hk_count_double_missing_file(self._h, ptr(bed), bps, int(n_snp)))
        return self.double_missing()

    def skfr_path(self, k: int, sparsities, mode: str = "global", max_iter: int = 50, tries: int = 1,
                  seed: int = 1) -> list:
        """hk_skfr_path (sparsekmeans_path): one result dict per sparsity (decreasing), each run after the
        first started from the previous one's labels."""
        if mode not in ("global", "local"):
            raise ValueError("mode must be 'global' or 'local'")
        sp = np.ascontiguousarray(sparsities, dtype=np.int32)
        n = sp.size
        clusters = np.zeros((n, self.I), dtype=np.int32)
        rooms = [2 * min(self.J, int(k) * int(s)) for s in sp]
        aims = np.zeros(max(1, sum(rooms)), dtype=np.int32)
        n_aims = np.zeros(n, dtype=np.int32)
        wss = np.zeros(n)
        iters = np.zeros(n, dtype=np.int32)
        self._check(self._L.hk_skfr_path(self._h, int(k), n, ptr(sp), int(mode == "local"), int(max_iter),
                                         int(tries), int(seed) & ((1 << 64) - 1), ptr(clusters), ptr(aims),
                                         ptr(n_aims), ptr(wss), ptr(iters)))
        out, off = [], 0
        for t in range(n):
            out.append(dict(clusters=clusters[t].copy(), aims=aims[off: off + n_aims[t]].copy(), wss=float(wss[t]),
                            iters=int(iters[t])))
            off += rooms[t]
        return out

    # -- runner level: the device arrays of CuAdmixData by name (cuda/structs.jl:1-15) --------------
    BUFFERS = {"q": 0, "p": 1, "q_next": 2, "p_next": 3, "XtX_q": 4, "Xtz_q": 5, "XtX_p": 6, "Xtz_p": 7}

    def buf_shape(self, name: str) -> tuple:
        K, I, J4 = self.K, self.I, 4 * self.J
        return {"q": (K, I), "q_next": (K, I), "p": (K, J4), "p_next": (K, J4), "XtX_q": (K, K, I),
                "Xtz_q": (K, I), "XtX_p": (K, K, J4), "Xtz_p": (K, J4)}[name]

    def buf_upload(self, name: str, a):
        """copyto_sync!([d_cu.<name>], [a])."""
        self._check(self._L.hk_buf_upload(self._h, self.BUFFERS[name], ptr(_F(a, self.buf_shape(name)))))

    def buf_download(self, name: str) -> np.ndarray:
        """copyto_sync!([a], [d_cu.<name>])."""
        a = np.zeros(self.buf_shape(name), order="F")
        self._check(self._L.hk_buf_download(self._h, self.BUFFERS[name], ptr(a)))
        return a

    def buf_copy(self, dst: str, src: str):
        """d_cu.<dst> .= d_cu.<src>."""
        self._check(self._L.hk_buf_copy(self._h, self.BUFFERS[dst], self.BUFFERS[src]))

    def loglikelihood_buf(self, q: str = "q", p: str = "p") -> float:
        """loglikelihood(d_cu, g_cu, d_cu.<q>, d_cu.<p>)."""
        ll = ctypes.c_double()
        self._check(self._L.hk_loglik_buf(self._h, self.BUFFERS[q], self.BUFFERS[p], ctypes.byref(ll)))
        return ll.value

    def em_buf(self, fix_q: bool = False):
        """em!(d_cu, g_cu) / em_p!(d_cu, g_cu): unnormalised sums into d_cu.q_next, d_cu.p_next."""
        self._check(self._L.hk_em_buf(self._h, int(bool(fix_q))))

    def update_q_buf(self):
        """update_q_cuda!(d_cu, g_cu): d_cu.XtX_q, d_cu.Xtz_q from d_cu.q, d_cu.p."""
        self._check(self._L.hk_update_q_buf(self._h))

    def update_p_buf(self):
        """update_p_cuda!(d_cu, g_cu): d_cu.XtX_p, d_cu.Xtz_p from d_cu.q, d_cu.p."""
        self._check(self._L.hk_update_p_buf(self._h))

    # -- device-resident row steps ----------------------------------------------------------------
    def q_step(self, penalty=False, rho=1e7):
        qn = np.zeros((self.K, self.I), order="F")
        self._check(self._L.hk_q_step(self._h, int(penalty), float(rho), ptr(qn)))
        return qn

    def p_step(self, q_pass=None, penalty=True, rho=1e7):
        pn = np.zeros((self.K, 4 * self.J), order="F")
        qp = None if q_pass is None else _F(q_pass, (self.K, self.I))
        self._check(self._L.hk_p_step(self._h, ptr(qp), int(penalty), float(rho), ptr(pn)))
        return pn

    # -- whole fit ----------------------------------------------------------------------------------
    @staticmethod
    def default_opts(**kw) -> FitOpts:
        o = FitOpts()
        _native.hip().hk_fit_opts_default(ctypes.byref(o))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown fit option {k}")
            setattr(o, k, v)
        return o

    def fit(self, opts: FitOpts | None = None, ll_start: float = math.nan, trace: bool = True, **kw):
        o = opts if opts is not None else self.default_opts(**kw)
        tr = np.full(max(int(o.n_iter), 1), np.nan) if trace else None
        tt = np.full(max(int(o.n_iter), 1), np.nan) if trace else None
        res = FitResult()
        self._check(self._L.hk_fit_ex(self._h, ctypes.byref(o), float(ll_start), ptr(tr), ptr(tt),
                                      ctypes.byref(res)))
        out = res.as_dict()
        if trace:
            out["trace"] = tr[: res.iters].copy()
            out["t_trace"] = tt[: res.iters].copy()
        return out

    # -- tuning / timing ----------------------------------------------------------------------------
    def set_tuning(self, q_chunks=0, p_chunks=0):
        self._check(self._L.hk_set_tuning(self._h, int(q_chunks), int(p_chunks)))

    def get_tuning(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.hk_get_tuning(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_pass_shape(self, q_shape=0, p_shape=0):
        """Curvature-pass variant per side (hk_set_pass_shape): 0 = default for K, 1 = the other staging form."""
        self._check(self._L.hk_set_pass_shape(self._h, int(q_shape), int(p_shape)))

    def get_pass_shape(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.hk_get_pass_shape(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def time_pass(self, kind: int, reps: int = 5) -> float:
        ms = ctypes.c_double()
        self._check(self._L.hk_time_pass(self._h, int(kind), int(reps), ctypes.byref(ms)))
        return ms.value

    def copy_bench(self, nbytes: int, reps: int = 3) -> float:
        """GB/s of a plain device copy of nbytes (read + write counted)."""
        g = ctypes.c_double()
        self._check(self._L.hk_copy_bench(self._h, int(nbytes), int(reps), ctypes.byref(g)))
        return g.value

    def autotune(self, reps: int = 3):
        """Time candidate chunk shapes of the passes on this data and keep the fastest (hk_autotune)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.hk_autotune(self._h, int(reps), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sync(self):
        self._check(self._L.hk_sync(self._h))

    # -- SKFR marker pre-selection ------------------------------------------------------------------
    def skfr(self, k: int, sparsity: int, mode: str = "global", max_iter: int = 50, tries: int = 1, seed: int = 1,
             init=None, profile: bool = False) -> dict:
        """hk_skfr on this context's J blocks: sparse k-means with feature ranking (mode "global" =
        sparsekmeans1, "local" = sparsekmeans2).  Returns clusters (I,), aims (0-based SNP indices, most
        important block first), wss and iters of the best of `tries` runs."""
        if mode not in ("global", "local"):
            raise ValueError("mode must be 'global' or 'local'")
        clusters = np.zeros(self.I, dtype=np.int32)
        aims = np.zeros(2 * max(1, min(self.J, int(k) * int(sparsity))), dtype=np.int32)
        n = ctypes.c_int()
        wss = ctypes.c_double()
        iters = ctypes.c_int()
        ini = None
        prof = np.zeros(8) if profile else None
        if init is not None:
            ini = np.ascontiguousarray(init, dtype=np.int32)
            if ini.shape != (int(k),):
                raise ValueError(f"init must hold k={k} individual indices")
        self._check(self._L.hk_skfr(self._h, int(k), int(sparsity), int(mode == "local"), int(max_iter), int(tries),
                                    int(seed) & ((1 << 64) - 1), ptr(ini), ptr(clusters), ptr(aims), ctypes.byref(n),
                                    ctypes.byref(wss), ctypes.byref(iters), ptr(prof)))
        out = dict(clusters=clusters, aims=aims[: n.value].copy(), wss=wss.value, iters=iters.value)
        if profile:
            out["profile"] = {"centers_ms": prof[0], "centers_launches": int(prof[1]), "dist_ms": prof[2],
                              "dist_launches": int(prof[3]), "assign_ms": prof[4], "assign_launches": int(prof[5]),
                              "device_ms": prof[6]}
        return out


def fit_batch(engines, opts_list, ll_start=None):
    """hk_fit_batch: fits of several contexts run concurrently (each enqueues its next
    iteration as soon as its previous one completed); returns (results, seconds_at_finish)."""
    n = len(engines)
    L = _native.hip()
    handles = (ctypes.c_void_p * n)(*[e._h for e in engines])
    opts = (FitOpts * n)(*opts_list)
    res = (FitResult * n)()
    fin = np.zeros(n)
    lls = None if ll_start is None else np.ascontiguousarray(ll_start, dtype=np.float64)
    rc = L.hk_fit_batch(handles, n, opts, ptr(lls), res, ptr(fin))
    if rc != 0:
        msgs = "; ".join(L.hk_last_error(e._h).decode() for e in engines if L.hk_last_error(e._h))
        raise HipError(rc, msgs or L.hk_create_error().decode())
    return [r.as_dict() for r in res], fin
