"""CPU oracle for the HaploADMIXTURE likelihood engine -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (built from ``oracle/hk_oracle.c`` by ``oracle/Makefile``),
the plain-C FP64 restatement of the reference's ``src/loops.jl`` passes and of the QN/SQP loop in
``src/algorithms_inner.jl`` / ``src/algorithms_outer.jl``.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg import this module, and only to check or to time the CPU path;
the product library never routes through it.

Each wrapper names the reference function it restates (file:line); the C file carries the same map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f64p_f = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double


class FitOpts(ctypes.Structure):
    """Mirror of ``hko_fit_opts``; defaults are run_admixture's (src/driver.jl:38-57,
    src/algorithms_outer.jl:58-61)."""

    _fields_ = [
        ("em_iters", _int),
        ("n_iter", _int),
        ("rtol", _dbl),
        ("Qn", _int),
        ("fix_q", _int),
        ("fix_p", _int),
        ("q_penalty", _int),
        ("p_penalty", _int),
        ("rho", _dbl),
        ("eps", _dbl),
        ("guarantee_increase", _int),
        ("qn_mode", _int),
    ]


def build() -> str:
    """Compile liboracle.so (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.hko_count_double_missing.argtypes = [_u8p, _i64, _i64, _i64, _i64p]
        L.hko_loglik.argtypes = [_u8p, _i64, _i64, _i64, _int, _f64p_f, _f64p_f]
        L.hko_loglik.restype = _dbl
        L.hko_loglik_rows.argtypes = [_u8p, _i64, _i64, _i64, _int, _f64p_f, _f64p_f, _f64p]
        L.hko_em.argtypes = [_u8p, _i64, _i64, _i64, _int, _f64p_f, _f64p_f, _f64p_f, _f64p_f, _int, _int]
        L.hko_update_q.argtypes = [_u8p, _i64, _i64, _i64, _int, _f64p_f, _f64p_f, _f64p_f, _f64p_f]
        L.hko_update_p.argtypes = [_u8p, _i64, _i64, _i64, _int, _f64p_f, _f64p_f, _f64p_f, _f64p_f]
        L.hko_project_simplex.argtypes = [_f64p, _int, _dbl]
        L.hko_project_q.argtypes = [_f64p_f, _int, _i64, _dbl]
        L.hko_project_p.argtypes = [_f64p_f, _int, _i64, _dbl]
        L.hko_qp_solve.argtypes = [_int, _f64p_f, _f64p, _f64p, _f64p, _int, _i32p, _f64p]
        L.hko_qp_solve.restype = _int
        L.hko_q_step.argtypes = [_int, _i64, _f64p_f, _f64p_f, _f64p_f, _f64p_f, _int, _dbl, _dbl]
        L.hko_q_step.restype = _int
        L.hko_p_step.argtypes = [_int, _i64, _f64p_f, _f64p_f, _f64p_f, _f64p_f, _int, _dbl, _dbl]
        L.hko_p_step.restype = _int
        L.hko_q_step_ex.argtypes = [_int, _i64, _f64p_f, _f64p_f, _f64p_f, _f64p_f, _int, _dbl, _dbl, _f64p_f]
        L.hko_q_step_ex.restype = _int
        L.hko_p_step_ex.argtypes = [_int, _i64, _f64p_f, _f64p_f, _f64p_f, _f64p_f, _int, _dbl, _dbl, _f64p_f]
        L.hko_p_step_ex.restype = _int
        L.hko_em_normalize.argtypes = [_int, _i64, _i64, _i64p, _f64p_f, _f64p_f, _int, _int, _dbl]
        L.hko_zal.argtypes = [_i64, _int, _int, _f64p_f, _f64p_f, _f64p, _f64p, _f64p, _f64p]
        L.hko_zal.restype = _int
        L.hko_qn.argtypes = [_i64, _int, _int, _int, _f64p_f, _f64p_f, _f64p, _f64p, _f64p, _f64p]
        L.hko_qn.restype = _int
        L.hko_fit.argtypes = [_u8p, _i64, _i64, _i64, _int, ctypes.POINTER(FitOpts), _f64p_f,
                              ctypes.POINTER(_dbl), _f64p]
        L.hko_fit.restype = _int
        L.hko_threads.restype = _int
        _lib = L
    return _lib


def _bps(I: int) -> int:
    return (I + 3) // 4


def _F(a):
    return np.asfortranarray(a, dtype=np.float64)


def count_double_missing(bed, I, J):
    """utils.jl:1-18 (first J blocks)."""
    dm = np.zeros(I, dtype=np.int64)
    lib().hko_count_double_missing(bed, _bps(I), I, J, dm)
    return dm


def loglik(bed, I, J, q, p):
    """loglikelihood_full2 (algorithms_inner.jl:20-27) -> loops.jl:236-258."""
    K = q.shape[0]
    return lib().hko_loglik(bed, _bps(I), I, J, K, _F(q), _F(p))


def loglik_rows(bed, I, J, q, p):
    K = q.shape[0]
    out = np.zeros(I)
    lib().hko_loglik_rows(bed, _bps(I), I, J, K, _F(q), _F(p), out)
    return out


def em(bed, I, J, q, p, fix_q=False, fix_p=False):
    """em_loop! / em_p_loop! (loops.jl:260-326): unnormalised q_next, p_next."""
    K = q.shape[0]
    qn = np.zeros((K, I), order="F")
    pn = np.zeros((K, 4 * J), order="F")
    lib().hko_em(bed, _bps(I), I, J, K, _F(q), _F(p), qn, pn, int(fix_q), int(fix_p))
    return qn, pn


def update_q(bed, I, J, q, p):
    """update_q_loop! (loops.jl:328-365): XtX (K,K,I), Xtz (K,I)."""
    K = q.shape[0]
    XtX = np.zeros((K, K, I), order="F")
    Xtz = np.zeros((K, I), order="F")
    lib().hko_update_q(bed, _bps(I), I, J, K, _F(q), _F(p), XtX, Xtz)
    return XtX, Xtz


def update_p(bed, I, J, q, p):
    """update_p_loop! (loops.jl:367-403): XtX (K,K,4J), Xtz (K,4J)."""
    K = q.shape[0]
    XtX = np.zeros((K, K, 4 * J), order="F")
    Xtz = np.zeros((K, 4 * J), order="F")
    lib().hko_update_p(bed, _bps(I), I, J, K, _F(q), _F(p), XtX, Xtz)
    return XtX, Xtz


def project_simplex(v, eps=1e-5):
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    lib().hko_project_simplex(v, v.size, eps)
    return v


def project_q(q, eps=1e-5):
    q = _F(q).copy(order="F")
    lib().hko_project_q(q, q.shape[0], q.shape[1], eps)
    return q


def project_p(p, eps=1e-5):
    p = _F(p).copy(order="F")
    lib().hko_project_p(p, p.shape[0], p.shape[1] // 4, eps)
    return p


def qp_solve(H, g, lo, hi, grp=None):
    """Dense active-set QP: min .5x'Hx + g'x, lo <= x <= hi, sum over each group of x = 0."""
    n = len(g)
    if grp is None:
        grp = -np.ones(n, dtype=np.int32)
        m = 0
    else:
        grp = np.ascontiguousarray(grp, dtype=np.int32)
        m = int(grp.max()) + 1 if (grp >= 0).any() else 0
    x = np.zeros(n)
    it = lib().hko_qp_solve(n, _F(H), np.ascontiguousarray(g, dtype=np.float64),
                            np.ascontiguousarray(lo, dtype=np.float64),
                            np.ascontiguousarray(hi, dtype=np.float64), m, grp, x)
    return x, it


def q_step(XtX, Xtz, q, penalty=False, rho=1e7, eps=1e-5):
    """update_q! QP + step + projection (algorithms_inner.jl:135-207)."""
    K, I = q.shape
    qn = np.zeros((K, I), order="F")
    fails = lib().hko_q_step(K, I, _F(XtX), _F(Xtz), _F(q), qn, int(penalty), rho, eps)
    return qn, fails


def p_step(XtX, Xtz, p, penalty=True, rho=1e7, eps=1e-5):
    """update_p! QP + step + projection (algorithms_inner.jl:259-351)."""
    K = p.shape[0]
    J = p.shape[1] // 4
    pn = np.zeros(p.shape, order="F")
    fails = lib().hko_p_step(K, J, _F(XtX), _F(Xtz), _F(p), pn, int(penalty), rho, eps)
    return pn, fails


def q_step_delta(XtX, Xtz, q, penalty=False, rho=1e7, eps=1e-5):
    """q_step plus the raw step qdiff before projection: (q_next, qdiff, fails)."""
    K, I = q.shape
    qn = np.zeros((K, I), order="F")
    dq = np.zeros((K, I), order="F")
    fails = lib().hko_q_step_ex(K, I, _F(XtX), _F(Xtz), _F(q), qn, int(penalty), rho, eps, dq)
    return qn, dq, fails


def p_step_delta(XtX, Xtz, p, penalty=True, rho=1e7, eps=1e-5):
    """p_step plus the raw step pdiff before projection: (p_next, pdiff, fails)."""
    K = p.shape[0]
    J = p.shape[1] // 4
    pn = np.zeros(p.shape, order="F")
    dp = np.zeros(p.shape, order="F")
    fails = lib().hko_p_step_ex(K, J, _F(XtX), _F(Xtz), _F(p), pn, int(penalty), rho, eps, dp)
    return pn, dp, fails


def em_normalize(qn, pn, dm, J, eps=1e-5):
    """em! normalisation + projections (algorithms_inner.jl:66-76)."""
    qn = _F(qn).copy(order="F")
    pn = _F(pn).copy(order="F")
    K, I = qn.shape
    lib().hko_em_normalize(K, I, J, np.ascontiguousarray(dm, dtype=np.int64), qn, pn, 0, 0, eps)
    return qn, pn


def zal(U, V, x0, x1, x2, it, Qn=3):
    """update_UV! + update_QN! (algorithms_outer.jl:128-133).  U, V are updated in place."""
    N = x0.size
    out = np.zeros(N)
    rc = lib().hko_zal(N, Qn, it, U, V, np.ascontiguousarray(x0), np.ascontiguousarray(x1),
                       np.ascontiguousarray(x2), out)
    return out, rc


def qn_step(U, V, x0, x1, x2, it, Qn=3, mode=0):
    """Quasi-Newton point of admixture_qn! (mode 0: ZAL, 1: LBQN); U, V are updated in place."""
    N = x0.size
    out = np.zeros(N)
    rc = lib().hko_qn(N, Qn, it, mode, U, V, np.ascontiguousarray(x0), np.ascontiguousarray(x1),
                      np.ascontiguousarray(x2), out)
    return out, rc


def fit(bed, I, J, q0, p0, em_iters=5, n_iter=1000, rtol=1e-7, Qn=3, fix_q=False, fix_p=False,
        q_penalty=False, p_penalty=True, rho=1e7, eps=1e-5, guarantee_increase=False, qn_mode=0):
    """init_em! + admixture_qn! (algorithms_outer.jl:1-172), with the step halving of update_q! /
    update_p! when guarantee_increase (algorithms_inner.jl:170-199,314-343).
    Returns (q, p, ll, iters, trace)."""
    K = q0.shape[0]
    x = np.asfortranarray(np.concatenate([q0, p0], axis=1))
    o = FitOpts(em_iters, n_iter, rtol, Qn, int(fix_q), int(fix_p), int(q_penalty), int(p_penalty), rho, eps,
                int(guarantee_increase), int(qn_mode))
    ll = _dbl(0.0)
    trace = np.zeros(max(n_iter, 1))
    iters = lib().hko_fit(bed, _bps(I), I, J, K, ctypes.byref(o), x, ctypes.byref(ll), trace)
    return x[:, :I].copy(order="F"), x[:, I:].copy(order="F"), ll.value, iters, trace[:iters].copy()


def threads() -> int:
    return lib().hko_threads()
