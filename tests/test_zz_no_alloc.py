"""GPU: hk_loglik_noalloc (the entry point that allocates nothing; its scratch is a fixed device array)
against hk_loglik on the same parameters, through the C ABI, on two calls in a row where the second has a
smaller input, and its rejection of a problem larger than the scratch.

Device buffers for the inputs and outputs come from the HIP runtime the library itself links
(libamdhip64.so.7), called through ctypes.  Tolerance: the FP64 engine bit for bit (same chunking, same
kernels), the MIXED engine 1e-12 relative (same FP32 products; the staged parameters are the same FP32
values)."""
import ctypes

import numpy as np
import pytest

from gen import case
from haplo_amd import HipAdmixData, _native
from haplo_amd._native import ptr

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return L


class DevBuf:
    def __init__(self, rt, nbytes):
        self.rt = rt
        self.p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert self.rt.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, _H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        if self.p:
            self.rt.hipFree(self.p)
            self.p = ctypes.c_void_p()


def _lib():
    return _native.hip()


def _noalloc(e, q, p, rows=True):
    L, rt = _lib(), _hip()
    qf = np.asfortranarray(q).ravel(order="F")
    pf = np.asfortranarray(p).ravel(order="F")
    bq, bp, bl = DevBuf(rt, qf.nbytes), DevBuf(rt, pf.nbytes), DevBuf(rt, 8)
    br = DevBuf(rt, 8 * e.I) if rows else None
    try:
        bq.put(qf)
        bp.put(pf)
        rc = L.hk_loglik_noalloc(e._h, bq.p, bp.p, bl.p, br.p if rows else None)
        if rc != 0:
            raise AssertionError(L.hk_last_error(e._h).decode())
        e.sync()
        ll = float(bl.get(np.float64, 1)[0])
        r = br.get(np.float64, e.I) if rows else None
    finally:
        for b in (bq, bp, bl, br):
            if b is not None:
                b.free()
    return ll, r


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matches_loglik_two_calls_second_smaller(prec):
    for (I, J, K, seed) in ((3000, 5000, 6, 3), (257, 130, 4, 8), (200, 90, 12, 5)):  # K = 12: the wide kernels
        print(f"noalloc {prec} I={I} J={J} K={K}", flush=True)
        s, q, p = case(I, J, K, seed=seed, miss=0.05)
        with HipAdmixData(s.bed, I, J, K, precision=prec) as e:
            e.set_params(q, p)
            ref, ref_rows = e.loglikelihood(rows=True)
            ll, rows = _noalloc(e, q, p, rows=True)
            ll2, _ = _noalloc(e, q, p, rows=False)  # rows kept in the scratch instead
            if prec == "fp64":
                assert ll == ref and ll2 == ref
                assert np.array_equal(rows, ref_rows)
            else:
                assert abs(ll - ref) <= 1e-12 * abs(ref) and abs(ll2 - ref) <= 1e-12 * abs(ref)
                assert np.max(np.abs(rows - ref_rows) / np.abs(ref_rows)) <= 1e-12


def test_noalloc_rejects_problem_beyond_scratch():
    I, J, K = 300, 1_500_000, 8  # p padded alone is 8 * 4 * 1.5M doubles = 384 MB > 64 MiB
    L = _lib()
    bed = np.zeros(2 * J * ((I + 3) // 4), dtype=np.uint8)
    with HipAdmixData(bed, I, J, K, precision="fp64") as e:
        dummy = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
        rc = L.hk_loglik_noalloc(e._h, dummy, dummy, dummy, None)
        assert rc == -1
        assert "maximum" in L.hk_last_error(e._h).decode()
