"""The C ABI from a plain C program (tests/c_abi/abi_driver.c, built with gcc against include/haplo_hip.h and
libhaplo_hip.so): the same calls a runner-level binding makes -- create, set parameters, log-likelihood, both
curvature passes, EM, the device-resident fit, read back -- with no Python in the process.  Its results must
be bitwise those of the same calls through ctypes (one library, deterministic passes), and the FP64 engine's
within 1e-12 of the oracle."""
import os
import subprocess

import numpy as np
import pytest

from gen import case, rel
from haplo_amd import HipAdmixData, _native

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c_abi", "abi_driver.c")


def _build(tmp_path):
    lib = os.path.dirname(_native.HIP_LIB)
    exe = str(tmp_path / "abi_driver")
    subprocess.run(["gcc", "-O2", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.dirname(_native.HEADER),
                    SRC, "-L", lib, "-lhaplo_hip", f"-Wl,-rpath,{lib}", "-Wl,-rpath-link,/opt/rocm/lib", "-lm",
                    "-o", exe], check=True)
    return exe


def test_c_driver_builds_against_header_and_library(tmp_path):
    """Compiles warning-free as C99 and links; with no inputs it stops before any HIP call."""
    if not os.path.exists(_native.HIP_LIB):
        pytest.skip("libhaplo_hip.so not built")
    exe = _build(tmp_path)
    r = subprocess.run([exe, str(tmp_path), "5", "4", "2", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, (r.returncode, r.stderr)


def _f64(path, n=None):
    a = np.fromfile(path, dtype="<f8")
    return a if n is None else a[:n]


@pytest.mark.gpu
@pytest.mark.parametrize("I,J,K,prec", [(301, 130, 5, "fp64"), (257, 65, 6, "mixed"), (130, 45, 12, "mixed")])
def test_c_driver_matches_ctypes_and_oracle(I, J, K, prec, tmp_path, oracle):
    print(f"C driver I={I} J={J} K={K} {prec}", flush=True)
    s, q, p = case(I, J, K)
    bps = (I + 3) // 4
    s.bed[: 2 * J * bps].tofile(tmp_path / "bed.bin")
    np.asarray(q, dtype="<f8").ravel(order="F").tofile(tmp_path / "q.bin")
    np.asarray(p, dtype="<f8").ravel(order="F").tofile(tmp_path / "p.bin")
    code = {"fp64": 0, "mixed": 1}[prec]
    r = subprocess.run([_build(tmp_path), str(tmp_path), str(I), str(J), str(K), str(code)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    nq, npp = K * I, K * 4 * J
    ll = _f64(tmp_path / "ll.bin")
    uq, up = _f64(tmp_path / "uq.bin"), _f64(tmp_path / "up.bin")
    em, fit, qp = _f64(tmp_path / "em.bin"), _f64(tmp_path / "fit.bin"), _f64(tmp_path / "qp.bin")

    with HipAdmixData(s.bed, I, J, K, precision=prec) as e:
        e.set_params(q, p)
        ll0, rows0 = e.loglikelihood(rows=True)
        XtXq, Xtzq = e.update_q()
        XtXp, Xtzp = e.update_p()
        qn, pn = e.em()
        res = e.fit()
        qf, pf = e.get_params()
    assert ll[0] == ll0 and np.array_equal(ll[1:], rows0)
    assert np.array_equal(uq, np.concatenate([XtXq.ravel(order="F"), Xtzq.ravel(order="F")]))
    assert np.array_equal(up, np.concatenate([XtXp.ravel(order="F"), Xtzp.ravel(order="F")]))
    assert np.array_equal(em, np.concatenate([qn.ravel(order="F"), pn.ravel(order="F")]))
    assert fit[0] == res["ll"] and fit[1] == res["ll_init"] and int(fit[2]) == res["iters"]
    assert int(fit[3]) == int(res["converged"]) and int(fit[4]) == res["qp_failures"]
    assert np.array_equal(fit[5:], res["trace"])
    assert np.array_equal(qp[:nq], qf.ravel(order="F")) and np.array_equal(qp[nq:], pf.ravel(order="F"))
    assert qp.size == nq + npp

    if prec == "fp64":
        assert abs(ll[0] - oracle.loglik(s.bed, I, J, q, p)) <= 1e-12 * abs(ll[0])
        oX, oz = oracle.update_q(s.bed, I, J, q, p)
        assert rel(uq[: nq * K], oX.ravel(order="F")) <= 1e-12 and rel(uq[nq * K:], oz.ravel(order="F")) <= 1e-12
