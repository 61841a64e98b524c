"""CPU-side checks of the native boundary and host logic (no GPU compute calls)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from gen import unpack_codes
import haplo_amd
from haplo_amd import _native, plink
from haplo_amd.sim import init_params, simulate


def test_hip_library_loads_and_exports_header_symbols():
    L = _native.hip()
    syms = _native.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _native.HIP_LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(syms) <= exported


def test_every_header_function_is_bound_with_argtypes():
    L = _native.hip()
    for s in _native.header_symbols():
        fn = getattr(L, s)
        assert fn.argtypes is not None or s in ("hk_create_error",), s


def test_create_rejects_bad_arguments_without_gpu_work():
    L = _native.hip()
    h = ctypes.c_void_p()
    bed = np.zeros(16, dtype=np.uint8)
    assert L.hk_ctx_create(ctypes.byref(h), 4, 2, 17, _native.ptr(bed), 1, 4, 0) == -1  # K > 16
    assert b"K in 1..16" in L.hk_create_error()
    assert L.hk_ctx_create(ctypes.byref(h), 4, 3, 2, _native.ptr(bed), 1, 4, 0) == -1  # n_snp < 2J
    assert L.hk_ctx_create(ctypes.byref(h), 4, 2, 2, _native.ptr(bed), 1, 4, 7) == -1  # precision


def test_simulator_deterministic_and_thread_independent():
    a = simulate(101, 57, 3, fst=0.1, miss=0.05, seed=42)
    b = simulate(101, 57, 3, fst=0.1, miss=0.05, seed=42)
    assert np.array_equal(a.bed, b.bed) and np.array_equal(a.q_true, b.q_true)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    code = ("import sys; sys.path.insert(0, %r); from haplo_amd.sim import simulate; "
            "import numpy as np; s = simulate(101, 57, 3, fst=0.1, miss=0.05, seed=42); "
            "sys.stdout.write(s.bed.tobytes().hex())") % _native.PKG_ROOT
    out = subprocess.run(["python", "-c", code], env=env, capture_output=True, text=True, check=True).stdout
    assert bytes.fromhex(out) == a.bed.tobytes()
    c = simulate(101, 57, 3, fst=0.1, miss=0.05, seed=43)
    assert not np.array_equal(a.bed, c.bed)


def test_simulator_statistics():
    """Genotype frequencies follow the model: missing rate, hom/het proportions implied by truth."""
    I, J, K = 400, 300, 3
    s = simulate(I, J, K, fst=0.1, miss=0.03, seed=7)
    codes = unpack_codes(s.bed, I, 2 * J)
    assert abs((codes == 1).mean() - 0.03) < 0.005
    assert np.allclose(s.q_true.sum(0), 1) and np.allclose(s.p_true.reshape(K, 4, J, order="F").sum(1), 1)
    # expected allele-2 dosage at SNP 2j: 2 * sum_k q_ki (p_k,j,10 + p_k,j,11)
    P = s.p_true.reshape(K, 4, J, order="F")
    f1 = np.einsum("ki,kj->ij", s.q_true, P[:, 2, :] + P[:, 3, :])
    g = codes[0::2].T.astype(float)
    g[g == 1] = np.nan
    dos = np.where(g == 0, 0.0, np.where(g == 2, 1.0, np.where(g == 3, 2.0, np.nan)))
    assert abs(np.nanmean(dos) - 2 * f1.mean()) < 0.02


def test_init_params_normalised():
    q, p = init_params(30, 20, 4, seed=3)
    assert np.allclose(q.sum(0), 1) and q.min() > 0
    assert np.allclose(p.reshape(4, 4, 20, order="F").sum(1), 1)


def test_plink_roundtrip(tmp_path):
    s = simulate(37, 10, 2, seed=5)
    f = str(tmp_path / "x.bed")
    plink.write_bed(f, s.bed, 37, 20)
    payload, n, m, bps = plink.read_bed(f)
    assert (n, m, bps) == (37, 20, 10)
    assert np.array_equal(np.asarray(payload), s.bed)
    with pytest.raises(ValueError):
        plink.read_bed(f[:-4] + ".txt")


def test_run_admixture_argument_errors(tmp_path):
    s = simulate(16, 4, 2, seed=5)
    f = str(tmp_path / "y.bed")
    plink.write_bed(f, s.bed, 16, 8)
    with pytest.raises(ValueError):
        haplo_amd.run_admixture(f[:-4] + ".txt", 16, 4, 2)
    with pytest.raises(ValueError):
        haplo_amd.run_admixture(f, 16, 4, 2, sparsity=2, skfr_mode="other")
    with pytest.raises(ValueError):
        haplo_amd._filter_SKFR(f, 2, [2, 3])
    with pytest.raises(NotImplementedError):
        haplo_amd.run_admixture(f, 16, 4, 2, use_gpu=False)


def test_product_path_has_no_oracle_dependency():
    """The package and the HIP sources never reference the oracle."""
    root = _native.PKG_ROOT
    for dirpath, _, files in os.walk(root):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".c")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "hko_" not in txt and "liboracle" not in txt, fn


def test_julia_binding_covers_the_runner_level():
    """integration/julia/HipBackend.jl: every library function it calls is declared in the header and
    exported; it defines each name the reference's host code calls with d_cu !== nothing (cuda/runners.jl's
    runners, _cu_admixture_base, copyto_sync! and `.=` on the device arrays); and HipAdmixData carries every
    d_cu field that algorithms_outer.jl / algorithms_inner.jl read (q, p, q_next, p_next: outer :9-11,50,
    122,149, inner :36,42,80,124,182,250,326; XtX_q, Xtz_q, XtX_p, Xtz_p: inner :126,252)."""
    import re

    path = os.path.join(_native.REPO_ROOT, "integration", "julia", "HipBackend.jl")
    src = open(path).read()
    called = set(re.findall(r"ccall\(\(:(hk_[a-z0-9_]+)", src))
    assert called and called <= set(_native.header_symbols()), called - set(_native.header_symbols())
    L = _native.hip()
    assert all(hasattr(L, s) for s in called)
    for sig in ("function loglikelihood(d::HipAdmixData, ::Nothing, q::HipBuf{2}, p::HipBuf{2})",
                "function em!(d::HipAdmixData, ::Nothing)", "function em_p!(d::HipAdmixData, ::Nothing)",
                "update_q_cuda!(d::HipAdmixData, ::Nothing)", "update_p_cuda!(d::HipAdmixData, ::Nothing)",
                "function _cu_admixture_base(d::AdmixData2{T,T2}, g_la::SnpLinAlg{T2}, I::Int, J::Int)",
                "function OpenADMIXTURE.copyto_sync!(dsts::Vector{<:HipBuf}, srcs::Vector{<:AbstractArray})",
                "function OpenADMIXTURE.copyto_sync!(dsts::Vector{<:AbstractArray}, srcs::Vector{<:HipBuf})",
                "function Base.copyto!(dst::HipBuf, bc::Broadcast.Broadcasted{HipStyle})"):
        assert sig in src, sig
    struct = src[src.index("mutable struct HipAdmixData"):]
    struct = struct[: struct.index("\nend\n")]
    for f in ("q", "p", "q_next", "p_next", "XtX_q", "Xtz_q", "XtX_p", "Xtz_p"):
        assert re.search(rf"^\s+{f}::HipBuf\{{[23]\}}$", struct, re.M), f
    # buffer ids agree with the header's enum
    hdr = dict((n, int(v)) for n, v in re.findall(r"(HK_BUF_[A-Z_]+) = (\d+)", open(_native.HEADER).read()))
    jl = {}
    for names, vals in re.findall(r"^const (HK_BUF_[A-Z_, ]+) = (.*)$", src, re.M):
        jl.update(zip(names.split(", "), (int(v) for v in re.findall(r"Cint\((\d+)\)", vals))))
    assert jl == hdr, (jl, hdr)


def test_result_and_option_structs_match_the_header(tmp_path):
    """ctypes' FitOpts / FitResult have the C structs' size and field offsets (compiled from the header)."""
    fields_r = [f for f, _ in _native.FitResult._fields_]
    fields_o = [f for f, _ in _native.FitOpts._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "haplo_hip.h"', "int main(void) {",
             'printf("%zu %zu\\n", sizeof(hk_fit_result), sizeof(hk_fit_opts));']
    lines += [f'printf("%zu\\n", offsetof(hk_fit_result, {f}));' for f in fields_r]
    lines += [f'printf("%zu\\n", offsetof(hk_fit_opts, {f}));' for f in fields_o]
    lines += ["return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(_native.HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == ctypes.sizeof(_native.FitResult) and vals[1] == ctypes.sizeof(_native.FitOpts)
    offs_r = vals[2: 2 + len(fields_r)]
    offs_o = vals[2 + len(fields_r):]
    assert offs_r == [getattr(_native.FitResult, f).offset for f in fields_r]
    assert offs_o == [getattr(_native.FitOpts, f).offset for f in fields_o]
