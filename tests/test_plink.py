"""PLINK fileset I/O: sidecar parsing, subsetting (SnpArrays.filter analogue), Q/P files.  CPU only."""
import numpy as np
import pytest

from gen import unpack_codes
from haplo_amd import plink
from haplo_amd.sim import simulate


def _fileset(tmp_path, I=37, J=10, seed=5, name="x"):
    s = simulate(I, J, 2, seed=seed, miss=0.1)
    f = str(tmp_path / f"{name}.bed")
    plink.write_bed(f, s.bed, I, 2 * J)
    return f, s


def test_sidecars_parse(tmp_path):
    f, s = _fileset(tmp_path)
    fs = plink.PlinkFileset(f)
    assert fs.shape == (37, 20) and fs.bytes_per_snp == 10
    assert fs.fam["iid"][3] == "ind3" and fs.fam["phenotype"][0] == "-9"
    assert fs.bim["snpid"][7] == "snp7" and fs.bim["pos"][7] == 8 and fs.bim["allele2"][0] == "G"
    with open(f[:-4] + ".bim", "a") as h:
        h.write("1 bad 0\n")
    with pytest.raises(ValueError):
        plink.read_bim(f[:-4] + ".bim")


def test_pack_unpack_exhaustive_bytes():
    # every byte value in every slot: unpack then pack is the identity on full bytes
    payload = np.arange(256, dtype=np.uint8)
    codes = plink.unpack_codes(payload, 4, np.arange(256))
    assert codes.shape == (256, 4)
    for r in range(4):
        assert np.array_equal(codes[:, r], (payload >> (2 * r)) & 3)
    assert np.array_equal(plink.pack_codes(codes), payload)


@pytest.mark.parametrize("rows,cols", [
    (np.arange(37), np.arange(20)),                         # identity
    (np.arange(13), np.arange(8)),                          # first I individuals, first 2J SNPs
    (np.array([36, 0, 5, 5, 17]), np.array([19, 3, 4])),    # reordered, repeated, ragged
    ("mask", "mask"),
])
def test_filter_matches_numpy_selection(tmp_path, rows, cols):
    f, s = _fileset(tmp_path)
    full = unpack_codes(s.bed, 37, 20)
    if isinstance(rows, str):
        rng = np.random.default_rng(1)
        rows = rng.random(37) < 0.5
        cols = rng.random(20) < 0.5
    d = plink.filter(f, rows, cols, str(tmp_path / "sub"))
    sub = plink.PlinkFileset(d)
    ri = np.nonzero(rows)[0] if np.asarray(rows).dtype == bool else rows
    ci = np.nonzero(cols)[0] if np.asarray(cols).dtype == bool else cols
    assert sub.shape == (ri.size, ci.size)
    assert np.array_equal(sub.codes(), full[ci][:, ri])
    assert list(sub.fam["iid"]) == [f"ind{i}" for i in ri]
    assert list(sub.bim["snpid"]) == [f"snp{c}" for c in ci]
    # padding bits of the last byte are zero, as PLINK writes them
    if ri.size % 4:
        last = np.asarray(sub.data).reshape(ci.size, -1)[:, -1]
        assert np.all((last >> (2 * (ri.size % 4))) == 0)


def test_filter_rejects_bad_selection(tmp_path):
    f, s = _fileset(tmp_path)
    with pytest.raises(ValueError):
        plink.filter(f, np.ones(36, bool), np.arange(4), str(tmp_path / "bad"))
    with pytest.raises(ValueError):
        plink.filter(f, np.arange(37), np.array([20]), str(tmp_path / "bad"))


def test_payload_first_individuals_and_snps(tmp_path):
    f, s = _fileset(tmp_path)
    fs = plink.PlinkFileset(f)
    pl, bps, nsnp = fs.payload(13, 3)
    assert bps == 10 and nsnp == 6 and pl.size == 60
    with pytest.raises(ValueError):
        fs.payload(38, 3)
    with pytest.raises(ValueError):
        fs.payload(10, 11)


def test_q_p_files_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    q = np.asfortranarray(rng.dirichlet(np.ones(3), size=11).T)
    p = np.asfortranarray(rng.random((3, 4 * 5)))
    fq, fp = plink.save_fit(str(tmp_path / "run"), q, p)
    assert fq.endswith("run.3.Q") and fp.endswith("run.3.P")
    assert np.array_equal(plink.read_q(fq), q)
    assert np.array_equal(plink.read_p(fp), p)
    lines = open(fq).read().splitlines()
    assert len(lines) == 11 and len(lines[0].split()) == 3
