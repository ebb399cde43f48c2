"""Phoebe's el-ph coupling file (prefix.phoebe.elph.hdf5), read and written
without libhdf5: the library's C++ reader / writer (through the C ABI, host
code only) against the oracle's independent Python decoding, pinned by the
reference's own fixture test/data/silicon.phoebe.elph.hdf5 (a libhdf5-written
format-1 file, copied to tests/golden/).
"""
import os
import shutil

import numpy as np
import pytest

import phoebe_amd as P
from oracle import h5_oracle as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "silicon.phoebe.elph.hdf5")
FIXTURE_SHA256 = "3ce7bad975d391a7c5ff0164047d96231cb013d884259bc6db228d26bdd16d22"


def _sha256(path):
    import hashlib
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def _random_coupling(rng, nw, nm, nRp, nRe):
    g = rng.standard_normal((nRe, nRp, nm, nw, nw)) + 1j * rng.standard_normal((nRe, nRp, nm, nw, nw))
    eR = rng.integers(-4, 5, (nRe, 3)).astype(float) * 5.1
    pR = rng.integers(-4, 5, (nRp, 3)).astype(float) * 5.1
    eD = rng.integers(1, 7, nRe).astype(float)
    pD = rng.integers(1, 7, nRp).astype(float)
    return g, eR, eD, pR, pD


# ---------------------------------------------------------------- oracle pinned by the fixture
def test_fixture_is_the_reference_file():
    assert _sha256(FIXTURE) == FIXTURE_SHA256


def test_oracle_reads_fixture_header():
    fi = H.read_structure(FIXTURE, strict=True)   # the libhdf5-written file passes the strict checks
    assert fi.superblock_version == 0 and (fi.leaf_k, fi.internal_k) == (4, 16)
    assert fi.eof == os.path.getsize(FIXTURE)
    assert sorted(fi.datasets) == sorted(
        ["fileFormat", "numElectrons", "numSpin", "numElBands", "numPhModes", "kMesh", "qMesh",
         "elBravaisVectors", "phBravaisVectors", "elDegeneracies", "phDegeneracies", "gWannier"])
    e = H.read_elph(FIXTURE)
    assert (e["fileFormat"], e["numElectrons"], e["numSpin"]) == (1, 8, 1)
    assert (e["nw"], e["nm"], e["nRe"], e["nRp"]) == (8, 6, 19, 19)
    assert e["kMesh"] == (2, 2, 2) and e["qMesh"] == (2, 2, 2)
    # HighFive's Eigen extents: (3, nR) matrices, (nR, 1) vectors, (1, N) coupling row
    assert fi.datasets["elBravaisVectors"].shape == (3, 19)
    assert fi.datasets["elDegeneracies"].shape == (19, 1)
    assert fi.datasets["gWannier"].shape == (1, 8 * 8 * 6 * 19 * 19)
    assert fi.datasets["gWannier"].dtype == np.complex128


def test_fixture_lattice_invariants():
    """What the physics fixes about the header: Wigner-Seitz weights 1/deg sum
    to the number of supercell points, and the vectors are Si fcc lattice
    vectors (a = 10.2 Bohr: integer multiples of a/2 with an even sum), the
    first one the origin with weight 1.  A transposed (3, nR) read breaks
    both the fcc test and the origin test."""
    e = H.read_elph(FIXTURE)
    for R, deg, mesh in ((e["elR"], e["elDeg"], e["kMesh"]), (e["phR"], e["phDeg"], e["qMesh"])):
        assert abs((1.0 / deg).sum() - np.prod(mesh)) < 1e-12
        n = R / 5.1
        assert np.allclose(n, np.round(n), atol=1e-9)
        assert np.all(np.round(n).sum(axis=1) % 2 == 0)
        assert np.array_equal(R[0], [0, 0, 0]) and deg[0] == 1
        # Wigner-Seitz degeneracy of a vector = the number of its images at equal length
        lens = np.round(np.linalg.norm(R, axis=1), 6)
        assert np.all(deg >= 1) and np.all(deg[lens == lens.max()] > 1)
    g = e["gW"]
    assert np.all(np.isfinite(g.view(float))) and np.abs(g).max() > 0


def test_oracle_strict_check_catches_planted_mistakes(tmp_path):
    """The strict structural checks are not vacuous: unsorted symbol-node
    entries and a wrong end-of-file address are both caught."""
    bad = tmp_path / "unsorted.h5"
    shutil.copy(FIXTURE, bad)
    fi = H.read_structure(FIXTURE)
    data = bytearray(open(bad, "rb").read())
    # the first symbol node: swap its first two 40-byte entries
    snod = data.index(b"SNOD")
    e0, e1 = snod + 8, snod + 48
    data[e0:e0 + 40], data[e1:e1 + 40] = data[e1:e1 + 40], data[e0:e0 + 40]
    open(bad, "wb").write(bytes(data))
    with pytest.raises(H.H5FormatError, match="sorted"):
        H.read_structure(str(bad), strict=True)
    short = tmp_path / "eof.h5"
    shutil.copy(FIXTURE, short)
    with open(short, "ab") as fh:
        fh.write(b"\0" * 8)
    with pytest.raises(H.H5FormatError, match="end-of-file"):
        H.read_structure(str(short), strict=True)
    assert fi.eof == os.path.getsize(FIXTURE)


# ---------------------------------------------------------------- library reader vs oracle
def test_library_reader_matches_oracle_on_fixture():
    o = H.read_elph(FIXTURE)
    with P.ElPhFile(FIXTURE) as f:
        h = f.header
        assert (h["fileFormat"], h["numElectrons"], h["numSpin"]) == (1, 8, 1)
        assert (h["numElBands"], h["numPhModes"], h["nRe"], h["nRp"]) == (8, 6, 19, 19)
        assert h["kMesh"] == o["kMesh"] and h["qMesh"] == o["qMesh"]
        eR, eD, pR, pD = f.bravais()
        for a, b in ((eR, o["elR"]), (eD, o["elDeg"]), (pR, o["phR"]), (pD, o["phDeg"])):
            assert np.array_equal(a, b)
        g = f.read_gw()
        assert g.shape == (19, 19, 6, 8, 8)
        assert np.array_equal(g.view(np.float64), o["gW"].view(np.float64))


@pytest.mark.parametrize("b,c", [(0, 1), (0, 19), (5, 7), (18, 1), (3, 0)])
def test_library_partial_reads(b, c):
    o = H.read_elph(FIXTURE)["gW"]
    with P.ElPhFile(FIXTURE) as f:
        assert np.array_equal(f.read_gw(b, c), o[b:b + c])


def test_library_read_argument_errors():
    with P.ElPhFile(FIXTURE) as f:
        with pytest.raises(P.PhxError, match="R_e range"):
            f.read_gw(18, 2)
        with pytest.raises(P.PhxError, match="R_e range"):
            f.read_gw(-1, 1)


def test_missing_and_foreign_files(tmp_path):
    missing = tmp_path / "absent.phoebe.elph.hdf5"
    with pytest.raises(P.PhxError, match="not found at"):
        P.ElPhFile(missing)
    junk = tmp_path / "junk.h5"
    junk.write_bytes(b"not an hdf5 file" * 100)
    with pytest.raises(P.PhxError, match="not an HDF5 file"):
        P.ElPhFile(junk)
    trunc = tmp_path / "truncated.h5"
    trunc.write_bytes(open(FIXTURE, "rb").read()[:4096])
    with pytest.raises(P.PhxError):
        P.ElPhFile(trunc)


def test_file_without_format_id_is_format_1(tmp_path):
    """Files that predate /fileFormat are format 1 (interaction_elph_parsing.cpp:651-663)."""
    old = tmp_path / "old.h5"
    data = bytearray(open(FIXTURE, "rb").read())
    i = data.index(b"fileFormat\0")
    data[i:i + 10] = b"fileFormaX"       # the dataset is no longer found by name
    old.write_bytes(bytes(data))
    with P.ElPhFile(old) as f:
        assert f.header["fileFormat"] == 1
        assert np.array_equal(f.read_gw(), H.read_elph(FIXTURE)["gW"])


# ---------------------------------------------------------------- writer
@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("dims", [(1, 1, 1, 1), (3, 2, 5, 4), (4, 6, 7, 13)])
def test_write_read_round_trip(tmp_path, fmt, dims):
    nw, nm, nRp, nRe = dims
    rng = np.random.default_rng(sum(dims) + 10 * fmt)
    g, eR, eD, pR, pD = _random_coupling(rng, nw, nm, nRp, nRe)
    path = tmp_path / f"rt{fmt}.phoebe.elph.hdf5"
    P.write_elph_file(path, g, eR, eD, pR, pD, numElectrons=7, numSpin=2, kMesh=(3, 4, 5),
                      qMesh=(2, 3, 4), fileFormat=fmt)
    fi = H.read_structure(str(path), strict=True)
    names = set(fi.datasets)
    if fmt == 1:
        assert "gWannier" in names and fi.datasets["gWannier"].shape == (1, g.size)
    else:
        assert "gWannier" not in names
        for e in range(nRe):
            assert fi.datasets[f"gWannier_{e}"].shape == (nw * nw * nm * nRp, 1)
    o = H.read_elph(str(path))
    assert (o["fileFormat"], o["numElectrons"], o["numSpin"]) == (fmt, 7, 2)
    assert o["kMesh"] == (3, 4, 5) and o["qMesh"] == (2, 3, 4)
    assert np.array_equal(o["gW"].view(float), g.view(float))
    assert np.array_equal(o["elR"], eR) and np.array_equal(o["phDeg"], pD)
    with P.ElPhFile(path) as f:
        assert f.header["fileFormat"] == fmt
        assert np.array_equal(f.read_gw().view(float), g.view(float))
        if nRe > 2:
            assert np.array_equal(f.read_gw(1, nRe - 2), g[1:nRe - 1])


def test_rewrite_of_fixture_matches_libhdf5_messages(tmp_path):
    """Writing the fixture's content back (format 1) reproduces, byte for byte,
    the dataspace and datatype messages libhdf5 wrote for every dataset, and
    the fill-value message of the header datasets (gWannier was written through
    MPI-IO, which allocates early)."""
    o = H.read_elph(FIXTURE)
    out = tmp_path / "rewrite.phoebe.elph.hdf5"
    P.write_elph_file(out, o["gW"], o["elR"], o["elDeg"], o["phR"], o["phDeg"],
                      numElectrons=o["numElectrons"], numSpin=o["numSpin"], kMesh=o["kMesh"],
                      qMesh=o["qMesh"], fileFormat=1)
    ref = H.read_structure(FIXTURE, strict=True)
    new = H.read_structure(str(out), strict=True)
    assert (new.superblock_version, new.sizeof_offsets, new.sizeof_lengths) == (0, 8, 8)
    assert (new.leaf_k, new.internal_k) == (ref.leaf_k, ref.internal_k)
    assert sorted(new.datasets) == sorted(ref.datasets)
    for name, d in ref.datasets.items():
        n = new.datasets[name]
        assert n.messages[0x01] == d.messages[0x01], name      # dataspace
        assert n.messages[0x03] == d.messages[0x03], name      # datatype
        assert n.messages[0x08][:2] == d.messages[0x08][:2]     # layout v3, contiguous
        if name != "gWannier":
            assert n.messages[0x05] == d.messages[0x05], name  # fill value
        assert n.shape == d.shape and n.nbytes == d.nbytes
        assert np.array_equal(H.read_dataset(str(out), n).view(np.uint8),
                              H.read_dataset(FIXTURE, d).view(np.uint8))


def test_many_datasets_grow_the_group_b_tree(tmp_path):
    """Format 2 with 300 R_e: 311 datasets need 39 symbol nodes, more than one
    default B-tree node holds (2K = 32); the writer widens K and the file
    still passes the strict checks."""
    rng = np.random.default_rng(3)
    g, eR, eD, pR, pD = _random_coupling(rng, 1, 1, 2, 300)
    path = tmp_path / "wide.phoebe.elph.hdf5"
    P.write_elph_file(path, g, eR, eD, pR, pD, numElectrons=1, fileFormat=2)
    fi = H.read_structure(str(path), strict=True)
    assert len(fi.datasets) == 311 and fi.internal_k >= 20 and len(fi.snod_entries) == 39
    with P.ElPhFile(path) as f:
        assert np.array_equal(f.read_gw(), g)


def test_writer_argument_errors(tmp_path):
    rng = np.random.default_rng(4)
    g, eR, eD, pR, pD = _random_coupling(rng, 2, 3, 2, 2)
    with pytest.raises(P.PhxError):
        P.write_elph_file(tmp_path / "x.h5", g, eR, eD, pR, pD, numElectrons=1, fileFormat=3)
    with pytest.raises(P.PhxError, match="inconsistent"):
        P.write_elph_file(tmp_path / "x.h5", g, eR[:1], eD, pR, pD, numElectrons=1)
    with pytest.raises(P.PhxError, match="cannot create"):
        P.write_elph_file(tmp_path / "no" / "such" / "dir.h5", g, eR, eD, pR, pD, numElectrons=1)
