"""The integral file (libimago_h5.so) without a GPU: the layout of the
reference's SCF file (hdf5SCF.F90:60-230, hdf5SCFIntg.F90:105-525) as
h5dump, an independent reader, reports it; bit-exact round trips of packed
matrices; the status attributes a resumed run reads; argument errors."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from imago_amd import h5store

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(h5store.H5LIB),
                                reason="libimago_h5.so not built (no HDF5 C library)")


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "imago_h5.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(imago_h5_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", h5store.H5LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (imago_h5_\w+)", out))
    assert len(declared_symbols()) == 7
    assert set(declared_symbols()) <= exported


def test_names_follow_the_reference():
    assert h5store.file_name() == "gs_scf-fb.hdf5"
    assert h5store.file_name("mb", "1s") == "1s_scf-mb.hdf5"
    assert h5store.kpoint_group((4, 4, 4)) == "00004_00004_00004"


def _dump_header(path):
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    return subprocess.run([H5DUMP, "-H", "-p", str(path)], capture_output=True, text=True,
                          check=True).stdout


def test_layout_as_an_independent_reader_sees_it(tmp_path):
    path = tmp_path / "gs_scf-fb.hdf5"
    nk, n, pot = 2, 5, 3
    with h5store.IntegralFile(str(path), "00002_00002_00002", nk, n, pot) as f:
        assert f.created
    text = _dump_header(path)
    base = '/00002_00002_00002/atomIntgGroup'
    # every group of initSCFIntegralHDF5 with coreDim = 0
    for g in ("atomOverlap", "atomKEOverlap", "atomMVOverlap", "atomNPOverlap",
              "atomDMOverlap", "atomMMOverlap", "atomKOverlap", "atomKOverlapPlusG",
              "atomPotOverlap"):
        assert 'GROUP "%s"' % g in text, g
    assert "atomOverlapCV" not in text
    for x in ("0000001", "0000002", "0000003"):
        assert text.count('GROUP "%s"' % x) >= 4
    # packed datasets: C shape (n(n+1)/2, 2), chunked whole, deflate 1
    npk = n * (n + 1) // 2
    m = re.search(r'GROUP "atomOverlap" \{(.*?)\n\s{9}\}', text, re.S)
    assert m, text[:2000]
    block = m.group(1)
    assert block.count('DATASET "') == nk
    assert "SIMPLE { ( %d, 2 ) / ( %d, 2 ) }" % (npk, npk) in block
    assert "CHUNKED ( %d, 2 )" % npk in block
    assert "COMPRESSION DEFLATE { LEVEL 1 }" in block
    assert 'ATTRIBUTE "status"' in block
    # the K overlaps hold full n x n real and imaginary parts
    assert "SIMPLE { ( %d, %d ) / ( %d, %d ) }" % (n, n, n, n) in text
    assert 'DATASET "real0000002"' in text and 'DATASET "imag0000001"' in text
    # one subgroup per potential term, each with a status
    pm = re.search(r'GROUP "atomPotOverlap" \{(.*)', text, re.S).group(1)
    for j in range(1, pot + 1):
        assert 'GROUP "%07d"' % j in pm
    assert base.split("/")[1] in text


def test_round_trip_and_status(tmp_path):
    path = str(tmp_path / "gs_scf-fb.hdf5")
    rng = np.random.default_rng(3)
    nk, n, pot = 3, 7, 4
    npk = n * (n + 1) // 2
    data = {}
    with h5store.IntegralFile(path, "00003_00001_00001", nk, n, pot) as f:
        for kind in ("ol", "ke", "np", "ep", "mv"):
            for term in (range(pot) if kind == "ep" else [0]):
                assert f.status(kind, term) == 0
                for k in range(nk):
                    a = rng.standard_normal(2 * npk)
                    f.write(kind, a, term, k)
                    data[kind, term, k] = a
        f.set_status("ol", 0, 1)
        f.set_status("ep", 2, 1)
    with h5store.IntegralFile(path, "00003_00001_00001", nk, n, pot) as f:
        assert not f.created
        for (kind, term, k), a in data.items():
            assert np.array_equal(f.read(kind, term, k), a)
        assert f.status("ol") == 1 and f.status("ke") == 0
        assert [f.status("ep", j) for j in range(pot)] == [0, 0, 1, 0]
    # a second k-point set in the same file gets groups of its own
    with h5store.IntegralFile(path, "00001_00001_00001", 1, n, pot, gamma=True) as f:
        assert f.created and f.status("ol") == 0
        g = rng.standard_normal(npk)
        f.write("ol", g)
        assert np.array_equal(f.read("ol"), g)


def test_errors_name_the_problem(tmp_path):
    path = str(tmp_path / "x.hdf5")
    with h5store.IntegralFile(path, "00001_00001_00002", 2, 4, 2) as f:
        with pytest.raises(RuntimeError, match="no such matrix"):
            f.read("ol", 0, 2)
        with pytest.raises(RuntimeError, match="no such matrix"):
            f.read("ep", 2, 0)
        with pytest.raises(ValueError):
            f.write("ol", np.zeros(3))
    # the same group opened with another valeDim is refused on access
    with h5store.IntegralFile(path, "00001_00001_00002", 2, 5, 2) as f:
        with pytest.raises(RuntimeError, match="unexpected shape"):
            f.read("ol", 0, 0)
