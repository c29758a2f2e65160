"""Matrix Market files (host side, no GPU): the library's reader gives TNL's MatrixReader result
(capacities, padding, setElement semantics) or its error, and the writer writes TNL's
MatrixWriter text, against fixtures made by TNL's own reader / writer and, where the reference
build is present, against it live."""
import numpy as np
import pytest

import pyoracle as O


@pytest.fixture(scope="module")
def mtx():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mtx.npz"))


def cases(d):
    return [str(n) for n in d["names"]]


def write(tmp_path, name, text):
    p = tmp_path / f"{name}.mtx"
    p.write_text(text)
    return p


def test_reader_matches_tnl_fixtures(tnl, mtx, tmp_path):
    for name in cases(mtx):
        path = write(tmp_path, name, str(mtx[f"{name}/text"]))
        if f"{name}/error" in mtx.files:
            with pytest.raises(tnl.TnlHipError):
                tnl.mtx_read_csr(path)
            continue
        rows, cols, rp, ci, v = tnl.mtx_read_csr(path)
        assert [rows, cols] == list(mtx[f"{name}/dims"]), name
        np.testing.assert_array_equal(rp, mtx[f"{name}/row_ptr"], err_msg=name)
        np.testing.assert_array_equal(ci, mtx[f"{name}/cols"], err_msg=name)
        np.testing.assert_array_equal(v, mtx[f"{name}/vals"], err_msg=name)


def test_writer_matches_tnl_fixtures(tnl, mtx, tmp_path):
    for name in cases(mtx):
        if f"{name}/error" in mtx.files:
            continue
        rows, cols = (int(x) for x in mtx[f"{name}/dims"])
        out = tmp_path / f"{name}.out.mtx"
        tnl.mtx_write_csr(out, rows, cols, mtx[f"{name}/row_ptr"], mtx[f"{name}/cols"], mtx[f"{name}/vals"])
        assert out.read_text() == str(mtx[f"{name}/written"]), name


def test_semantics_by_hand(tnl, tmp_path):
    """Capacities count repeated entries, the last repeated value wins, explicit zeros are
    stored, a symmetric file fills both triangles (MatrixReader.hpp readMtx, SparseMatrixView
    addElement)."""
    p = write(tmp_path, "h", "%%MatrixMarket matrix coordinate real symmetric\n3 3 5\n2 1 1\n2 1 7\n"
                             "3 3 0\n1 1 2\n3 1 -4\n")
    rows, cols, rp, ci, v = tnl.mtx_read_csr(p)
    assert (rows, cols) == (3, 3)
    # row 0: (0,0)=2, (0,1)=7 mirrored twice, (0,2)=-4 -> capacity 4; row 1: (1,0) twice -> 2
    np.testing.assert_array_equal(rp, [0, 4, 6, 8])
    np.testing.assert_array_equal(ci, [0, 1, 2, -1, 0, -1, 0, 2])
    np.testing.assert_array_equal(v, [2, 7, -4, 0, 7, 0, -4, 0])


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reader_and_writer_against_live_tnl(tnl, tmp_path):
    rng = np.random.default_rng(3)
    for trial in range(5):
        n, m = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        sym = trial % 2 == 1 and n == m
        lines = []
        for _ in range(int(rng.integers(0, 200))):
            r, c = int(rng.integers(1, n + 1)), int(rng.integers(1, m + 1))
            if sym and c > r:
                r, c = c, r
            val = 0.0 if rng.random() < 0.1 else rng.standard_normal()
            lines.append(f"{r} {c} {val:.17g}")
        kind = "symmetric" if sym else "general"
        p = write(tmp_path, f"t{trial}", f"%%MatrixMarket matrix coordinate real {kind}\n{n} {m} {len(lines)}\n" +
                  "\n".join(lines) + "\n")
        try:
            want = O.ref_read_mtx(p)
        except ValueError:
            with pytest.raises(tnl.TnlHipError):
                tnl.mtx_read_csr(p)
            continue
        got = tnl.mtx_read_csr(p)
        assert got[:2] == want[:2]
        for a, b in zip(got[2:], want[2:]):
            np.testing.assert_array_equal(a, b)
        o1, o2 = tmp_path / "ours.mtx", tmp_path / "tnl.mtx"
        tnl.mtx_write_csr(o1, *got)
        O.ref_write_mtx(o2, *want)
        assert o1.read_text() == o2.read_text()
