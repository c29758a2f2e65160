"""TNL binary matrix files on the host (no GPU): the library's reader accepts every file TNL's own
SparseMatrix::save wrote (tests/golden/tnlfile.npz, from the reference build: all matrix types,
both organizations, double and float) and reports its type, and it rejects damaged files with
TNL's deserialization messages or a metadata check before anything could reach the device."""
import os

import numpy as np
import pytest

import pyoracle as O
import tnl_files as TF

KIND_OF_VARIANT = {"CSRScalar": "CSRScalar", "CSRVector": "CSRVector", "CSRHybrid": "CSRHybrid",
                   "CSRLight": "CSRLight", "CSRAdaptive": "CSRAdaptive", "Ellpack": "Ellpack",
                   "SlicedEllpack": "SlicedEllpack", "ChunkedEllpack": "ChunkedEllpack", "BiEllpack": "BiEllpack"}


@pytest.fixture(scope="module")
def files():
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tnlfile.npz"))


def entries(files):
    for name in (str(n) for n in files["names"]):
        for variant in O.TNL_FILE_VARIANTS:
            for org in ("cm", "rm"):
                for dt in ("f64", "f32"):
                    key = f"{name}/{variant}/{org}/{dt}"
                    if key in files.files:
                        yield name, variant, org, dt, files[key].tobytes()


def layout_of(org):
    return "cuda" if org == "cm" else "host"


def write(tmp_path, data, name="m.tnl"):
    p = tmp_path / name
    p.write_bytes(data)
    return p


def test_reader_accepts_every_tnl_file(tnl, files, tmp_path):
    count = 0
    for name, variant, org, dt, data in entries(files):
        info = tnl.file_inspect(write(tmp_path, data), layout_of(org))
        f = TF.fields(data)
        rp = files[f"{name}/row_ptr"]
        assert info["format"] == KIND_OF_VARIANT[variant], (name, variant)
        assert info["dtype"] == dt
        assert (info["rows"], info["columns"]) == (len(rp) - 1, int(files[f"{name}/ncols"][0]))
        assert info["storage_size"] == f["values"][1]
        assert info["type"] == f["type"]
        count += 1
    assert count == 4 * (5 * 2 + 4 * 4)


def test_type_strings_name_the_kernel_and_real_type(tnl, files, tmp_path):
    data = files["tiny_5x5/CSRHybrid/cm/f32"].tobytes()
    t = tnl.file_inspect(write(tmp_path, data))["type"]
    assert t == ("Matrices::SparseMatrix< float, CSR< [any_device], int, TNL::Algorithms::Segments::"
                 "CSRHybridKernel<int, TNL::Devices::Host, 128> >, [any_device], int, General, [any_allocator], "
                 "[any_allocator] >")


def test_ellpack_organization_is_checked(tnl, files, tmp_path):
    # 70 rows: the row-major file is not aligned to 32, the column-major one is
    rm = write(tmp_path, files["ragged_70x60/Ellpack/rm/f64"].tobytes(), "rm.tnl")
    cm = write(tmp_path, files["ragged_70x60/Ellpack/cm/f64"].tobytes(), "cm.tnl")
    with pytest.raises(tnl.TnlHipError, match="alignedSize = 70, expected 96"):
        tnl.file_inspect(rm, "cuda")
    with pytest.raises(tnl.TnlHipError, match="alignedSize = 96, expected 70"):
        tnl.file_inspect(cm, "host")


def deserialization_error(tnl, path, layout="cuda", match=None):
    with pytest.raises(tnl.TnlHipError) as e:
        tnl.file_inspect(path, layout)
    assert "Failed to deserialize an object from the file" in str(e.value)
    if match:
        assert match in str(e.value), str(e.value)


def test_truncated_and_foreign_files_are_rejected(tnl, files, tmp_path):
    data = files["ragged_70x60/SlicedEllpack/cm/f64"].tobytes()
    for cut in (0, 3, 7, 40, 200, len(data) // 2, len(data) - 1):
        deserialization_error(tnl, write(tmp_path, data[:cut]), match="unexpected end of file")
    deserialization_error(tnl, write(tmp_path, b"TNLMX" + data[5:]), match="wrong magic number")
    deserialization_error(tnl, write(tmp_path, data + b"\0"), match="trailing bytes")
    # a file that is not a sparse matrix: the column-index array of the same file on its own
    f = TF.fields(data)
    off = f["values"][0] + 8 * f["values"][1]
    deserialization_error(tnl, write(tmp_path, data[off:]), match="object type does not match")


def retyped(data, old, new):
    f = TF.fields(data)
    t = f["type"].replace(old, new)
    enc = t.encode()
    return b"TNLMN" + len(enc).to_bytes(4, "little", signed=True) + enc + data[9 + len(f["type"]):]


def test_unsupported_types_are_refused(tnl, files, tmp_path):
    data = files["tiny_5x5/BiEllpack/cm/f64"].tobytes()
    assert tnl.file_inspect(write(tmp_path, retyped(data, "General", "Symmetric")))["symmetric"]
    for old, new, match in (("General", "Hermitian", "only General and Symmetric matrices"),
                            (", int, General", ", long int, General", "unsupported Index type"),
                            ("BiEllpack<", "SlicedEllpackX<", "unknown segments type")):
        with pytest.raises(tnl.TnlHipError, match=match):
            tnl.file_inspect(write(tmp_path, retyped(data, old, new)))
    with pytest.raises(tnl.TnlHipError, match="unsupported Real type"):
        tnl.file_inspect(write(tmp_path, retyped(data, "SparseMatrix< double", "SparseMatrix< long double")))


def damaged(data, field, index, value):
    f = TF.fields(data)
    off, n, eb = f[field]
    assert index < n
    return TF.poke_int(data, off + eb * index, value)


@pytest.mark.parametrize("variant,field,index,value,match", [
    ("CSRScalar", "offsets", 3, 1, "offsets decrease"),
    ("CSRScalar", "offsets", 0, 1, "offsets[0]"),
    ("CSRScalar", "columnIndexes", 5, 60, "column index 60"),
    ("CSRScalar", "columnIndexes", 5, -2, "column index -2"),
    ("Ellpack", "segmentsCount", 0, 69, "segmentsCount"),
    ("Ellpack", "segmentSize", 0, 1000, "stored values = 2400, expected 96000"),
    ("SlicedEllpack", "sliceOffsets", 1, 0, "sliceOffsets"),
    ("SlicedEllpack", "sliceSegmentSizes", 0, 3, "sliceOffsets"),
    ("ChunkedEllpack", "rowToChunkMapping", 3, 0, "chunks"),
    ("ChunkedEllpack", "rowToSliceMapping", 2, 1, "chunks"),
    ("ChunkedEllpack", "chunksToSegmentsMapping", 9, 0, "chunksToSegmentsMapping"),
    ("ChunkedEllpack", "numberOfSlices", 0, 2, "chunksToSegmentsMapping size"),
    ("ChunkedEllpack", "storageSize", 0, 1, "storageSize"),
    ("BiEllpack", "rowPermArray", 1, 0, "rowPermArray"),
    ("BiEllpack", "groupPointers", 1, 3, "group"),
    ("BiEllpack", "virtualRows", 0, 64, "virtualRows"),
])
def test_damaged_metadata_is_rejected(tnl, files, tmp_path, variant, field, index, value, match):
    data = files[f"ragged_70x60/{variant}/cm/f64"].tobytes()
    deserialization_error(tnl, write(tmp_path, damaged(data, field, index, value)), match=match)


def test_tnl_chunked_load_swaps_two_arrays(files, tmp_path):
    """ChunkedEllpack::load reads rowToSliceMapping where the view's save wrote
    chunksToSegmentsMapping (ChunkedEllpack.hpp:392-401 vs ChunkedEllpackView.hpp:375-376), so
    TNL's own save / load round trip changes the product of a multi-row-slice matrix; the
    library's reader follows the order of the writes."""
    if not O.ref_available():
        pytest.skip("TNL reference build not present")
    name = "ragged_70x60"
    rp, ci, v = files[f"{name}/row_ptr"], files[f"{name}/cols"], files[f"{name}/vals"]
    nc = int(files[f"{name}/ncols"][0])
    x = np.linspace(-1.0, 1.0, nc)
    _, _, y_converted = O.ref_convert_spmv("ChunkedEllpack", rp, ci, v, nc, x)
    for variant in ("SlicedEllpack", "BiEllpack"):
        p = write(tmp_path, files[f"{name}/{variant}/cm/f64"].tobytes(), f"{variant}.tnl")
        np.testing.assert_array_equal(O.ref_load_tnl(variant, p, x)[2], y_converted)
    p = write(tmp_path, files[f"{name}/ChunkedEllpack/cm/f64"].tobytes(), "chunked.tnl")
    # the swapped arrays send TNL's kernel through slice entries it never initialized, so the
    # product runs in a child process: it either ends abnormally or differs
    import subprocess
    import sys
    out = tmp_path / "y.npy"
    code = ("import sys, numpy as np; sys.path.insert(0, sys.argv[1]); import pyoracle as O; "
            "x = np.linspace(-1.0, 1.0, int(sys.argv[3])); np.save(sys.argv[4], O.ref_load_tnl('ChunkedEllpack', "
            "sys.argv[2], x)[2])")
    rc = subprocess.run([sys.executable, "-c", code, os.path.dirname(O.__file__), str(p), str(nc), str(out)],
                        timeout=120).returncode
    assert rc != 0 or np.abs(np.load(out) - y_converted).max() > 1.0


def test_fixtures_are_tnls_current_output(files, tmp_path):
    """The committed files equal what the reference build writes now (unused slice entries masked)."""
    if not O.ref_available():
        pytest.skip("TNL reference build not present")
    name = "ragged_70x60"
    rp, ci, v = files[f"{name}/row_ptr"], files[f"{name}/cols"], files[f"{name}/vals"]
    nc = int(files[f"{name}/ncols"][0])
    for variant in O.TNL_FILE_VARIANTS:
        p = tmp_path / f"{variant}.tnl"
        O.ref_save_tnl(variant, rp, ci, v, nc, p, column_major=False, single=True)
        org = "cm" if variant.startswith("CSR") else "rm"
        assert TF.masked(p.read_bytes()) == TF.masked(files[f"{name}/{variant}/{org}/f32"].tobytes()), variant
