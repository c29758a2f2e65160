"""TNL binary matrix files through the device.  A matrix built on the GPU and saved gives the
bytes TNL's own SparseMatrix::save writes for the same matrix (tests/golden/tnlfile.npz, made
from the reference build), in both layouts; loading TNL's files gives, element for element, the
matrix the library builds from the same CSR arrays, and the same product."""
import os

import numpy as np
import pytest

import tnl_files as TF

pytestmark = pytest.mark.gpu

VARIANTS = ["CSRScalar", "CSRVector", "CSRHybrid", "CSRLight", "CSRAdaptive", "Ellpack", "SlicedEllpack",
            "ChunkedEllpack", "BiEllpack"]
METADATA = {"CSR": ["offsets"], "Ellpack": ["ellpack"], "SlicedEllpack": ["sliceOffsets", "sliceSegmentSizes", "sliced"],
            "ChunkedEllpack": ["rowToChunkMapping", "rowToSliceMapping", "chunksToSegmentsMapping", "rowPointers",
                               "slices", "chunked"],
            "BiEllpack": ["rowPermArray", "groupPointers", "biellpack"]}


@pytest.fixture(scope="module")
def files():
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tnlfile.npz"))


def names(files):
    return [str(n) for n in files["names"]]


def case(files, name):
    return (files[f"{name}/row_ptr"], files[f"{name}/cols"], files[f"{name}/vals"], int(files[f"{name}/ncols"][0]))


def tnl_bytes(files, name, variant, layout, dt):
    """TNL's file: the row-major one for the host layout; for the CUDA layout the column-major one,
    whose CSR kernel names TNL's CUDA matrices write with Devices::Cuda."""
    org = "cm" if layout == "cuda" or variant.startswith("CSR") else "rm"
    data = files[f"{name}/{variant}/{org}/{dt}"].tobytes()
    if layout == "cuda" and variant.startswith("CSR"):
        t = TF.fields(data)["type"]
        enc = t.replace("TNL::Devices::Host", "TNL::Devices::Cuda").encode()
        data = b"TNLMN" + len(enc).to_bytes(4, "little", signed=True) + enc + data[9 + len(t):]
    return data


def family(variant):
    return "CSR" if variant.startswith("CSR") else variant


def product(tnl, A, x):
    import torch
    y = torch.zeros(A.getRows(), dtype=A.vector_dtype, device="cuda:0")
    A.vectorProduct(x.to(A.vector_dtype), y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("layout", ["cuda", "host"])
def test_save_writes_tnls_bytes(tnl, dev, files, tmp_path, layout):
    for name in names(files):
        rp, ci, v, nc = case(files, name)
        for variant in VARIANTS:
            for dt in ("f64", "f32"):
                print(f"save {name} {variant} {dt} {layout}", flush=True)
                A = tnl.SparseMatrix.from_csr(rp, ci, v, columns=nc, format=variant, dtype=dt)
                p = tmp_path / "m.tnl"
                A.save(p, layout)
                assert TF.masked(p.read_bytes()) == TF.masked(tnl_bytes(files, name, variant, layout, dt)), \
                    (name, variant, dt, layout)


@pytest.mark.parametrize("layout", ["cuda", "host"])
def test_load_gives_the_matrix_built_from_csr(tnl, dev, files, tmp_path, layout):
    import torch
    for name in names(files):
        rp, ci, v, nc = case(files, name)
        x = torch.linspace(-1.0, 1.0, max(nc, 1), dtype=torch.float64, device=dev)[:nc]
        for variant in VARIANTS:
            for dt in ("f64", "f32"):
                print(f"load {name} {variant} {dt} {layout}", flush=True)
                p = tmp_path / "m.tnl"
                p.write_bytes(tnl_bytes(files, name, variant, layout, dt))
                B = tnl.SparseMatrix.load(p, layout=layout)
                A = tnl.SparseMatrix.from_csr(rp, ci, v, columns=nc, format=variant, dtype=dt)
                key = (name, variant, dt, layout)
                assert (B.format, B.value_dtype, B.getRows(), B.getColumns()) == (variant, dt, A.getRows(), nc), key
                for field in ("storage_size", "allocated", "nonzeros", "lanes_per_row"):
                    assert getattr(B.info, field) == getattr(A.info, field), (key, field)
                for got, want in zip(B.storage(), A.storage()):
                    np.testing.assert_array_equal(got, want, err_msg=str(key))
                for m in METADATA[family(variant)]:
                    np.testing.assert_array_equal(B.metadata(m), A.metadata(m), err_msg=str((key, m)))
                if variant == "CSRAdaptive":
                    np.testing.assert_array_equal(B.metadata("adaptiveBlocks"), A.metadata("adaptiveBlocks"))
                np.testing.assert_array_equal(product(tnl, B, x), product(tnl, A, x), err_msg=str(key))


def test_round_trip_and_type_checks(tnl, dev, files, tmp_path):
    import torch
    rp, ci, v, nc = case(files, "many_slices_600x400")
    x = torch.linspace(-1.0, 1.0, nc, dtype=torch.float64, device=dev)
    for variant in VARIANTS:
        A = tnl.SparseMatrix.from_csr(rp, ci, v, columns=nc, format=variant, dtype="f64")
        for layout in ("cuda", "host"):
            p = tmp_path / f"{variant}_{layout}.tnl"
            A.save(p, layout)
            B = tnl.SparseMatrix.load(p, format=variant, dtype="f64", layout=layout)
            np.testing.assert_array_equal(product(tnl, B, x), product(tnl, A, x), err_msg=f"{variant} {layout}")
    # any CSR kernel reads a CSR file; another segments type or Real type is refused
    p = tmp_path / "csr.tnl"
    tnl.SparseMatrix.from_csr(rp, ci, v, columns=nc, format="CSRScalar").save(p)
    B = tnl.SparseMatrix.load(p, format="CSRAdaptive")
    np.testing.assert_array_equal(product(tnl, B, x), product(tnl, tnl.SparseMatrix.from_csr(
        rp, ci, v, columns=nc, format="CSRAdaptive"), x))
    with pytest.raises(tnl.TnlHipError, match="object type does not match"):
        tnl.SparseMatrix.load(p, format="Ellpack")
    with pytest.raises(tnl.TnlHipError, match="object type does not match"):
        tnl.SparseMatrix.load(p, dtype="f32")
    # a SymmetricMatrix keeps its type through the file ("Symmetric" in the type string)
    n = 300
    lrp = np.concatenate([[0, 1], np.arange(3, 2 * n, 2)]).astype(np.int32)
    lc = np.concatenate([[0], np.stack([np.arange(n - 1), np.arange(1, n)], 1).ravel()]).astype(np.int32)
    lv = np.concatenate([[4.0], np.tile([-1.0, 4.0], n - 1)])
    xs = torch.linspace(-1.0, 1.0, n, dtype=torch.float64, device=dev)
    S = tnl.SparseMatrix.symmetric_from_csr(lrp, lc, lv, format="BiEllpack")
    ps = tmp_path / "sym.tnl"
    S.save(ps)
    assert tnl.file_inspect(ps)["symmetric"] and "Symmetric" in tnl.file_inspect(ps)["type"]
    T = tnl.SparseMatrix.load(ps)
    assert T.symmetric
    np.testing.assert_array_equal(product(tnl, T, xs), product(tnl, S, xs))
    with pytest.raises(tnl.TnlHipError, match="double or float values only"):
        tnl.SparseMatrix.from_csr(rp, ci, v, columns=nc, format="SlicedEllpack", dtype="bf16").save(p)


def test_chunked_file_with_padding_inside_a_chunk(tnl, dev, tmp_path):
    """TNL's save writes a chunk's padding after its elements, and the ChunkedEllpack kernel stops
    a wavefront once every lane's chunk has reached padding.  A file whose chunk holds padding
    before an element (slot 7 emptied, slots 8-11 kept; never written by TNL's save nor by this
    library) is loaded with the full walk, and the product skips the padding slot and counts the
    elements after it, as every kernel here does with a stored -1 (TNL's Host traversal would
    end the row at it)."""
    import torch
    n = 200
    lens = np.full(n, 5)
    lens[0] = 1000  # the long row sets the slice's chunk size (18)
    lens[150] = 12
    rng = np.random.default_rng(4)
    ncols = 2000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(ncols, k, replace=False)) for k in lens]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(cols))
    A = tnl.SparseMatrix.from_csr(rp, cols, vals, columns=ncols, format="ChunkedEllpack", dtype="f64")
    p = tmp_path / "chunked.tnl"
    A.save(p)
    data = bytearray(p.read_bytes())
    f = TF.fields(data)
    r2c = TF.read(data, f, "rowToChunkMapping")
    slices = TF.read(data, f, "slices").reshape(-1, 4)
    assert slices[0, 0] == n, slices[0]  # one slice holds every row
    chunk_size, pointer = int(slices[0, 1]), int(slices[0, 3])
    c150 = int(r2c[149])
    # row 150 owns one chunk, in a wavefront (chunks 192-255) whose other rows end before slot 7
    assert r2c[150] - r2c[149] == 1 and chunk_size >= 12 and c150 >= 192, (r2c[148:151], chunk_size)
    stored = TF.read(data, f, "columnIndexes").copy()
    g = pointer + c150 + 256 * 7  # slot 7 of row 150's chunk
    assert stored[g] >= 0 and stored[g + 256] >= 0
    removed_col, removed_val = int(stored[g]), float(TF.read(data, f, "values")[g])
    stored[g] = -1
    off = f["columnIndexes"][0]
    data[off:off + 4 * len(stored)] = stored.astype(np.int32).tobytes()
    p.write_bytes(bytes(data))
    B = tnl.SparseMatrix.load(p)
    x = torch.linspace(-1.0, 1.0, ncols, dtype=torch.float64, device=dev)
    got = product(tnl, B, x)
    want = product(tnl, A, x)
    want[150] -= removed_val * float(x[removed_col])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
