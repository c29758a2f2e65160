"""Field map of TNL binary matrix files, for the tests: where each field of a file written by
SparseMatrix::save lies (Object.hpp:17-75, SparseMatrixView.hpp:911-920 and the segments' save
methods), so that tests can damage one field or mask bytes TNL leaves uninitialized."""
import struct

import numpy as np

# segments fields in file order: "i" an int32, "a" an int array, "s" the ChunkedEllpack slice-info array
SEGMENT_FIELDS = {
    "CSR": [("offsets", "a")],
    "Ellpack": [("segmentSize", "i"), ("segmentsCount", "i"), ("alignedSize", "i")],
    "SlicedEllpack": [("size", "i"), ("alignedSize", "i"), ("segmentsCount", "i"), ("sliceOffsets", "a"),
                      ("sliceSegmentSizes", "a")],
    "ChunkedEllpack": [("size", "i"), ("storageSize", "i"), ("chunksInSlice", "i"), ("desiredChunkSize", "i"),
                       ("rowToChunkMapping", "a"), ("chunksToSegmentsMapping", "a"), ("rowToSliceMapping", "a"),
                       ("rowPointers", "a"), ("slices", "s"), ("numberOfSlices", "i")],
    "BiEllpack": [("size", "i"), ("storageSize", "i"), ("virtualRows", "i"), ("rowPermArray", "a"),
                  ("groupPointers", "a")],
}


def fields(data):
    """{name: (byte offset of the elements, element count, element bytes)} plus "type"."""
    b = bytes(data)
    pos = 0

    def object_type():
        nonlocal pos
        assert b[pos:pos + 5] == b"TNLMN"
        (n,) = struct.unpack_from("<i", b, pos + 5)
        t = b[pos + 9:pos + 9 + n].decode()
        pos += 9 + n
        return t

    def int32(name):
        nonlocal pos
        out[name] = (pos, 1, 4)
        pos += 4

    def array(name, eb):
        nonlocal pos
        object_type()
        (n,) = struct.unpack_from("<i", b, pos)
        out[name + ".size"] = (pos, 1, 4)
        out[name] = (pos + 4, n, eb)
        pos += 4 + n * eb

    out = {}
    out["type"] = object_type()
    int32("rows")
    int32("columns")
    array("values", 8 if out["type"].startswith("Matrices::SparseMatrix< double") else 4)
    array("columnIndexes", 4)
    seg = out["type"].split(", ", 1)[1].split("<", 1)[0]
    for name, kind in SEGMENT_FIELDS[seg]:
        if kind == "i":
            int32(name)
        else:
            array(name, 16 if kind == "s" else 4)
    assert pos == len(b)
    return out


def read(data, f, name):
    off, n, eb = f[name]
    return np.frombuffer(bytes(data)[off:off + n * eb], np.float64 if eb == 8 else np.int32 if eb == 4 else np.int32)


def poke_int(data, offset, value):
    """A copy of the file with the int32 at offset replaced."""
    out = bytearray(data)
    out[offset:offset + 4] = struct.pack("<i", value)
    return bytes(out)


def masked(data):
    """The file with the unused ChunkedEllpack slice entries (allocated but never written by TNL,
    ChunkedEllpack.hpp:198) set to zero."""
    f = fields(data)
    if "slices" not in f:
        return bytes(data)
    off, n, _ = f["slices"]
    (used,) = struct.unpack_from("<i", bytes(data), f["numberOfSlices"][0])
    out = bytearray(data)
    out[off + 16 * used:off + 16 * n] = bytes(16 * (n - used))
    return bytes(out)
