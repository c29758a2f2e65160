"""The C-ABI library on the host: it loads, exports every function include/tnlhip.h declares,
and its host-side segments metadata is TNL's, bit for bit (no GPU needed)."""
import os
import re

import numpy as np
import pytest

import pyoracle as O
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORMAT_OF_KIND = {"CSR": "CSRScalar", "Ellpack": "Ellpack", "SlicedEllpack": "SlicedEllpack",
                  "ChunkedEllpack": "ChunkedEllpack", "BiEllpack": "BiEllpack"}


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tnlhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tnlhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(tnl):
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(tnl.lib, name), name
    assert sorted(tnl.EXPORTS) == names


def test_status_strings(tnl):
    assert tnl.lib.tnlhip_status_string(0) == b"success"
    assert tnl.lib.tnlhip_status_string(6) == b"solver breakdown"


@pytest.mark.parametrize("kind", list(FORMAT_OF_KIND))
def test_host_metadata_matches_tnl_fixtures(tnl, golden, kind):
    d = golden["segments"]
    cases = sorted({k.split("/")[0] for k in d.files})
    for case in cases:
        caps = np.diff(d[f"{case}/row_ptr"]).astype(np.int32)
        for name in O.METADATA_NAMES[O.KINDS[kind]]:
            np.testing.assert_array_equal(tnl.segments_metadata(FORMAT_OF_KIND[kind], caps, name),
                                          d[f"{case}/{kind}/col/{name}"], err_msg=f"{case} {name}")
        assert tnl.segments_metadata(FORMAT_OF_KIND[kind], caps, "storage") == int(d[f"{case}/{kind}/col/storage"][0])


@pytest.mark.parametrize("kind", list(FORMAT_OF_KIND))
def test_host_metadata_at_full_size(tnl, kind):
    """160^3 Laplacian (the SpMV benchmark matrix) and a 200k irregular matrix: bit-exact
    against the oracle."""
    for rp in (W.laplacian_3d(160)[0], W.irregular(200000, seed=3)[0]):
        caps = np.diff(rp).astype(np.int32)
        s = O.Segments(kind, caps)
        for name in O.METADATA_NAMES[O.KINDS[kind]]:
            np.testing.assert_array_equal(tnl.segments_metadata(FORMAT_OF_KIND[kind], caps, name), s.get(name))
        assert tnl.segments_metadata(FORMAT_OF_KIND[kind], caps, "storage") == s.storage_size()


def test_edge_capacities(tnl):
    """Empty matrix, all-empty rows, a single row, one-row last strip, rows longer than a slice."""
    for caps in ([], [0], [0] * 40, [5], [3] * 33, [0, 7, 0, 4096, 1, 0], [300] * 3 + [1] * 70):
        caps = np.asarray(caps, np.int32)
        for kind, fmt in FORMAT_OF_KIND.items():
            s = O.Segments(kind, caps)
            for name in O.METADATA_NAMES[O.KINDS[kind]]:
                np.testing.assert_array_equal(tnl.segments_metadata(fmt, caps, name), s.get(name),
                                              err_msg=f"{kind} {name} {caps[:5]}")


def test_adaptive_blocks_follow_the_tnl_split(tnl):
    """With wave64 parameters the product's CSR-Adaptive split is the same greedy algorithm
    (CSRAdaptiveKernel::initValueSize) as the oracle's with the same parameters."""
    rp, _, _ = W.irregular(20000, seed=4)
    caps = np.diff(rp).astype(np.int32)
    got = tnl.segments_metadata("CSRAdaptive", caps, "adaptiveBlocks").reshape(-1, 6)
    want = O.csr_adaptive_blocks(rp, 8, warp_size=64, block_size=256, streamed_shared_bytes=4 * 4096,
                                 max_adaptive_elements=1024)
    np.testing.assert_array_equal(got, want)


def test_solver_argument_checks(tnl):
    """tnlhip_solver_set_params refuses what would mis-size the solver's arrays or name an unknown
    preconditioner, and the history getters refuse a negative capacity; no GPU is touched."""
    import ctypes as C
    lib = tnl.lib
    h = C.c_void_p()
    assert lib.tnlhip_solver_create(tnl.METHODS["gmres"], C.byref(h)) == 0
    try:
        def status(**kw):
            p = tnl.SolverParams()
            lib.tnlhip_solver_params_default(C.byref(p))
            for k, v in kw.items():
                setattr(p, k, v)
            return lib.tnlhip_solver_set_params(h, C.byref(p))
        assert status() == 0
        assert status(preconditioner=3, poly_degree=0) == 0
        assert status(preconditioner=3, poly_degree=32) == 0
        assert status(preconditioner=4) == 1
        assert status(preconditioner=3, poly_degree=33) == 1
        assert status(preconditioner=3, poly_degree=-1) == 1
        assert status(restarting_min=20, restarting_max=10) == 1
        assert status(restarting_min=0, restarting_max=10) == 1
        assert status(restarting_min=10, restarting_max=32) == 1
        assert status(restarting_step_min=4, restarting_step_max=3) == 1
        assert status(bicgstab_ell=0) == 1
        buf = (C.c_double * 4)()
        assert lib.tnlhip_solver_get_history(h, C.cast(buf, C.c_void_p), -1) == -1
        ibuf = (C.c_int32 * 4)()
        assert lib.tnlhip_solver_get_history_iterations(h, C.cast(ibuf, C.c_void_p), -1) == -1
        assert lib.tnlhip_solver_get_history(h, None, 0) == 0
    finally:
        lib.tnlhip_solver_destroy(h)


def test_noalloc_multidot_refuses_before_touching_the_gpu(tnl):
    """tnlhip_multidot_noalloc rejects a column count its fixed device array cannot hold, and bad
    shapes, with a clear error and without a device call."""
    import ctypes as C
    lib = tnl.lib
    dummy = C.c_void_p(16)  # never dereferenced: every case below fails its argument check
    for count in (0, 33, 1000):
        assert lib.tnlhip_multidot_noalloc(0, 64, count, dummy, 64, dummy, dummy, None) == 1
        assert b"1 to 32 columns" in lib.tnlhip_last_error()
    assert lib.tnlhip_multidot_noalloc(0, 64, 4, dummy, 32, dummy, dummy, None) == 1  # ld < n
    assert lib.tnlhip_multidot_noalloc(0, 64, 4, None, 64, dummy, dummy, None) == 1
    assert lib.tnlhip_multidot_noalloc(2, 64, 4, dummy, 64, dummy, dummy, None) == 1  # BF16 vectors
