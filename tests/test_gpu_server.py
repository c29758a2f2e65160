"""The solve server on the GPU: FP64-refined results against a dense FP64 solve, and the same
bits for a request whether it runs alone or among others admitted and retired around it."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def spmv(rp, c, v, X, shift):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    out = np.zeros_like(X)
    np.add.at(out, rows, v[:, None] * X[c])
    return out + shift * X


def rel_residuals(rp, c, v, shift, B, X):
    R = B - spmv(rp, c, v, X, shift)
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


def systems():
    return {"l2_40": W.laplacian_2d(40), "l2_23": W.laplacian_2d(23), "l3_12": W.laplacian_3d(12),
            "l3_7": W.laplacian_3d(7), "l2_64": W.laplacian_2d(64)}


def requests(seed=3):
    rng = np.random.default_rng(seed)
    names = list(systems())
    out = []
    for i, k in enumerate([1, 2, 3, 5, 8, 16, 4, 7, 1, 9, 16, 2]):
        name = names[i % len(names)]
        n = len(systems()[name][0]) - 1
        out.append((name, float(rng.uniform(0.0, 1.0)), rng.uniform(-1, 1, (n, k))))
    return out


def test_single_request_against_dense_solve(tnl):
    rp, c, v = W.laplacian_2d(40)
    n = len(rp) - 1
    rng = np.random.default_rng(1)
    B = rng.uniform(-1, 1, (n, 3))
    S = tnl.Server(vector_capacity=1 << 20, matrix_capacity=1 << 20, max_requests=8)
    mid = S.add_matrix(rp, c, v)
    t = S.submit(mid, B, shift=0.3)
    done = S.run_all()
    X, info = done[t]
    print(info, flush=True)
    assert info["converged"] and info["residue"] < 1e-10
    assert np.all(rel_residuals(rp, c, v, 0.3, B, X) < 1.01e-10)
    A = W.csr_to_dense(rp, c, v, n) + 0.3 * np.eye(n)
    Xref = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-9
    S.close()


def _run(tnl, reqs, vector_capacity, late=0, use_graph=True, iterations=8):
    S = tnl.Server(vector_capacity=vector_capacity, matrix_capacity=1 << 21, max_requests=8,
                   iterations_per_launch=iterations, use_graph=use_graph)
    mids = {name: S.add_matrix(*sysm) for name, sysm in systems().items()}
    tickets = []
    done = {}
    for i, (name, shift, B) in enumerate(reqs):
        tickets.append(S.submit(mids[name], B, shift=shift))
        if late and i % late == late - 1:  # let some work run before the next arrivals
            for t in S.progress():
                done[t] = S.result(t)
    done.update(S.run_all())
    S.close()
    return [done[t] for t in tickets]


def test_same_bits_alone_and_in_the_mix(tnl):
    """Bit-identical X (tolerance 0) for every request: alone in a server, all submitted together
    into a server too small to hold them at once (FIFO admission as vector memory frees up), and
    arriving while others run."""
    reqs = requests()
    alone = [_run(tnl, [r], 1 << 20)[0] for r in reqs]
    for (name, shift, B), (X, info) in zip(reqs, alone):
        rp, c, v = systems()[name]
        assert info["converged"], (name, info)
        assert np.all(rel_residuals(rp, c, v, shift, B, X) < 1.01e-10), name
    biggest = max(B.size for _, _, B in reqs)
    mixed = _run(tnl, reqs, vector_capacity=2 * biggest + 5000)
    staggered = _run(tnl, reqs, vector_capacity=1 << 20, late=3)
    for i, ((Xa, ia), (Xm, im), (Xs, is_)) in enumerate(zip(alone, mixed, staggered)):
        assert np.array_equal(Xa, Xm), i
        assert np.array_equal(Xa, Xs), i
        assert ia["outer_iterations"] == im["outer_iterations"] == is_["outer_iterations"]
        assert ia["inner_iterations"] == im["inner_iterations"] == is_["inner_iterations"]


def test_graph_and_direct_launches_agree(tnl):
    reqs = requests(5)[:5]
    a = _run(tnl, reqs, 1 << 20, use_graph=True, iterations=4)
    b = _run(tnl, reqs, 1 << 20, use_graph=False, iterations=3)
    for (Xa, _), (Xb, _) in zip(a, b):
        assert np.array_equal(Xa, Xb)


def test_edge_cases_and_capacity(tnl):
    rp, c, v = W.laplacian_2d(30)
    n = len(rp) - 1
    S = tnl.Server(vector_capacity=4 * n, matrix_capacity=1 << 16, max_requests=4)
    mid = S.add_matrix(rp, c, v)
    # b = 0 -> x = 0 at once
    t0 = S.submit(mid, np.zeros((n, 2)))
    # one right-hand side identically zero beside a normal one
    B = np.zeros((n, 2))
    B[:, 1] = np.linspace(-1, 1, n)
    t1 = S.submit(mid, B, shift=0.5)
    done = S.run_all()
    X0, i0 = done[t0]
    assert i0["converged"] and np.all(X0 == 0)
    X1, i1 = done[t1]
    assert i1["converged"] and np.all(X1[:, 0] == 0)
    assert rel_residuals(rp, c, v, 0.5, B[:, 1:], X1[:, 1:])[0] < 1.01e-10
    with pytest.raises(tnl.CapacityError):
        S.submit(mid, np.ones((n, 5)))  # larger than the vector arena
    with pytest.raises(tnl.CapacityError):
        S.add_matrix(*W.laplacian_2d(200))  # larger than the matrix arena
    with pytest.raises(tnl.TnlHipError):
        S.submit(mid, np.ones((n, 17)))
    S.close()


def test_gmres_requests_against_dense_solve(tnl):
    """GMRES requests (worker track) on a nonsymmetric matrix: FP64-refined X against a dense solve."""
    rp, c, v = W.convection_diffusion_3d(8, 10.0)
    n = len(rp) - 1
    rng = np.random.default_rng(2)
    B = rng.uniform(-1, 1, (n, 4))
    S = tnl.Server(vector_capacity=1 << 20, matrix_capacity=1 << 20, max_requests=8, gmres_workers=2)
    mid = S.add_matrix(rp, c, v)
    t = S.submit(mid, B, shift=0.05, method="gmres", inner_tolerance=1e-2)
    done = S.run_all()
    X, info = done[t]
    print(info, flush=True)
    assert info["converged"] and info["residue"] < 1e-10
    assert np.all(rel_residuals(rp, c, v, 0.05, B, X) < 1.01e-10)
    A = W.csr_to_dense(rp, c, v, n) + 0.05 * np.eye(n)
    Xref = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-8
    S.close()


def test_gmres_same_bits_alone_and_in_the_mix(tnl):
    """A GMRES request gives the same bits alone and among CG requests and other GMRES requests
    running at the same time on other streams; GMRES with no workers is refused."""
    reqs = requests(7)[:6]
    cd = W.convection_diffusion_3d(7, 30.0)
    rng = np.random.default_rng(9)
    gm = [(0.1 * (i + 1), rng.uniform(-1, 1, (len(cd[0]) - 1, k))) for i, k in enumerate([1, 3, 2, 5])]

    def run(with_cg, with_other_gmres, workers=3):
        S = tnl.Server(vector_capacity=1 << 20, matrix_capacity=1 << 21, max_requests=8, gmres_workers=workers)
        mids = {name: S.add_matrix(*sysm) for name, sysm in systems().items()}
        mcd = S.add_matrix(*cd)
        tickets = []
        if with_cg:
            for name, shift, B in reqs:
                S.submit(mids[name], B, shift=shift)
        for i, (shift, B) in enumerate(gm):
            if with_other_gmres or i == 0:
                tickets.append(S.submit(mcd, B, shift=shift, method="gmres", inner_tolerance=1e-2))
        done = S.run_all()
        S.close()
        return [done[t] for t in tickets]

    alone = run(False, False, workers=1)[0]
    mixed = run(True, True)[0]
    assert alone[1]["converged"]
    assert np.array_equal(alone[0], mixed[0])
    assert alone[1]["inner_iterations"] == mixed[1]["inner_iterations"]
    rp, c, v = cd
    for (shift, B), (X, info) in zip(gm, run(True, True)):
        assert info["converged"]
        assert np.all(rel_residuals(rp, c, v, shift, B, X) < 1.01e-10)
    S = tnl.Server(vector_capacity=1 << 16, matrix_capacity=1 << 16, max_requests=2, gmres_workers=0)
    mid = S.add_matrix(*W.laplacian_2d(10))
    with pytest.raises(tnl.TnlHipError):
        S.submit(mid, np.ones((100, 1)), method="gmres")
    S.close()
