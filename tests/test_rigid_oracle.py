"""CPU pins of the rigid-body aggregation in the oracle (learned_oracle.rigid_hierarchy), the
restatement the GPU's rigid-body hierarchy (RS_AGG_RIGID) is checked against:

* frame_shift: the modes about the origin are the modes about p times frame_shift(p);
* mgs2: orthonormal columns, B = Q R with R upper triangular, and dependent modes found dead
  (one node: only its 3 translations live; collinear nodes: the rotation about the line dead);
* the hierarchy on the cantilever: orthonormal prolongations, the global rigid-body motions lie
  in every composite coarse space P_0 ... P_k (a frame mistake between levels breaks this), the
  level operators are P^T A P, and flexible CG with the cycle needs far fewer iterations than
  with per-component aggregation (which keeps only the translations).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import learned_oracle as lo

EOD = 0xFFFFFFFF


def test_frame_shift():
    rng = np.random.default_rng(3)
    r, p = rng.standard_normal((30, 3)), rng.standard_normal(3)
    comp = rng.integers(0, 3, 30)
    assert np.allclose(lo.rigid_modes(r + p, comp), lo.rigid_modes(r, comp) @ lo.frame_shift(p), atol=1e-14)
    assert np.allclose(lo.frame_shift(p) @ lo.frame_shift(-p), np.eye(6), atol=1e-14)


def test_mgs2_orthonormal_and_dead_columns():
    rng = np.random.default_rng(4)
    B = rng.standard_normal((24, 6))
    Q, R, dead = lo.mgs2(B)
    assert not dead.any()
    assert np.allclose(Q.T @ Q, np.eye(6), atol=1e-13) and np.allclose(Q @ R, B, atol=1e-13)
    assert np.allclose(R, np.triu(R))
    # one node: the three translations only
    Q, R, dead = lo.mgs2(lo.rigid_modes(np.zeros((3, 3)), np.arange(3)))
    assert dead.tolist() == [False] * 3 + [True] * 3
    assert np.allclose(R[3:], 0.0) and np.allclose(Q[:, 3:], 0.0)
    # nodes on the x axis: the rotation about x does nothing there
    pts = np.repeat(np.array([[0.0, 0, 0], [1.0, 0, 0], [2.5, 0, 0]]), 3, axis=0)
    _, _, dead = lo.mgs2(lo.rigid_modes(pts, np.tile(np.arange(3), 3)))
    assert dead.tolist() == [False] * 4 + [True, False]


@pytest.fixture(scope="module")
def cantilever(rs, orc):
    nx, ny, nz = 32, 8, 8
    xyz, tets = rs.kuhn_box(nx, ny, nz, 4.0, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    E = rs.field_lognormal(3, ne, 0.0, 1.0)
    tris = rs.plane_faces(xyz, tets, 0, 4.0)
    trac = np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1))
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    A, b = orc.stress_assemble(xyz, tets, E, np.full(ne, 0.3), np.ones(ne), [0.0, 0.0, 0.0], np.zeros(3 * nn), tris,
                               trac, book, n_dof)
    return xyz, np.asarray(book), A.tocsr(), b, (nx // 2, ny // 2, nz // 2)


def test_rigid_hierarchy_keeps_rigid_motions(cantilever):
    xyz, book, A, b, bins = cantilever
    levels, Ps, deads = lo.rigid_hierarchy(A, xyz, book, bins, 512)
    assert len(levels) >= 3 and levels[-1].shape[0] <= 512
    n = A.shape[0]
    pos = np.nonzero(book.astype(np.uint64) != EOD)[0]
    node, comp = np.empty(n, np.int64), np.empty(n, np.int64)
    node[book[pos].astype(np.int64)] = pos // 3
    comp[book[pos].astype(np.int64)] = pos % 3
    U = lo.rigid_modes(xyz[node], comp)           # the 6 global rigid motions at the free DOFs
    T = sp.eye(n, format="csr")
    for l, P in enumerate(Ps):
        assert np.allclose((P.T @ P).toarray(), np.diag(~deads[l]).astype(float), atol=1e-12)
        ref = (P.T @ levels[l] @ P).toarray()
        ref[deads[l], deads[l]] = 1.0
        assert np.allclose(levels[l + 1].toarray(), ref, atol=1e-12 * abs(ref).max())
        T = T @ P
        Qc, _ = np.linalg.qr(T.toarray())
        resid = U - Qc @ (Qc.T @ U)
        assert np.linalg.norm(resid) <= 1e-10 * np.linalg.norm(U), l


def test_rigid_cycle_beats_per_component_cycle(cantilever):
    xyz, book, A, b, bins = cantilever

    def iterations(levels, Ps):
        lams = [lo.gershgorin(L) for L in levels[:-1]]
        cinv = np.linalg.inv(levels[-1].toarray())
        _, its = lo.fcg(A, b, lambda r: lo.vcycle(levels, Ps, r, 1, 1, lambda Ac, rc: cinv @ rc, lams), tol=1e-8)
        return its

    levels, Ps, _ = lo.rigid_hierarchy(A, xyz, book, bins, 512)
    rigid = iterations(levels, Ps)
    maps = lo.bins_hierarchy(xyz, book, 3, bins, 512, by_component=True)
    comp_levels, comp_Ps = lo.hierarchy(A, maps)
    per_comp = iterations(comp_levels, comp_Ps)
    assert rigid <= 40 and per_comp >= 100 and 2 * rigid < per_comp
