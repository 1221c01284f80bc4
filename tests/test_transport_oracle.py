"""CPU pins of the scalar-transport restatement (RSolverFluidHeat / RSolverFluidParticle: SUPG
advection-diffusion of a node field by the element velocity, heat-style lifting):

* an independent dense formulation with the integrals in closed form (int N_m = V/4,
  int N_m N_n = V (1 + d_mn) / 20, int 1 = V), barycentric gradients and the published
  doubly-asymptotic SUPG parameter tau = h / (2 |v|) min(Pe / 3, 1), Pe = |v| h / (2 k), with the
  flow-direction length h = 2 / sum_a |s . grad N_a|; steady and unsteady, both Peclet regimes;
* pure diffusion (v = 0, no source) reproduces a linear field from its boundary values (the
  reference's diffusion block is the negated Laplacian, which leaves the solution unchanged);
* an unsteady step at rest keeps a constant field constant;
* the write-back clamps at zero.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from helpers import bary_gradients, kuhn_box_np

EOD = 0xFFFFFFFF


def tau(h, speed, k):
    if speed <= 0.0:
        return 0.0
    pe = speed * h / (2.0 * k) if k != 0.0 else 0.0
    return h / (2.0 * speed) * (pe / 3.0 if 0.0 < pe <= 3.0 else 1.0)


def dense_transport(xyz, tets, ca, k, vel, q, field, alpha, dt, unsteady):
    nn = xyz.shape[0]
    A = np.zeros((nn, nn))
    f = np.zeros(nn)
    Mmn = (np.ones((4, 4)) + np.eye(4)) / 20.0
    for e, t in enumerate(tets):
        g, V = bary_gradients(xyz[t])
        v = vel[e]
        speed = np.linalg.norm(v)
        s = v / speed if speed > 0 else np.zeros(3)
        sg = np.abs(g @ s).sum()
        h = 2.0 / sg if sg != 0 else 0.0
        ts = tau(h, speed, k[e])
        adv = g @ v                                           # v . grad N_m
        # advection with the Galerkin test function, negated diffusion, SUPG streamline
        # diffusion, and the SUPG term that does not depend on the trial function
        S = ca[e] * V / 4 * np.outer(np.ones(4), adv) - k[e] * V * (g @ g.T) \
            + ts * ca[e] * V * np.outer(adv, adv) + ts * V * np.outer(adv, np.ones(4))
        fv = q[t] * V / 4
        if unsteady:
            M = ca[e] * V * Mmn
            Ae = M + alpha * dt * S
            be = dt * fv + (M - (1 - alpha) * dt * S) @ field[t]
        else:
            Ae, be = S, fv
        A[np.ix_(t, t)] += Ae
        f[t] += be
    return A, f


def eliminate(A, f, book, field):
    free = book != EOD
    fi = np.nonzero(free)[0][np.argsort(book[free])]
    fx = np.nonzero(~free)[0]
    return A[np.ix_(fi, fi)], f[fi] - A[np.ix_(fi, fx)] @ field[fx]


def boundary(xyz):
    return ((np.abs(xyz) < 1e-12) | (np.abs(xyz - 1) < 1e-12)).any(1)


@pytest.mark.parametrize("unsteady", [0, 1])
@pytest.mark.parametrize("speed", [1e-2, 500.0])
def test_transport_dense(orc, unsteady, speed):
    xyz, tets = kuhn_box_np(3, 2, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(31 + unsteady)
    ca = rng.uniform(0.5, 3.0, ne)
    k = rng.uniform(0.5, 2.0, ne)
    vel = rng.standard_normal((ne, 3))
    vel *= speed / np.linalg.norm(vel, axis=1, keepdims=True)
    q = rng.standard_normal(nn)
    field = rng.uniform(0.0, 2.0, nn)
    fixed = boundary(xyz) & (xyz[:, 0] < 0.5)
    book, n_dof = orc.node_book(nn, tets, 1, fixed.astype(np.uint8))
    A, b = orc.transport_assemble(xyz, tets, ca, k, vel, q, field, 0.7, 0.05, unsteady, book, n_dof)
    Ad, fd = eliminate(*dense_transport(xyz, tets, ca, k, vel, q, field, 0.7, 0.05, unsteady), book, field)
    Ao = A.toarray()
    assert np.abs(Ao - Ad).max() <= 1e-12 * np.abs(Ad).max()
    assert np.abs(b - fd).max() <= 1e-12 * np.abs(fd).max()
    # slow: every element Peclet number below 3; fast: every one above
    pe = []
    for e, t in enumerate(tets):
        g, _ = bary_gradients(xyz[t])
        sp_ = np.linalg.norm(vel[e])
        pe.append(sp_ * (2.0 / np.abs(g @ (vel[e] / sp_)).sum()) / (2 * k[e]))
    pe = np.array(pe)
    assert (pe <= 3).all() if speed < 1 else (pe > 3).all()


def test_transport_pure_diffusion_linear_field(orc):
    xyz, tets = kuhn_box_np(4, 3, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    T = 2.0 + xyz @ np.array([1.5, -0.5, 0.75])
    fixed = boundary(xyz)
    book, n_dof = orc.node_book(nn, tets, 1, fixed.astype(np.uint8))
    rng = np.random.default_rng(2)
    A, b = orc.transport_assemble(xyz, tets, rng.uniform(1, 2, ne), np.full(ne, 0.3), np.zeros((ne, 3)),
                                  np.zeros(nn), np.where(fixed, T, 0.0), 1.0, 1.0, 0, book, n_dof)
    x = spla.spsolve(A.tocsc(), b)
    free = ~fixed
    np.testing.assert_allclose(x[book[free]], T[free], rtol=1e-12)


def test_transport_unsteady_rest_keeps_constant(orc):
    xyz, tets = kuhn_box_np(3, 3, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    fixed = boundary(xyz)
    book, n_dof = orc.node_book(nn, tets, 1, fixed.astype(np.uint8))
    field = np.full(nn, 3.25)
    A, b = orc.transport_assemble(xyz, tets, np.full(ne, 2.0), np.full(ne, 0.4), np.zeros((ne, 3)), np.zeros(nn),
                                  field, 0.5, 0.01, 1, book, n_dof)
    x = spla.spsolve(A.tocsc(), b)
    np.testing.assert_allclose(x, 3.25, rtol=1e-12)


def test_transport_update_clamps(orc):
    rng = np.random.default_rng(4)
    nn = 30
    book = np.full(nn, EOD, np.uint32)
    free = rng.random(nn) < 0.7
    book[free] = np.arange(free.sum(), dtype=np.uint32)
    x = rng.standard_normal(int(free.sum()))
    field = rng.uniform(0, 1, nn)
    f0 = field.copy()
    norms = orc.transport_update(book, x, field)
    expect = f0.copy()
    expect[free] = np.maximum(x, 0.0)
    np.testing.assert_array_equal(field, expect)
    np.testing.assert_allclose(norms, [np.linalg.norm(f0), np.linalg.norm(expect)], rtol=1e-14)
