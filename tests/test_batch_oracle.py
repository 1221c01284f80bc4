"""CPU pins of the batched transient heat path: the oracle's time loop against an independent
dense formulation, and the right-hand-side recurrence the batch forms on the device."""
import numpy as np
import pytest

from helpers import dense_heat


def _problem(rs, n, seed=7):
    xyz, tets = rs.kuhn_box(n, n, n)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    t0 = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
    ne = tets.shape[0]
    k = rs.field_lognormal(seed, ne, 0.0, 1.5)
    q = rs.field_uniform(seed + 1, ne)
    rhoc = 0.5 + rs.field_uniform(seed + 2, ne)
    return xyz, tets, fixed, t0, k, q, rhoc


def test_transient_oracle_matches_dense_steps(rs, orc):
    """Three implicit-Euler steps: the oracle (sparse assembly + Jacobi-CG from x = 0 each
    step) against dense assembly from barycentric gradients + a direct solve."""
    xyz, tets, fixed, t0, k, q, rhoc = _problem(rs, 3)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    dt, steps = 1e-2, 3
    xs, its = orc.transient_heat(xyz, tets, k, q, rhoc, t0, book, n_dof, dt, steps, cvg=1e-14)
    t = t0.copy()
    free = book != 0xFFFFFFFF
    for s in range(steps):
        Aff, bf, _, _ = dense_heat(xyz, tets, k, q, t, book, n_dof, rhoc=rhoc, transient=True, alpha=1.0, dt=dt)
        x = np.linalg.solve(Aff, bf)
        t[free] = x[book[free]]
        np.testing.assert_allclose(xs[s], x, rtol=1e-11, atol=1e-13)
        assert its[s] > 0
    # the temperatures move toward the hot wall: every step changes the solution
    assert np.linalg.norm(xs[1] - xs[0]) > 0 and np.linalg.norm(xs[2] - xs[1]) > 0


def test_rhs_recurrence(rs, orc):
    """b(T) assembled with free temperatures T_f equals c + M_ff T_f, where c = b(0) and M_ff
    is the matrix assembled with dt = 0 -- the identity the batch uses instead of assembling
    every step.  A dropped lifting term or a wrong sign breaks it."""
    xyz, tets, fixed, t0, k, q, rhoc = _problem(rs, 3, seed=11)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    free = book != 0xFFFFFFFF
    dt = 1e-3
    M, _ = orc.heat_assemble(xyz, tets, k, q, t0, book, n_dof, rhoc=rhoc, transient=True, alpha=1.0, dt=0.0)
    A, c = orc.heat_assemble(xyz, tets, k, q, t0, book, n_dof, rhoc=rhoc, transient=True, alpha=1.0, dt=dt)
    rng = np.random.default_rng(3)
    t = t0.copy()
    t[free] = rng.random(int(free.sum()))
    A2, b = orc.heat_assemble(xyz, tets, k, q, t, book, n_dof, rhoc=rhoc, transient=True, alpha=1.0, dt=dt)
    tf = np.empty(n_dof)
    tf[book[free]] = t[free]
    np.testing.assert_allclose(b, c + M @ tf, rtol=1e-12, atol=1e-15)
    assert abs(A2 - A).max() <= 1e-14 * abs(A).max()
    # M is the consistent mass matrix: rows of the full (undeleted) M sum to rho*c*V/4 per tet,
    # so for interior rows M_ff rows sum to the nodal share of rho*c volume (> 0)
    assert (M.toarray() > -1e-18).all() and M.diagonal().min() > 0


def test_system_sizes(rs):
    tr = __import__("__graft_entry__").load_module("transient")
    s = tr.system_sizes(rs, 512, 8, 48, 4)
    assert s.shape == (512,) and s.min() >= 8 and s.max() <= 48
    assert len(set(s.tolist())) > 30
    assert np.array_equal(s, tr.system_sizes(rs, 512, 8, 48, 4))
