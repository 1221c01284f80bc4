"""CPU pins of the fluid restatement (RSolverFluid, stabilised linear tetrahedra; the
system is for the increments of (vx, vy, vz, p) per node).

* an independent dense formulation, written from the stabilised weak form with the integrals
  in closed form (int N_m = V/4, int N_m N_n = V (1 + d_mn) / 20, int 1 = V), barycentric
  gradients from the 4x4 affine inverse, and the published Tezduyar parameters
  (tau = h / (2 |u|) above element Reynolds number 3, rho h^2 / (12 mu) below it, with the
  flow-direction length h = 2 / sum_a |s . grad N_a| for SUPG and the equal-volume sphere
  diameter for PSPG), steady and unsteady, in both Reynolds regimes;
* the hydrostatic rest state (v = 0, grad p = rho g): zero right-hand side on every interior
  momentum row and every continuity row;
* consistency of residual and matrix: at rho = 1, g = 0 and a state of amplitude eps, the
  right-hand side equals -eps A [v; p] up to O(eps^2);
* a uniform flow at constant pressure has zero residual on the interior rows;
* the write-back (increments added through the node book, acceleration, norms).
"""
import numpy as np
import pytest

from helpers import bary_gradients, kuhn_box_np

EOD = 0xFFFFFFFF


def channel_fixed(xyz):
    """Walls y = 0, 1 and z = 0, 1 and the inflow x = 0 fix the velocity; the outflow x = 1
    fixes the pressure."""
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    near = lambda a, v: np.abs(a - v) < 1e-12
    wall = near(y, 0) | near(y, 1) | near(z, 0) | near(z, 1) | near(x, 0)
    fixed = np.zeros((xyz.shape[0], 4), np.uint8)
    fixed[wall, :3] = 1
    fixed[near(x, 1), 3] = 1
    return fixed.reshape(-1)


def tau(h, speed, rho, mu):
    """Tezduyar's stabilisation parameter with the switch at element Reynolds number 3."""
    if speed <= 0.0:
        return 0.0
    re = rho * speed * h / (2.0 * mu)
    return rho * h * h / (12.0 * mu) if 0.0 < re <= 3.0 else h / (2.0 * speed)


def dense_fluid(xyz, tets, rho, mu, grav, state, v_old, acc, U, alpha, dt, unsteady):
    """Global dense matrix and right-hand side over all 4 nn positions."""
    nn = xyz.shape[0]
    A = np.zeros((4 * nn, 4 * nn))
    f = np.zeros(4 * nn)
    I3 = np.eye(3)
    Mmn = (np.ones((4, 4)) + np.eye(4)) / 20.0           # int N_m N_n / V
    for e, t in enumerate(tets):
        g, V = bary_gradients(xyz[t])                      # g[m] = grad N_m
        r, m_, gr = rho[e], mu[e], grav[e]
        st = state[t]                                      # (4, 4)
        v, p = st[:, :3], st[:, 3]
        ve, pe = v.mean(0), p.mean()
        vo = v_old[t].mean(0)
        an = acc[t]
        Jv = v.T @ g                                       # Jv[c, a] = d v_c / d x_a
        gp = g.T @ p                                       # grad p
        conv = Jv @ ve                                     # (ve . grad) v
        divv = np.trace(Jv)
        speed = np.linalg.norm(vo)
        s = vo / speed if speed > 0 else np.zeros(3)
        sg = np.abs(g @ s).sum()
        h = 2.0 / sg if sg != 0 else 0.0
        hn = (6.0 * V / np.pi) ** (1.0 / 3.0)
        ts, tp, tl = tau(h, speed, r, m_), tau(hn, U, r, m_), speed * h / 2.0
        adv = g @ ve                                       # ve . grad N_m
        Ae = np.zeros((4, 4, 4, 4))                        # [m, i, n, j]
        be = np.zeros((4, 4))
        for m in range(4):
            for n in range(4):
                # momentum-velocity: Newton-linearised convection (reaction with int N N, advection
                # with the element velocity), viscous rho mu (grad w : (grad u + grad u^T)),
                # SUPG on convection, reaction, and the test-function perturbation terms, LSIC
                react = r * Mmn[m, n] * V * Jv
                advect = r * V / 4 * adv[n] * I3
                visc = r * m_ * V * (np.dot(g[m], g[n]) * I3 + np.outer(g[n], g[m]))
                supg = ts * r * V * adv[m] * adv[n] * I3 + ts * r * adv[m] * V / 4 * Jv \
                    + ts * V / 4 * (r * np.outer(conv, g[m]) + np.outer(gp, g[m]))
                lsic = tl * r * V * np.outer(g[m], g[n])
                steady = react + visc + supg + lsic
                if unsteady:
                    mass = (r * Mmn[m, n] * V + ts * r * adv[m] * V / 4) * I3 \
                        + alpha * dt * ts * r * V / 4 * np.outer(an[n], g[m])
                    Ae[m, :3, n, :3] = mass + alpha * dt * (steady + advect - lsic) + dt * lsic
                else:
                    Ae[m, :3, n, :3] = steady + advect
                # momentum-pressure: -int p div w and the (reference-signed) SUPG pressure term
                Ae[m, :3, n, 3] = -(dt if unsteady else 1.0) * (g[m] * V / 4 + ts * adv[m] * V * g[n])
                # continuity-velocity: int q div u, PSPG on convection and on the reaction term
                cont = g[n] * V / 4
                pspg = tp * V * adv[n] * g[m] + tp * V / 4 * (Jv.T @ g[m])
                if unsteady:
                    Ae[m, 3, n, :3] = tp * V / 4 * g[m] + dt * cont + alpha * dt * pspg
                else:
                    Ae[m, 3, n, :3] = cont + pspg
                Ae[m, 3, n, 3] = (dt if unsteady else 1.0) * tp * V * np.dot(g[m], g[n]) / r
            force = r * V / 4 * gr + ts * r * adv[m] * V * gr
            res = r * V / 4 * conv + m_ * V * (Jv + Jv.T) @ g[m] - g[m] * pe * V \
                + ts * r * adv[m] * V * conv - ts * adv[m] * V * gp + tl * r * divv * V * g[m]
            cres = tp * V * np.dot(g[m], gr) - (V / 4 * divv + tp * V * np.dot(g[m], conv)
                                                 + tp * V * np.dot(g[m], gp) / r)
            if unsteady:
                be[m, :3] = dt * (force - (r * V / 4 * an[m] + ts * r * adv[m] * V * an[m] + res))
                be[m, 3] = dt * (cres - tp * V * np.dot(g[m], an[m]))
            else:
                be[m, :3] = force - res
                be[m, 3] = cres
        idx = (4 * t[:, None] + np.arange(4)[None, :]).reshape(-1)
        A[np.ix_(idx, idx)] += Ae.reshape(16, 16)
        f[idx] += be.reshape(16)
    return A, f


def reduce(A, f, book):
    free = book != EOD
    fi = np.nonzero(free)[0][np.argsort(book[free])]
    return A[np.ix_(fi, fi)], f[fi]


def make_state(rng, nn, ne, speed, p_scale=1.0):
    state = np.zeros((nn, 4))
    state[:, :3] = rng.standard_normal((nn, 3)) * speed
    state[:, 3] = rng.standard_normal(nn) * p_scale
    v_old = state[:, :3] + rng.standard_normal((nn, 3)) * 0.1 * speed
    acc = rng.standard_normal((nn, 3)) * speed
    return state, v_old, acc


@pytest.mark.parametrize("unsteady", [0, 1])
@pytest.mark.parametrize("speed", [1e-3, 20.0])
def test_fluid_assembly_dense(orc, unsteady, speed):
    xyz, tets = kuhn_box_np(2, 3, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(5 + unsteady)
    rho = rng.uniform(0.5, 2.0, ne)
    mu = rng.uniform(0.01, 0.1, ne)
    grav = np.tile([0.0, 0.0, -9.81], (ne, 1)) + rng.standard_normal((ne, 3))
    state, v_old, acc = make_state(rng, nn, ne, speed)
    fixed = channel_fixed(xyz)
    book, n_dof = orc.node_book(nn, tets, 4, fixed)
    args = (rho, mu, grav, state, v_old, acc, 1.7, 0.6, 0.05, unsteady)
    A, b = orc.fluid_assemble(xyz, tets, *args, book, n_dof)
    Ad, fd = reduce(*dense_fluid(xyz, tets, *args), book)
    Ao = A.toarray()
    assert np.abs(Ao - Ad).max() <= 1e-12 * np.abs(Ad).max()
    assert np.abs(b - fd).max() <= 1e-12 * np.abs(fd).max()
    # the slow state has every SUPG element Reynolds number below 3, the fast one above; the
    # PSPG numbers (stream velocity 1.7) fall on both sides
    re_s, re_p = [], []
    for e, t in enumerate(tets):
        g, V = bary_gradients(xyz[t])
        vo = v_old[t].mean(0)
        sp_ = np.linalg.norm(vo)
        h = 2.0 / np.abs(g @ (vo / sp_)).sum()
        re_s.append(rho[e] * sp_ * h / (2 * mu[e]))
        re_p.append(rho[e] * 1.7 * (6 * V / np.pi) ** (1 / 3) / (2 * mu[e]))
    re_s, re_p = np.array(re_s), np.array(re_p)
    assert (re_s <= 3).all() if speed < 1 else (re_s > 3).all()
    assert (re_p <= 3).any() and (re_p > 3).any()


def interior_momentum_rows(xyz, book):
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    inner = (x > 1e-12) & (x < 1 - 1e-12) & (y > 1e-12) & (y < 1 - 1e-12) & (z > 1e-12) & (z < 1 - 1e-12)
    rows = [book[4 * i + c] for i in np.nonzero(inner)[0] for c in range(3) if book[4 * i + c] != EOD]
    return np.array(rows, dtype=np.int64)


@pytest.mark.parametrize("unsteady", [0, 1])
def test_fluid_hydrostatic_rest(orc, unsteady):
    xyz, tets = kuhn_box_np(3, 3, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    rho0, g = 1.3, np.array([0.4, -1.1, -9.81])
    state = np.zeros((nn, 4))
    state[:, 3] = rho0 * xyz @ g + 2.0                     # grad p = rho g
    zeros3 = np.zeros((nn, 3))
    fixed = np.zeros((nn, 4), np.uint8)
    fixed[0, 3] = 1                                         # pressure level
    book, n_dof = orc.node_book(nn, tets, 4, fixed.reshape(-1))
    A, b = orc.fluid_assemble(xyz, tets, np.full(ne, rho0), np.full(ne, 0.05), np.tile(g, (ne, 1)), state, zeros3,
                              zeros3, 2.0, 0.5, 0.01, unsteady, book, n_dof)
    scale = rho0 * np.linalg.norm(g) / ne
    cont = np.array([book[4 * i + 3] for i in range(nn) if book[4 * i + 3] != EOD])
    mom = interior_momentum_rows(xyz, book)
    assert mom.size > 0
    assert np.abs(b[mom]).max() <= 1e-12 * scale * (0.01 if unsteady else 1.0) * 1e3
    assert np.abs(b[cont]).max() <= 1e-12 * scale * 1e3
    # the boundary momentum rows carry the unbalanced pressure force, so b is not trivially zero
    assert np.abs(b).max() > 1e3 * np.abs(b[mom]).max()


def test_fluid_residual_matches_matrix(orc):
    """rho = 1, g = 0: b(eps v, eps p) = -eps A(eps v) [v; p] + O(eps^2)."""
    xyz, tets = kuhn_box_np(2, 2, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(9)
    fixed = channel_fixed(xyz)
    book, n_dof = orc.node_book(nn, tets, 4, fixed)
    v = rng.standard_normal((nn, 4))
    v[fixed.reshape(nn, 4) == 1] = 0.0                      # increments live on the free positions
    eps = 1e-7
    state = eps * v
    A, b = orc.fluid_assemble(xyz, tets, np.ones(ne), rng.uniform(0.02, 0.2, ne), np.zeros((ne, 3)), state,
                                  state[:, :3].copy(), np.zeros((nn, 3)), 1.0, 1.0, 1.0, 0, book, n_dof)
    free = book != EOD
    y = np.zeros(n_dof)
    y[book[free]] = v.reshape(-1)[free]
    lin = -eps * (A @ y)
    assert np.linalg.norm(b) > 0
    assert np.linalg.norm(b - lin) <= 1e-5 * np.linalg.norm(b)


def test_fluid_uniform_flow(orc):
    xyz, tets = kuhn_box_np(3, 3, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    state = np.zeros((nn, 4))
    state[:, :3] = [1.5, -0.4, 0.8]
    state[:, 3] = 3.0
    fixed = np.zeros((nn, 4), np.uint8)
    fixed[0, 3] = 1
    book, n_dof = orc.node_book(nn, tets, 4, fixed.reshape(-1))
    _, b = orc.fluid_assemble(xyz, tets, np.ones(ne), np.full(ne, 0.01), np.zeros((ne, 3)), state,
                              state[:, :3].copy(), np.zeros((nn, 3)), 1.0, 1.0, 1.0, 0, book, n_dof)
    cont = np.array([book[4 * i + 3] for i in range(nn) if book[4 * i + 3] != EOD])
    mom = interior_momentum_rows(xyz, book)
    assert np.abs(b[mom]).max() <= 1e-13 and np.abs(b[cont]).max() <= 1e-13


@pytest.mark.parametrize("unsteady", [0, 1])
def test_fluid_update(orc, unsteady):
    rng = np.random.default_rng(3)
    nn = 40
    fixed = (rng.random(4 * nn) < 0.3).astype(np.uint8)
    book = np.full(4 * nn, EOD, np.uint32)
    book[fixed == 0] = np.arange(int((fixed == 0).sum()), dtype=np.uint32)
    x = rng.standard_normal(int((fixed == 0).sum()))
    state = rng.standard_normal((nn, 4))
    v_old = rng.standard_normal((nn, 3))
    acc = np.zeros((nn, 3))
    s0, vo0 = state.copy(), v_old.copy()
    norms = orc.fluid_update(book, x, unsteady, 0.25, state, v_old, acc)
    inc = np.zeros(4 * nn)
    inc[fixed == 0] = x
    s1 = s0 + inc.reshape(nn, 4)
    np.testing.assert_array_equal(state, s1)
    if unsteady:
        np.testing.assert_array_equal(v_old, vo0)
        np.testing.assert_allclose(acc, (s1[:, :3] - vo0) / 0.25, rtol=1e-15)
    else:
        np.testing.assert_array_equal(v_old, s0[:, :3])
    np.testing.assert_allclose(norms, [np.linalg.norm(s0[:, :3]), np.linalg.norm(s1[:, :3]),
                                       np.linalg.norm(s0[:, 3]), np.linalg.norm(s1[:, 3])], rtol=1e-14)
