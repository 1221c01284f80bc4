"""CPU oracle of the learned coarse-correction network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward/backward kernels of the learned
preconditioner.  The learned preconditioner has no counterpart in the reference; this file
pins its mathematics: the backward pass here is checked against central finite differences of
the forward pass (tests/test_learned_oracle.py), and the GPU gradients are then checked against
this backward pass with the same BF16 rounding points (`bf16=True`).

Network (per coarse graph with N nodes, dense coarse operator Ac, coarse residual rc; with
`bf16_points` the weights wn, the activations H, M and the back-propagated dY are rounded to
BF16 where the GPU kernels round them):
    d_I = Ac[I,I];  u_I = rc_I / d_I;  sigma = sqrt(mean(u^2)) + tiny
    wn_IJ = -Ac[I,J] / d_I  (J != I), 0 on the diagonal
    F = [u/sigma, log(d/geomean(d)), 1, 0]                       (N x 4)
    H0 = relu(F Win + bin)
    for l in 0..L-1:  M = wn H;  Y = H Ws[l] + M Wn[l] + b[l];  H = H + relu(Y)
    o = H wo + bo;   e = u + sigma * o
Loss for a (residual, error) pair with s = e_k.r_k = ||e_k||_A^2:
    L = (e.Ac.e - 2 e.rc) / s + 1   (the A-norm error fraction left after the correction P e)

Multilevel cycle (the preconditioner the network sits in): piecewise-constant prolongations
P_l from level maps, Galerkin operators A_{l+1} = P_l^T A_l P_l, Chebyshev smoothing in D^-1 A_l
over [lambda/30, 1.05 lambda] with lambda the Gershgorin bound of D^-1 A_l (Adams, Brezina, Hu,
Tuminaro 2003), pre/post steps, a coarsest solver.  Pinned by closed forms in
tests/test_learned_oracle.py: the smoother on D^-1 A = I against the Chebyshev residual
polynomial, the two-level projection property, symmetry of the symmetric cycle.
"""
from __future__ import annotations

import numpy as np

HID = 64
NIN = 4
LAYERS = 6
TINY = 1e-30


def param_count(layers: int = LAYERS, hid: int = HID) -> int:
    return NIN * hid + hid + layers * (2 * hid * hid + hid) + hid + 1


def unpack(theta: np.ndarray, layers: int = LAYERS, hid: int = HID) -> dict:
    t = np.asarray(theta, dtype=np.float64)
    o = 0

    def take(n, shape):
        nonlocal o
        a = t[o:o + n].reshape(shape)
        o += n
        return a
    p = {"Win": take(NIN * hid, (NIN, hid)), "bin": take(hid, (hid,)), "Ws": [], "Wn": [], "b": []}
    for _ in range(layers):
        p["Ws"].append(take(hid * hid, (hid, hid)))
        p["Wn"].append(take(hid * hid, (hid, hid)))
        p["b"].append(take(hid, (hid,)))
    p["wo"] = take(hid, (hid,))
    p["bo"] = take(1, (1,))
    assert o == t.size
    return p


def pack(g: dict) -> np.ndarray:
    parts = [g["Win"].ravel(), g["bin"].ravel()]
    for l in range(len(g["Ws"])):
        parts += [g["Ws"][l].ravel(), g["Wn"][l].ravel(), g["b"][l].ravel()]
    parts += [g["wo"].ravel(), g["bo"].ravel()]
    return np.concatenate(parts)


def init_params(seed: int, layers: int = LAYERS, hid: int = HID) -> np.ndarray:
    """He-style init for the hidden maps, zero output head (the untrained network is the
    coarse Jacobi estimate e = u)."""
    rng = np.random.default_rng(seed)
    g = {"Win": rng.standard_normal((NIN, hid)) * np.sqrt(2.0 / NIN), "bin": np.zeros(hid),
         "Ws": [], "Wn": [], "b": []}
    for _ in range(layers):
        g["Ws"].append(rng.standard_normal((hid, hid)) * np.sqrt(1.0 / hid) * 0.5)
        g["Wn"].append(rng.standard_normal((hid, hid)) * np.sqrt(1.0 / hid) * 0.5)
        g["b"].append(np.zeros(hid))
    g["wo"] = np.zeros(hid)
    g["bo"] = np.zeros(1)
    return pack(g)


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest even) and back to float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def _inputs(Ac, rc):
    d = np.diag(Ac).copy()
    u = rc / d
    sigma = np.sqrt(np.mean(u * u)) + TINY
    logd = np.log(d)
    F = np.zeros((Ac.shape[0], NIN))
    F[:, 0] = u / sigma
    F[:, 1] = logd - logd.mean()
    F[:, 2] = 1.0
    wn = -Ac / d[:, None]
    np.fill_diagonal(wn, 0.0)
    return d, u, sigma, F, wn


def forward(theta, Ac, rc, s=None, bf16_points: bool = False, layers: int = LAYERS):
    """Returns (e, cache) or (e, loss, cache) when s is given."""
    p = unpack(theta, layers)
    q = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    d, u, sigma, F, wn = _inputs(Ac, rc)
    wn = q(wn)      # the message weights are BF16 operands as well
    Z0 = F @ p["Win"] + p["bin"]
    H = [q(np.maximum(Z0, 0.0))]
    Ms, Ys = [], []
    for l in range(layers):
        M = q(wn @ H[l])
        Y = q(H[l]) @ q(p["Ws"][l]) + M @ q(p["Wn"][l]) + p["b"][l]
        Ms.append(M)
        Ys.append(Y)
        H.append(q(H[l] + np.maximum(Y, 0.0)))
    o = H[-1] @ p["wo"] + p["bo"][0]
    e = u + sigma * o
    cache = dict(p=p, d=d, u=u, sigma=sigma, F=F, wn=wn, Z0=Z0, H=H, M=Ms, Y=Ys, o=o, e=e)
    if s is None:
        return e, cache
    loss = (e @ Ac @ e - 2.0 * e @ rc) / s + 1.0
    return e, loss, cache


def backward(theta, Ac, rc, s, bf16_points: bool = False, layers: int = LAYERS):
    """Gradient of the loss with respect to theta (sigma and the inputs are constants)."""
    q = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    e, loss, c = forward(theta, Ac, rc, s, bf16_points, layers)
    p = c["p"]
    ge = 2.0 * (Ac @ e - rc) / s
    go = ge * c["sigma"]
    g = {"Ws": [None] * layers, "Wn": [None] * layers, "b": [None] * layers}
    H = c["H"]
    g["wo"] = H[-1].T @ go
    g["bo"] = np.array([go.sum()])
    dH = np.outer(go, p["wo"])
    for l in reversed(range(layers)):
        dY = q(dH * (c["Y"][l] > 0))
        g["Ws"][l] = q(H[l]).T @ dY
        g["Wn"][l] = c["M"][l].T @ dY
        g["b"][l] = dY.sum(axis=0)
        dM = dY @ q(p["Wn"][l]).T
        dH = dH + dY @ q(p["Ws"][l]).T + c["wn"].T @ dM
    dZ0 = dH * (c["Z0"] > 0)
    g["Win"] = c["F"].T @ dZ0
    g["bin"] = dZ0.sum(axis=0)
    return pack(g), loss


def galerkin(A, agg, n_agg):
    """Dense Pt A P for piecewise-constant aggregation (A scipy sparse)."""
    import scipy.sparse as sp
    n = A.shape[0]
    P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, n_agg))
    return (P.T @ A @ P).toarray()


# ------------------------------------------------------------------------------ multilevel cycle
def prolongation(m, n_next):
    import scipy.sparse as sp
    m = np.asarray(m, dtype=np.int64)
    return sp.csr_matrix((np.ones(m.size), (np.arange(m.size), m)), shape=(m.size, n_next))


def bins_hierarchy(xyz, book, dpn, bins, max_coarse, max_levels=16, eod=0xFFFFFFFF, by_component=False):
    """Nested coordinate bins over the enabled positions of `book` (numpy restatement that pins
    the library's host and GPU builders).  Level k bins the nodes of level k (level 0: the DOFs,
    each at its node's coordinates; a coarser node at the coordinates of its first member) into
    bins >> k cells per axis of the bounding box of all nodes; ids in first-appearance order;
    a level that merges nothing is skipped (the bins halve again).  by_component: the key of a
    node is dpn * bin + the component of its first member, so an aggregate holds one component
    only (unknown-based aggregation).  Returns the level maps."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    book = np.asarray(book, dtype=np.uint64).ravel()
    pos = np.nonzero(book != eod)[0]
    n_dof = pos.size
    node_of = np.empty(n_dof, dtype=np.int64)
    node_of[book[pos].astype(np.int64)] = pos // dpn
    comp_of = np.empty(n_dof, dtype=np.int64)
    comp_of[book[pos].astype(np.int64)] = pos % dpn
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    rep = np.arange(n_dof)
    cur, maps, k = n_dof, [], 0
    while cur > max_coarse or not maps:
        if len(maps) == max_levels:
            raise ValueError("coarsest level still too large")
        nb = np.maximum(1, np.asarray(bins, dtype=np.int64) >> min(k, 31))
        k += 1
        p = xyz[node_of[rep]]
        span = hi - lo
        t = np.where(span > 0, (p - lo) / np.where(span > 0, span, 1.0), 0.0)
        c = np.minimum(nb - 1, (t * nb).astype(np.uint64).astype(np.int64))
        b = c[:, 0] + nb[0] * (c[:, 1] + nb[1] * c[:, 2])
        if by_component:
            b = b * dpn + comp_of[rep]
        _, first, inv = np.unique(b, return_index=True, return_inverse=True)
        order = np.argsort(first)
        rank = np.empty_like(order)
        rank[order] = np.arange(order.size)
        if order.size == cur:
            if (nb == 1).all():
                raise ValueError("cannot coarsen further")
            continue
        maps.append(rank[inv.ravel()].astype(np.uint32))
        rep = rep[first[order]]
        cur = order.size
    return maps


def hierarchy(A, maps):
    """Level operators [A_0, A_1, ...] and prolongations for level maps (scipy sparse A)."""
    levels, Ps = [A.tocsr()], []
    for m in maps:
        P = prolongation(m, int(np.max(m)) + 1)
        Ps.append(P)
        levels.append((P.T @ levels[-1] @ P).tocsr())
    return levels, Ps


def gershgorin(A):
    A = A.tocsr()
    d = A.diagonal()
    s = np.asarray(abs(A).sum(axis=1)).ravel()
    ok = d != 0
    return float(np.max(s[ok] / np.abs(d[ok])))


def cheb_coefs(lam, steps=4):
    """(c1, c2) per step: d_k = c1 d_{k-1} + c2 D^-1 res_k over [lam/30, 1.05 lam]."""
    hi, lo = 1.05 * lam, lam / 30.0
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sig = theta / delta
    rho = 1.0 / sig
    out = [(0.0, 1.0 / theta)]
    for _ in range(1, steps):
        rn = 1.0 / (2.0 * sig - rho)
        out.append((rn * rho, 2.0 * rn / delta))
        rho = rn
    return out


def cheb_smooth(A, dinv, coefs, r, z, steps):
    """`steps` Chebyshev steps from z (z None: zero start)."""
    d = None
    for k in range(steps):
        c1, c2 = coefs[k]
        res = r if z is None else r - A @ z
        dd = c2 * dinv * res
        if k > 0:
            dd = dd + c1 * d
        d = dd
        z = dd if z is None else z + dd
    return z


def compressed(A):
    """D fp16(D^-1 A): the operator a level smooths with when it uses the FP16 copy of D^-1 A
    (values rounded to binary16, round to nearest even)."""
    import scipy.sparse as sp
    A = A.tocsr()
    dg = A.diagonal()
    S = sp.diags(1.0 / dg) @ A
    S = S.tocsr()
    S.data = S.data.astype(np.float16).astype(np.float64)
    return (sp.diags(dg) @ S).tocsr()


def vcycle(levels, Ps, r, pre, post, coarse, lams=None, level=0, fp16_levels=(), coarse_steps=None):
    """z = B r; coarse(Ac, r) solves the coarsest level; levels listed in fp16_levels smooth
    (and form the restricted residual) with compressed(A_l); coarse_steps = (pre, post) on the
    levels below the finest (None: the finest level's)."""
    A = levels[level]
    if level == len(levels) - 1:
        return coarse(A, r)
    if coarse_steps is not None and level > 0:
        pre, post = coarse_steps
    lam = gershgorin(A) if lams is None else lams[level]
    coefs = cheb_coefs(lam)
    dg = A.diagonal()
    dinv = np.where(dg != 0, 1.0 / np.where(dg != 0, dg, 1.0), 0.0)
    As = compressed(A) if level in fp16_levels else A
    z = cheb_smooth(As, dinv, coefs, r, None, pre) if pre > 0 else None
    res = r if z is None else r - As @ z
    e = vcycle(levels, Ps, Ps[level].T @ res, pre, post, coarse, lams, level + 1, fp16_levels, coarse_steps)
    z = Ps[level] @ e if z is None else z + Ps[level] @ e
    return cheb_smooth(As, dinv, coefs, r, z, post)


def coarse_exact(Ac, r):
    return np.linalg.solve(Ac.toarray(), r)


def coarse_jacobi(Ac, r):
    return r / Ac.diagonal()


def fcg(A, b, prec, tol=1e-8, maxit=100000):
    """Flexible CG, FCG(1) (Notay 2000): p = z - (z.q_prev / p_prev.q_prev) p_prev, alpha =
    p.r / p.q, stop when ||r|| / ||b|| < tol.  FP64 throughout; returns (x, iterations).  Pinned
    by its closed forms: B = A^-1 converges in one iteration, B = I is plain CG (finite
    termination on few distinct eigenvalues)."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    r = b.copy()
    nb = np.linalg.norm(b)
    if nb == 0:
        return x, 0
    p_prev = q_prev = None
    for k in range(1, maxit + 1):
        z = prec(r)
        p = z if p_prev is None else z - (z @ q_prev) / (p_prev @ q_prev) * p_prev
        q = A @ p
        a = (p @ r) / (p @ q)
        x = x + a * p
        r = r - a * q
        if np.linalg.norm(r) / nb < tol:
            return x, k
        p_prev, q_prev = p, q
    return x, maxit


# ------------------------------------------------------------------------ rigid-body aggregation
def rigid_modes(xyz, comp):
    """The six rigid-body modes at DOFs with coordinates xyz (m x 3, relative to a reference
    point) and components comp: translations x, y, z, then the rotations about z, x and y
    (u = w x r)."""
    m = xyz.shape[0]
    B = np.zeros((m, 6))
    B[np.arange(m), comp] = 1.0
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    B[:, 3] = np.where(comp == 0, -y, np.where(comp == 1, x, 0.0))
    B[:, 4] = np.where(comp == 1, -z, np.where(comp == 2, y, 0.0))
    B[:, 5] = np.where(comp == 2, -x, np.where(comp == 0, z, 0.0))
    return B


def frame_shift(p):
    """T (6 x 6) with modes(r + p) = modes(r) T: the rigid-body modes about the origin from those
    about the point p (translations unchanged, each rotation gains the translation p x w)."""
    T = np.eye(6)
    x0, y0, z0 = p
    T[0, 3], T[1, 3] = -y0, x0
    T[1, 4], T[2, 4] = -z0, y0
    T[0, 5], T[2, 5] = z0, -x0
    return T


def mgs2(B, tol=1e-10):
    """Modified Gram-Schmidt, every column orthogonalised twice against the previous live
    columns, in column order.  Returns Q (m x 6, dead columns zero), R (6 x 6 upper triangular,
    the rows of dead columns zero) and the dead flags: a column whose norm after the two passes
    is at most tol times its norm before (or zero) is dead."""
    m, k = B.shape
    Q = np.zeros((m, k))
    R = np.zeros((k, k))
    dead = np.zeros(k, dtype=bool)
    for c in range(k):
        v = B[:, c].copy()
        n0 = np.sqrt(v @ v)
        for _ in range(2):
            for cp in range(c):
                if dead[cp]:
                    continue
                d = Q[:, cp] @ v
                v = v - d * Q[:, cp]
                R[cp, c] += d
        nv = np.sqrt(v @ v)
        if n0 == 0.0 or nv <= tol * n0:
            dead[c] = True
        else:
            Q[:, c] = v / nv
            R[c, c] = nv
    return Q, R, dead


def rigid_hierarchy(A, xyz, book, bins, max_coarse, max_levels=16, eod=0xFFFFFFFF):
    """Smoothed-aggregation's tentative prolongations with the rigid-body modes (Vanek, Mandel,
    Brezina 1996), unsmoothed, for 3-DOF-per-node elasticity: the node aggregates of the nested
    bins (bins_hierarchy with the DOFs of a node in one aggregate, coarsest at most max_coarse /
    6 aggregates); each aggregate J has a reference point p_J (the node of its first fine DOF);
    its members' rigid-body modes about p_J (level 0: from the node coordinates; coarser levels:
    the near-null-space rows B times frame_shift(-p_J)) are orthonormalised by mgs2 over the
    members in row order, B_J = Q_J R_J; the 6 coarse rows of J (6 J + a) take the rows of
    R_J frame_shift(p_J), the modes about the origin again, as their near-null-space rows, so that
    all aggregates share one frame.  A dead column leaves a zero prolongation column and a unit
    diagonal on its coarse row.
    Returns the level operators [A_0, A_1, ...], the prolongations and the dead flags per level."""
    import scipy.sparse as sp
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    book = np.asarray(book, dtype=np.uint64).ravel()
    maps = bins_hierarchy(xyz, book, 3, bins, max(1, max_coarse // 6), max_levels, eod)
    pos = np.nonzero(book != eod)[0]
    n0 = pos.size
    node = np.empty(n0, dtype=np.int64)
    comp = np.empty(n0, dtype=np.int64)
    node[book[pos].astype(np.int64)] = pos // 3
    comp[book[pos].astype(np.int64)] = pos % 3
    levels, Ps, deads = [A.tocsr()], [], []
    Bl = None
    ref = xyz[node]          # reference point of every row of the current level
    n = n0
    for lvl, m in enumerate(maps):
        m = np.asarray(m, dtype=np.int64)
        rowmap = m if lvl == 0 else np.repeat(m, 6)
        na = int(m.max()) + 1
        rows, cols, vals = [], [], []
        Bn = np.zeros((6 * na, 6))
        dead_n = np.zeros(6 * na, dtype=bool)
        ref_n = np.zeros((6 * na, 3))
        order = np.argsort(rowmap, kind="stable")
        bounds = np.searchsorted(rowmap[order], np.arange(na + 1))
        for J in range(na):
            ids = order[bounds[J]:bounds[J + 1]]
            pJ = ref[ids[0]]
            if lvl == 0:
                B = rigid_modes(xyz[node[ids]] - pJ, comp[ids])
            else:
                B = Bl[ids] @ frame_shift(-pJ)
            Q, R, dead = mgs2(B)
            for a in range(6):
                rows.append(ids)
                cols.append(np.full(ids.size, 6 * J + a))
                vals.append(Q[:, a])
            Bn[6 * J:6 * J + 6] = R @ frame_shift(pJ)
            dead_n[6 * J:6 * J + 6] = dead
            ref_n[6 * J:6 * J + 6] = pJ
        P = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, 6 * na))
        P.eliminate_zeros()
        Ac = (P.T @ levels[-1] @ P).tolil()
        for r in np.nonzero(dead_n)[0]:
            Ac[r, r] = 1.0
        Ps.append(P.tocsr())
        levels.append(Ac.tocsr())
        deads.append(dead_n)
        Bl = Bn
        ref = ref_n
        n = 6 * na
    return levels, Ps, deads
