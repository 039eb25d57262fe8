"""FP64 NumPy restatement of one PBF substep with frozen neighbour pairs, and
its adjoint: the checker of the library's adjoint kernels.

Test infrastructure, not product code.  The forward follows the step the
library runs (src/pbf/pbf_solver.cu:186-328 and the kernels it launches):

    v1 = v + g dt;  p = x + v1 dt                          integrate
    I x:  lambda_i = -C_i / (S_i + eps_l)                   computeLambda
          p_i += (m / rho0) sum_j (l_i + l_j + s_ij) G(p_i - p_j), then the
          container projection                              Delta-p, collisions
    v2 = (p - x) / dt                                       velocity update
    v3_i = v2_i + c sum_j (v2_j - v2_i) W(p_i - p_j)        XSPH
    w_i = sum_j G(p_i - p_j) x (v3_j - v3_i)                vorticity
    v4_i = v3_i + dt e (N_i x w_i),  N = eta / |eta|        confinement
    v5 = contact velocities of the container walls          resolveVelocities

with W the poly6 kernel, G the spiky gradient and s_ij the artificial
pressure, each neighbour set given as directed pairs (i, j) of one grid
build (both directions present).  Only the container is modelled (no
spheres, boxes or planes).  The adjoint is the exact reverse of this
function with the pairs held fixed; test_adjoint.py pins it against
central differences.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Params:
    dt: float           # substep duration
    h: float
    mass: float
    rho0: float
    eps_lambda: float
    gravity: tuple
    scorr_k: float
    scorr_n: int
    scorr_dq: float
    radius: float
    box_min: tuple
    box_max: tuple
    restitution: float = 0.0
    friction: float = 0.0
    iterations: int = 4


# ---------------------------------------------------------------------------
# Kernels and their derivatives (include/pbf/sph_helpers.cuh:13-95)
# ---------------------------------------------------------------------------
def poly6(d, P):
    h2 = P.h * P.h
    d2 = np.einsum("ij,ij->i", d, d)
    norm = 315.0 / (64.0 * np.pi) / P.h ** 3
    q = np.where(d2 < h2, (h2 - d2) / h2, 0.0)
    return norm * q ** 3


def poly6_grad(d, P):
    """dW/dd."""
    h2 = P.h * P.h
    d2 = np.einsum("ij,ij->i", d, d)
    norm = 315.0 / (64.0 * np.pi) / P.h ** 3
    q = np.where(d2 < h2, (h2 - d2) / h2, 0.0)
    return (norm * 3.0 * q ** 2 * (-2.0 / h2))[:, None] * d


def spiky(d, P):
    """G(d) = A (1 - r/h)^2 d / r for 0 < r < h, A = -45 / (pi h^4)."""
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    A = -45.0 / np.pi / P.h ** 4
    ok = (r > 0) & (r < P.h)
    rs = np.where(ok, r, 1.0)
    s = np.where(ok, A * (1.0 - rs / P.h) ** 2 / rs, 0.0)
    return s[:, None] * d


def spiky_jac_t(d, P, a):
    """J_G(d)^T a (J_G is symmetric) for vectors a, one per pair."""
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    A = -45.0 / np.pi / P.h ** 4
    ok = (r > 0) & (r < P.h)
    rs = np.where(ok, r, 1.0)
    u = d / rs[:, None]
    w = 1.0 - rs / P.h
    ua = np.einsum("ij,ij->i", u, a)
    # J = A [ w^2 (I - u u^T) / r - (2 / h) w u u^T ]
    out = A * ((w ** 2 / rs)[:, None] * (a - ua[:, None] * u) - (2.0 / P.h * w * ua)[:, None] * u)
    return np.where(ok[:, None], out, 0.0)


def scorr(d, P):
    h2 = P.h * P.h
    dq2 = P.scorr_dq ** 2
    d2 = np.einsum("ij,ij->i", d, d)
    if P.scorr_k == 0.0 or P.scorr_n <= 0 or dq2 >= h2:
        return np.zeros(len(d))
    q = np.where(d2 < h2, (h2 - d2) / (h2 - dq2), 0.0)
    return -P.scorr_k * q ** (3 * P.scorr_n)


def scorr_grad(d, P):
    h2 = P.h * P.h
    dq2 = P.scorr_dq ** 2
    d2 = np.einsum("ij,ij->i", d, d)
    if P.scorr_k == 0.0 or P.scorr_n <= 0 or dq2 >= h2:
        return np.zeros_like(d)
    den = h2 - dq2
    q = np.where(d2 < h2, (h2 - d2) / den, 0.0)
    k3n = 3 * P.scorr_n
    return (-P.scorr_k * k3n * q ** (k3n - 1) * (-2.0 / den))[:, None] * d


def scatter(n, idx, vals):
    out = np.zeros((n,) + vals.shape[1:])
    np.add.at(out, idx, vals)
    return out


def cross(a, b):
    return np.cross(a, b)


# ---------------------------------------------------------------------------
# Forward
# ---------------------------------------------------------------------------
def container(p, P):
    lo = np.array(P.box_min) + P.radius
    hi = np.array(P.box_max) - P.radius
    q = p.copy()
    q[:, 0] = np.minimum(np.maximum(q[:, 0], lo[0]), hi[0])
    q[:, 1] = np.maximum(q[:, 1], lo[1])
    q[:, 2] = np.minimum(np.maximum(q[:, 2], lo[2]), hi[2])
    # d(out)/d(in) per coordinate: 1 where the coordinate was kept
    keep = (q == p).astype(float)
    return q, keep


WALLS = [(0, 1.0), (0, -1.0), (1, 1.0), (2, 1.0), (2, -1.0)]  # axis, normal sign


def contacts(xp, P):
    """Walls each particle touches, in the reference's order."""
    eps = 1e-4
    lo = np.array(P.box_min) + P.radius
    hi = np.array(P.box_max) - P.radius
    t = np.zeros((len(xp), 5), bool)
    t[:, 0] = xp[:, 0] <= lo[0] + eps
    t[:, 1] = ~t[:, 0] & (xp[:, 0] >= hi[0] - eps)
    t[:, 2] = xp[:, 1] <= lo[1] + eps
    t[:, 3] = xp[:, 2] <= lo[2] + eps
    t[:, 4] = ~t[:, 3] & (xp[:, 2] >= hi[2] - eps)
    return t


def lambda_terms(p, I, J, P):
    n = len(p)
    d = p[I] - p[J]
    W = poly6(d, P)
    G = spiky(d, P)
    gs = P.mass / P.rho0
    rho = P.mass * 315.0 / (64.0 * np.pi) / P.h ** 3 + scatter(n, I, P.mass * W)
    C = rho / P.rho0 - 1.0
    g = -gs * G
    gsum = scatter(n, I, g)  # sum_j g_ij
    S = scatter(n, I, np.einsum("ij,ij->i", g, g)) + np.einsum("ij,ij->i", gsum, gsum)
    lam = -C / (S + P.eps_lambda)
    return dict(d=d, W=W, G=G, g=g, gsum=gsum, C=C, S=S, lam=lam)


def delta_p(p, lam, I, J, P):
    n = len(p)
    d = p[I] - p[J]
    G = spiky(d, P)
    s = scorr(d, P)
    coef = lam[I] + lam[J] + s
    return (P.mass / P.rho0) * scatter(n, I, coef[:, None] * G)


def forward(x, v, pairs, P, c, e, keep_aux=False, frozen=None):
    """One substep.  pairs: P.iterations + 1 (I, J) arrays, one per
    constraint iteration, then the post-solve build.  `frozen` (the aux of a base evaluation) holds the
    discrete decisions fixed as well: which coordinates the container
    projects, which walls a particle touches and which contacts act, and
    whether |eta| passes the normalisation threshold -- the function the
    adjoint differentiates."""
    g = np.array(P.gravity)
    v1 = v + g * P.dt
    p = x + v1 * P.dt
    aux = {"x": x, "v1": v1, "it": []}
    iters = len(pairs) - 1
    for k in range(iters):
        I, J = pairs[k]
        L = lambda_terms(p, I, J, P)
        dp = delta_p(p, L["lam"], I, J, P)
        pn, keep = container(p + dp, P)
        if frozen is not None:
            fk = frozen["it"][k]
            keep = fk["keep"]
            pn = np.where(keep > 0, p + dp, fk["out"])
        aux["it"].append(dict(p=p, lam=L["lam"], keep=keep, out=pn))
        p = pn
    I, J = pairs[iters]
    n = len(p)
    d = p[I] - p[J]
    W = poly6(d, P)
    G = spiky(d, P)
    v2 = (p - x) / P.dt
    v3 = v2 + c * scatter(n, I, (v2[J] - v2[I]) * W[:, None])
    u = v3[J] - v3[I]
    w = scatter(n, I, cross(G, u))
    wl = np.linalg.norm(w, axis=1)
    eta = scatter(n, I, (wl[J] - wl[I])[:, None] * G)
    el = np.linalg.norm(eta, axis=1)
    ok = el > 1e-6 if frozen is None else frozen["ok"]
    N = np.where(ok[:, None], eta / np.where(ok, el, 1.0)[:, None], 0.0)
    v4 = v3 + P.dt * e * cross(N, w)
    t = contacts(p, P) if frozen is None else frozen["touch"]
    v5 = v4.copy()
    acts = np.zeros_like(t)
    for k, (ax, sgn) in enumerate(WALLS):
        nrm = np.zeros(3)
        nrm[ax] = sgn
        vn = v5 @ nrm
        vinn = v1 @ nrm
        target = np.where(vinn < 0.0, -P.restitution * vinn, 0.0)
        act = t[:, k] & ~((vn > target) | ((vn == target) & (vinn >= 0.0)))
        if frozen is not None:
            act = frozen["acts"][:, k]
            target = np.where(frozen["vin_neg"][:, k], -P.restitution * vinn, 0.0)
        tang = v5 - vn[:, None] * nrm
        v5 = np.where(act[:, None], target[:, None] * nrm + (1.0 - P.friction) * tang, v5)
        acts[:, k] = act
    vin_neg = np.stack([(v1 @ np.eye(3)[ax]) * sgn < 0.0 for ax, sgn in WALLS], 1)
    aux.update(p=p, v2=v2, v3=v3, w=w, wl=wl, eta=eta, el=el, N=N, ok=ok, v4=v4, acts=acts,
               touch=t, vin_neg=vin_neg)
    return (p, v5, aux) if keep_aux else (p, v5)


# ---------------------------------------------------------------------------
# Adjoint
# ---------------------------------------------------------------------------
def backward(aux, pairs, P, c, e, xb, vb):
    """Given the adjoints of the substep's outputs (xb of positions, vb of
    velocities), the adjoints of its inputs x, v and of c, e."""
    x, v1, p = aux["x"], aux["v1"], aux["p"]
    n = len(x)
    iters = len(aux["it"])
    I, J = pairs[iters]
    d = p[I] - p[J]
    W = poly6(d, P)
    G = spiky(d, P)
    v2, v3, w, wl, N, ok, el = aux["v2"], aux["v3"], aux["w"], aux["wl"], aux["N"], aux["ok"], aux["el"]
    pb = xb.copy()
    v1b = np.zeros_like(v1)
    # contact velocities, walls in reverse
    vcur = vb.copy()
    for k in reversed(range(5)):
        ax, sgn = WALLS[k]
        nrm = np.zeros(3)
        nrm[ax] = sgn
        act = aux["acts"][:, k]
        # v_out = target n + ts (v - (v.n) n), target = -e_r (vin.n) when vin.n < 0
        ts = 1.0 - P.friction
        vn_b = vcur @ nrm
        tang_b = ts * (vcur - vn_b[:, None] * nrm)
        vinn = v1 @ nrm
        tb = vn_b  # adjoint of target
        v1b += np.where((act & (vinn < 0.0))[:, None], (-P.restitution * tb)[:, None] * nrm, 0.0)
        vcur = np.where(act[:, None], tang_b, vcur)
    v4b = vcur
    # confinement: v4 = v3 + dt e (N x w)
    v3b = v4b.copy()
    cb = P.dt * e * v4b
    eb = P.dt * np.sum(cross(N, w) * v4b)
    Nb = cross(w, cb)
    wb = cross(cb, N)
    etab = np.where(ok[:, None], (Nb - N * np.einsum("ij,ij->i", N, Nb)[:, None]) /
                    np.where(ok, el, 1.0)[:, None], 0.0)
    # eta_i = sum_j (|w_j| - |w_i|) G(d_ij)
    a = wl[J] - wl[I]
    gdot = np.einsum("ij,ij->i", etab[I], G)
    wlb = scatter(n, J, gdot) - scatter(n, I, gdot)
    db = a[:, None] * spiky_jac_t(d, P, etab[I])
    wsafe = np.where(wl > 0, wl, 1.0)
    wb += np.where((wl > 0)[:, None], wlb[:, None] * w / wsafe[:, None], 0.0)
    # vorticity: w_i = sum_j G(d_ij) x (v3_j - v3_i)
    u = v3[J] - v3[I]
    ub = cross(wb[I], G)          # adjoint of u for c = G x u:  u_b = w_b x G
    Gb = cross(u, wb[I])          # G_b = u x w_b
    v3b += scatter(n, J, ub) - scatter(n, I, ub)
    db += spiky_jac_t(d, P, Gb)
    # XSPH: v3_i = v2_i + c sum_j (v2_j - v2_i) W_ij
    v2b = v3b.copy()
    t = c * W[:, None] * v3b[I]
    v2b += scatter(n, J, t) - scatter(n, I, t)
    dv = v2[J] - v2[I]
    cb_x = np.sum(v3b[I] * dv * W[:, None])
    db += (c * np.einsum("ij,ij->i", dv, v3b[I]))[:, None] * poly6_grad(d, P)
    # velocity update v2 = (p - x) / dt
    pb += v2b / P.dt
    xb0 = -v2b / P.dt
    pb += scatter(n, I, db) - scatter(n, J, db)
    # constraint iterations in reverse
    gs = P.mass / P.rho0
    for k in reversed(range(iters)):
        it = aux["it"][k]
        pk, lam, keep = it["p"], it["lam"], it["keep"]
        I, J = pairs[k]
        prb = keep * pb  # adjoint of p + dp before the projection
        pkb = prb.copy()
        d = pk[I] - pk[J]
        G = spiky(d, P)
        s = scorr(d, P)
        coef = lam[I] + lam[J] + s
        # dp_i = gs sum_j coef_ij G_ij
        q = gs * np.einsum("ij,ij->i", prb[I], G)
        lamb = scatter(n, I, q) + scatter(n, J, q)
        db = gs * coef[:, None] * spiky_jac_t(d, P, prb[I]) + q[:, None] * scorr_grad(d, P)
        # lambda_i = -C_i / (S_i + eps)
        L = lambda_terms(pk, I, J, P)
        den = L["S"] + P.eps_lambda
        Cb = -lamb / den
        Sb = lamb * L["C"] / den ** 2
        db += (Cb[I] * gs)[:, None] * poly6_grad(d, P)
        gb = 2.0 * Sb[I][:, None] * (L["g"] + L["gsum"][I])
        db += -gs * spiky_jac_t(d, P, gb)
        pkb += scatter(n, I, db) - scatter(n, J, db)
        pb = pkb
    # integrate: p = x + v1 dt, v1 = v + g dt
    xb_out = xb0 + pb
    v1b += pb * P.dt
    return xb_out, v1b, cb_x, eb


# ---------------------------------------------------------------------------
# Full FP64 forward: neighbour pairs rebuilt at every grid build
# ---------------------------------------------------------------------------
def pairs_within(p, h):
    """Directed pairs (i, j), i != j, with |p_i - p_j|^2 <= h^2 (the radius
    test of findNeighbors; the 27-cell stencil holds every such pair when
    the cell size is h)."""
    from scipy.spatial import cKDTree

    und = cKDTree(p).query_pairs(h * (1.0 + 1e-12), output_type="ndarray")
    if len(und) == 0:
        return np.zeros(0, int), np.zeros(0, int)
    d = p[und[:, 0]] - p[und[:, 1]]
    und = und[np.einsum("ij,ij->i", d, d) <= h * h]
    I = np.concatenate([und[:, 0], und[:, 1]])
    J = np.concatenate([und[:, 1], und[:, 0]])
    return I, J


def substep_full(x, v, P, c, e):
    """One substep in FP64 with the pairs of every build taken from the
    positions entering it: the function central differences see."""
    g = np.array(P.gravity)
    p = x + (v + g * P.dt) * P.dt
    pairs = []
    for _ in range(P.iterations):
        pairs.append(pairs_within(p, P.h))
        L = lambda_terms(p, *pairs[-1], P)
        p, _ = container(p + delta_p(p, L["lam"], *pairs[-1], P), P)
    pairs.append(pairs_within(p, P.h))
    return forward(x, v, pairs, P, c, e)


def run_full(x, v, substeps, P, c, e):
    for _ in range(substeps):
        x, v = substep_full(x, v, P, c, e)
    return x, v

