"""FP64 NumPy restatement of one PBF substep with frozen neighbour pairs, and
its adjoint: the checker of the library's adjoint kernels.

Test infrastructure, not product code.  The forward follows the step the
library runs (src/pbf/pbf_solver.cu:186-328 and the kernels it launches):

    v1 = v + g dt;  p = x + v1 dt                          integrate
    I x:  lambda_i = -C_i / (S_i + eps_l)                   computeLambda
          p_i += (m / rho0) sum_j (l_i + l_j + s_ij) G(p_i - p_j), then the
          collision fixed point                             Delta-p, collisions
    v2 = (p - x) / dt                                       velocity update
    v3_i = v2_i + c sum_j (v2_j - v2_i) W(p_i - p_j)        XSPH
    w_i = sum_j G(p_i - p_j) x (v3_j - v3_i)                vorticity
    v4_i = v3_i + dt e (N_i x w_i),  N = eta / |eta|        confinement
    v5 = contact velocities of walls and colliders          resolveVelocities

with W the poly6 kernel, G the spiky gradient and s_ij the artificial
pressure, each neighbour set given as directed pairs (i, j) of one grid
build (both directions present).  The collision solve projects against the
container, spheres, rounded boxes and planes; the contacts follow the
container's walls, then every sphere, box and plane.  The adjoint is the
exact reverse of this function with the pairs and every discrete decision
held fixed; test_adjoint.py pins it against central differences.
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
    spheres: tuple = ()  # ((cx, cy, cz), radius), in array order
    boxes: tuple = ()    # ((cx, cy, cz), (hx, hy, hz))
    planes: tuple = ()   # ((px, py, pz), (nx, ny, nz)), normal not normalised


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
    """out[k] = sum of vals[m] over idx[m] == k (np.bincount per column)."""
    if vals.ndim == 1:
        return np.bincount(idx, weights=vals, minlength=n).astype(np.float64)
    out = np.empty((n,) + vals.shape[1:])
    for c in range(vals.shape[1]):
        out[:, c] = np.bincount(idx, weights=vals[:, c], minlength=n)
    return out


def cross(a, b):
    return np.cross(a, b)


# ---------------------------------------------------------------------------
# Forward
# ---------------------------------------------------------------------------
WALLS = [(0, 1.0), (0, -1.0), (1, 1.0), (2, 1.0), (2, -1.0)]  # axis, normal sign
FLT_EPS = float(np.finfo(np.float32).eps)
CONTACT_EPS = 1e-4  # collisionEpsilon (src/pbf/collision_kernels.cu:254)


def _tol(q, o, radius):
    """projectionTolerance (src/pbf/collision_kernels.cu:9-14): 8 FLT_EPSILON
    times the largest of the radius and the coordinates of both points."""
    return 8.0 * FLT_EPS * np.maximum(radius, np.maximum(np.abs(q).max(1), np.abs(o).max(1)))


def _eye(n):
    return np.broadcast_to(np.eye(3), (n, 3, 3)).copy()


def _radial_jac(u, scale):
    """scale (I - u u^T), one per row."""
    return scale[:, None, None] * (np.eye(3)[None] - u[:, :, None] * u[:, None, :])


def _project(kind, k, q, P, rec=None):
    """One projection of CollisionSystem::solve (src/pbf/collision_kernels.cu:
    47-234) on every row of q: (projected q, rows it moved, their Jacobians,
    the record that replays it).  With rec, the record's decisions are
    applied instead of evaluated (the frozen function)."""
    n = len(q)
    r = P.radius
    J = _eye(n)
    out = q.copy()
    if kind == "container":
        lo = np.array(P.box_min) + r
        hi = np.array(P.box_max) - r
        c = q.copy()
        c[:, 0] = np.minimum(np.maximum(c[:, 0], lo[0]), hi[0])
        c[:, 1] = np.maximum(c[:, 1], lo[1])
        c[:, 2] = np.minimum(np.maximum(c[:, 2], lo[2]), hi[2])
        changed = (c != q) if rec is None else rec["changed"]
        vals = c if rec is None else rec["vals"]
        out = np.where(changed, vals, q)
        J = J * (~changed)[:, :, None]
        return out, changed.any(1), J, {"kind": kind, "k": k, "changed": changed, "vals": vals}
    if kind == "sphere":
        ctr, rad = P.spheres[k]
        ctr = np.asarray(ctr, float)
        md = rad + r
        o = q - ctr
        d2 = np.einsum("ij,ij->i", o, o)
        if rec is None:
            cond = (d2 < md * md) & (md - np.sqrt(d2) > _tol(q, np.broadcast_to(ctr, q.shape), md))
            pos, zero = cond & (d2 > 0), cond & (d2 == 0)
        else:
            pos, zero = rec["pos"], rec["zero"]
        ln = np.sqrt(np.where(pos, d2, 1.0))
        u = o / ln[:, None]
        out[pos] = ctr + u[pos] * md
        out[zero, 0] = ctr[0] + md
        J[pos] = _radial_jac(u[pos], md / ln[pos])
        J[zero, 0, :] = 0.0
        return out, pos | zero, J, {"kind": kind, "k": k, "pos": pos, "zero": zero}
    if kind == "box":
        ctr, he = (np.asarray(a, float) for a in P.boxes[k])
        mn, mx = ctr - he, ctr + he
        cp = np.minimum(np.maximum(q, mn), mx)
        if rec is None:
            inside = (q >= mn) & (q <= mx)
            o = q - cp
            d2 = np.einsum("ij,ij->i", o, o)
            A = (d2 > 0) & (d2 < r * r) & (r - np.sqrt(np.where(d2 > 0, d2, 1.0)) > _tol(q, cp, r))
            B = d2 == 0
            dd = np.minimum(q - mn, mx - q)
            axis = np.where((dd[:, 0] <= dd[:, 1]) & (dd[:, 0] <= dd[:, 2]), 0,
                            np.where(dd[:, 1] <= dd[:, 2], 1, 2))
            lower = np.take_along_axis(q - mn, axis[:, None], 1)[:, 0] < \
                np.take_along_axis(mx - q, axis[:, None], 1)[:, 0]
            val = np.where(lower, mn[axis] - r, mx[axis] + r)
            cpr = cp
        else:
            A, B, inside, axis, val, cpr = (rec[k_] for k_ in ("A", "B", "inside", "axis", "val", "cp"))
        cpf = np.where(inside, q, cpr)  # clamp decisions frozen
        o = q - cpf
        ln = np.sqrt(np.where(A, np.einsum("ij,ij->i", o, o), 1.0))
        u = o / ln[:, None]
        out[A] = cpf[A] + u[A] * r
        out[B, axis[B]] = val[B]
        JA = _radial_jac(u, r / ln) * (~inside)[:, None, :]
        JA = JA + inside[:, :, None] * np.eye(3)[None]
        J[A] = JA[A]
        J[B, axis[B], :] = 0.0
        return out, A | B, J, {"kind": kind, "k": k, "A": A, "B": B, "inside": inside,
                               "axis": axis, "val": val, "cp": cpr}
    # plane
    pt, nrm = (np.asarray(a, float) for a in P.planes[k])
    l2 = float(nrm @ nrm)
    if l2 == 0.0:
        return out, np.zeros(n, bool), J, {"kind": kind, "k": k, "m": np.zeros(n, bool)}
    nu = nrm / np.sqrt(l2)
    dist = (q - pt) @ nu
    m = ((dist < r) & (r - dist > _tol(q, np.broadcast_to(pt, q.shape), r))) if rec is None \
        else rec["m"]
    out[m] = q[m] + nu * (r - dist[m])[:, None]
    J[m] = np.eye(3) - np.outer(nu, nu)
    return out, m, J, {"kind": kind, "k": k, "m": m}


def _projections(P):
    return ([("container", 0)] + [("sphere", k) for k in range(len(P.spheres))] +
            [("box", k) for k in range(len(P.boxes))] + [("plane", k) for k in range(len(P.planes))])


def collide(p, P, events=None):
    """CollisionSystem::solve (src/pbf/collision_system.cu:133-172) in FP64:
    passes of container, spheres, boxes, planes until a pass moves nothing
    (at most 32; every row in lockstep, which a row already at its fixed
    point does not notice).  Returns (positions, the Jacobian of each row
    (n, 3, 3), the list of projections that acted); with `events` (that
    list from a base evaluation) the same projections act, in the same
    order, with their decisions held fixed."""
    n = len(p)
    q = p.copy()
    J = _eye(n)
    if events is not None:
        for rec in events:
            q, m, Jk, _ = _project(rec["kind"], rec["k"], q, P, rec)
            J[m] = Jk[m] @ J[m]
        return q, J, events
    events = []
    for _ in range(32):
        marked = np.zeros(n, bool)
        for kind, k in _projections(P):
            q, m, Jk, rec = _project(kind, k, q, P)
            if m.any():
                J[m] = Jk[m] @ J[m]
                events.append(rec)
            marked |= m
        if not marked.any():
            return q, J, events
    raise RuntimeError("Collision constraints did not converge")


def container(p, P):
    """The container projection alone (one pass): positions and, per
    coordinate, 1 where it was kept."""
    q, _, J, _ = _project("container", 0, p, P)
    return q, np.diagonal(J, axis1=1, axis2=2).copy()


def contact_candidates(xp, P, frozen=None):
    """The contacts of resolveVelocitiesKernel (src/pbf/collision_kernels.cu:
    236-371) in the reference's order: five container walls, then spheres,
    boxes, planes.  Each: the rows it touches, its normal per row, and what
    the normal's derivative needs (sphere and box normals follow the
    position).  With `frozen`, the base evaluation's touches and box
    branches are kept and only the normals are re-evaluated."""
    eps = CONTACT_EPS
    r = P.radius
    n = len(xp)
    lo = np.array(P.box_min) + r
    hi = np.array(P.box_max) - r
    out = []
    t0 = xp[:, 0] <= lo[0] + eps
    t3 = xp[:, 2] <= lo[2] + eps
    walls = [t0, ~t0 & (xp[:, 0] >= hi[0] - eps), xp[:, 1] <= lo[1] + eps, t3,
             ~t3 & (xp[:, 2] >= hi[2] - eps)]
    for w, (ax, sgn) in enumerate(WALLS):
        nrm = np.zeros((n, 3))
        nrm[:, ax] = sgn
        touch = walls[w] if frozen is None else frozen[len(out)]["touch"]
        out.append({"kind": "wall", "touch": touch, "n": nrm})
    for k, (ctr, rad) in enumerate(P.spheres):
        o = xp - np.asarray(ctr, float)
        d2 = np.einsum("ij,ij->i", o, o)
        cd = rad + r + eps
        touch = ((d2 > 0) & (d2 <= cd * cd)) if frozen is None else frozen[len(out)]["touch"]
        ln = np.sqrt(np.where(d2 > 0, d2, 1.0))
        out.append({"kind": "sphere", "touch": touch, "n": o / ln[:, None], "len": ln,
                    "outside": np.ones((n, 3), bool)})
    for k, (ctr, he) in enumerate(P.boxes):
        ctr, he = np.asarray(ctr, float), np.asarray(he, float)
        mn, mx = ctr - he, ctr + he
        rc = r + eps
        if frozen is None:
            inside = (xp >= mn) & (xp <= mx)
            cp = np.minimum(np.maximum(xp, mn), mx)
            o = xp - cp
            d2 = np.einsum("ij,ij->i", o, o)
            A = (d2 > 0) & (d2 <= rc * rc)
            B = d2 == 0
        else:
            f = frozen[len(out)]
            inside, cp, A, B = f["inside"], f["cp"], f["A"], f["B"]
        cpf = np.where(inside, xp, cp)
        o = xp - cpf
        ln = np.sqrt(np.where(A, np.einsum("ij,ij->i", o, o), 1.0))
        nrm = o / ln[:, None]
        dd = np.minimum(xp - mn, mx - xp)
        axis = np.where((dd[:, 0] <= dd[:, 1]) & (dd[:, 0] <= dd[:, 2]), 0,
                        np.where(dd[:, 1] <= dd[:, 2], 1, 2))
        if frozen is not None:
            axis = frozen[len(out)]["axis"]
        lower = np.take_along_axis(xp - mn, axis[:, None], 1)[:, 0] < \
            np.take_along_axis(mx - xp, axis[:, None], 1)[:, 0]
        if frozen is not None:
            lower = frozen[len(out)]["lower"]
        nb = np.zeros((n, 3))
        nb[np.arange(n), axis] = np.where(lower, -1.0, 1.0)
        nrm = np.where(A[:, None], nrm, nb)
        out.append({"kind": "box", "touch": A | B, "n": nrm, "len": ln, "outside": ~inside,
                    "A": A, "B": B, "inside": inside, "cp": cp, "axis": axis, "lower": lower,
                    "radial": A})
    for k, (pt, nrm) in enumerate(P.planes):
        pt, nrm = np.asarray(pt, float), np.asarray(nrm, float)
        l2 = float(nrm @ nrm)
        nu = nrm / np.sqrt(l2) if l2 > 0 else np.zeros(3)
        dist = (xp - pt) @ nu
        touch = ((dist <= r + eps) & (l2 > 0)) if frozen is None else frozen[len(out)]["touch"]
        out.append({"kind": "plane", "touch": touch, "n": np.broadcast_to(nu, (n, 3)).copy()})
    return out


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
    constraint iteration, then the post-solve build; or a function of the
    positions entering each build that returns its pairs (the pairs are
    then rebuilt at every build, as the reference does).  `frozen` (the aux of a base evaluation) holds the
    discrete decisions fixed as well: which coordinates the container
    projects, which walls a particle touches and which contacts act, and
    whether |eta| passes the normalisation threshold -- the function the
    adjoint differentiates."""
    g = np.array(P.gravity)
    v1 = v + g * P.dt
    p = x + v1 * P.dt
    aux = {"x": x, "v1": v1, "it": []}
    build = pairs if callable(pairs) else None
    iters = P.iterations if build is not None else len(pairs) - 1
    for k in range(iters):
        I, J = build(p) if build is not None else pairs[k]
        L = lambda_terms(p, I, J, P)
        dp = delta_p(p, L["lam"], I, J, P)
        ev = frozen["it"][k]["events"] if frozen is not None else None
        pn, Jc, ev = collide(p + dp, P, ev)
        aux["it"].append(dict(p=p, lam=L["lam"], J=Jc, events=ev, out=pn,
                              acted=np.any(Jc != np.eye(3), axis=(1, 2))))
        p = pn
    I, J = build(p) if build is not None else pairs[iters]
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
    cands = contact_candidates(p, P, frozen["contacts"] if frozen is not None else None)
    v5 = v4.copy()
    for k, cd in enumerate(cands):
        nrm = cd["n"]
        vn = np.einsum("ij,ij->i", v5, nrm)
        vinn = np.einsum("ij,ij->i", v1, nrm)
        target = np.where(vinn < 0.0, -P.restitution * vinn, 0.0)
        act = cd["touch"] & ~((vn > target) | ((vn == target) & (vinn >= 0.0)))
        vin_neg = vinn < 0.0
        if frozen is not None:
            act = frozen["contacts"][k]["act"]
            vin_neg = frozen["contacts"][k]["vin_neg"]
            target = np.where(vin_neg, -P.restitution * vinn, 0.0)
        cd.update(act=act, vin_neg=vin_neg, target=target, vin=v5.copy())
        tang = v5 - vn[:, None] * nrm
        v5 = np.where(act[:, None], target[:, None] * nrm + (1.0 - P.friction) * tang, v5)
    acts = np.stack([cd["act"] for cd in cands], 1)
    aux.update(p=p, v2=v2, v3=v3, w=w, wl=wl, eta=eta, el=el, N=N, ok=ok, v4=v4, acts=acts,
               contacts=cands)
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
    # contact velocities in reverse: v_out = t n + ts (v - (v.n) n) with
    # t = -e_r (v1.n) when v1.n < 0; sphere and box normals follow the
    # position (n = o / |o|), so their contacts also feed the positions.
    vcur = vb.copy()
    ts = 1.0 - P.friction
    for cd in reversed(aux["contacts"]):
        nrm, act, neg = cd["n"], cd["act"], cd["vin_neg"]
        vn_b = np.einsum("ij,ij->i", vcur, nrm)  # adjoint of t
        tang_b = ts * (vcur - vn_b[:, None] * nrm)
        v1b += np.where((act & neg)[:, None], (-P.restitution * vn_b)[:, None] * nrm, 0.0)
        if cd["kind"] in ("sphere", "box"):
            vin = cd["vin"]
            radial = act if cd["kind"] == "sphere" else act & cd["radial"]
            nbar = (np.where(neg, -P.restitution * vn_b, 0.0)[:, None] * v1
                    + cd["target"][:, None] * vcur - ts * vn_b[:, None] * vin
                    - ts * np.einsum("ij,ij->i", vin, nrm)[:, None] * vcur)
            g = (nbar - nrm * np.einsum("ij,ij->i", nrm, nbar)[:, None]) / cd["len"][:, None]
            pb += np.where(radial[:, None], g * cd["outside"], 0.0)
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
        pk, lam = it["p"], it["lam"]
        I, J = pairs[k]
        prb = np.einsum("nji,nj->ni", it["J"], pb)  # adjoint of p + dp before the projections
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
    return forward(x, v, lambda p: pairs_within(p, P.h), P, c, e)


def run_full(x, v, substeps, P, c, e):
    for _ in range(substeps):
        x, v = substep_full(x, v, P, c, e)
    return x, v

