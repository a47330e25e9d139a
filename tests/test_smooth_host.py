"""Smooth (limit-surface) shell-shell contact on the host, no GPU.

The oracle's limit surface at any point of an element (SubdivEvaluation.C:
14-187, here with second derivatives) is pinned by brute-force Loop
subdivision, by its own finite differences and against the library's
weights.  The closest-point search of SmoothShellShellContact::compute
(SmoothShellShellContact.C:21-204) minimises |p2 - p1|^2 over the two
triangles; the reference's minimiser (ConstrainedMinimizer.h) is not in its
sources, so the search is pinned to the constrained minimum it must find:
the triangle-triangle distance for planar patches (whose limit surface is
the triangle itself), brute-force sampling and the KKT conditions for curved
ones.  The library's host search equals the oracle's; copies of the oracle
with a planted slip fail these pins.
"""
import numpy as np
import pytest

import oracle as O
import shell_amd as S
from conftest import imperfect
from test_subdiv_host import _limit, _loop_subdivide

# lattice offsets of the 12 patch slots of a regular element (v0, its ring
# from v1 to v2, then (1,1), (2,0), (2,-1), (0,2), (-1,2))
LAT = [(0, 0), (1, 0), (1, -1), (0, -1), (-1, 0), (-1, 1), (0, 1), (1, 1), (2, 0), (2, -1), (0, 2), (-1, 2)]


def planar_patch(origin, a, b):
    return np.array([origin + i * a + j * b for i, j in LAT])


def seg_hits_tri(p, q, A, B, C):
    n = np.cross(B - A, C - A)
    dp, dq = np.dot(p - A, n), np.dot(q - A, n)
    if dp * dq > 0 or dp == dq:
        return False
    x = p + dp / (dp - dq) * (q - p)
    return all(np.dot(np.cross(V - U, x - U), n) >= 0 for U, V in ((A, B), (B, C), (C, A)))


def triangles_intersect(T1, T2):
    return (any(seg_hits_tri(T1[i], T1[(i + 1) % 3], *T2) for i in range(3)) or
            any(seg_hits_tri(T2[i], T2[(i + 1) % 3], *T1) for i in range(3)))


@pytest.fixture(scope="module")
def squashed():
    """an L3 sphere pancaked to 8 % of its height: its top and bottom sheets
    face each other at about 0.15, with curvature, over most elements"""
    X, T = O.icosphere(3)
    X = imperfect(X, 1e-4, seed=3)
    o = O.Oracle(X, T)
    idx, tri, N = o.patches()
    o.close()
    Xs = X.copy()
    Xs[:, 2] *= 0.08
    return Xs, tri, idx, N


def facing_pairs(sq, n, seed):
    """(valence, points) of n element pairs, one on each sheet, near each
    other in x, y"""
    Xs, tri, idx, N = sq
    cent = Xs[tri].mean(axis=1)
    top = np.where(cent[:, 2] > 0.03)[0]
    bot = np.where(cent[:, 2] < -0.03)[0]
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        e1 = rng.choice(top)
        d = np.linalg.norm(cent[bot, :2] - cent[e1, :2], axis=1)
        e2 = bot[np.argsort(d)[rng.integers(0, 6)]]
        out.append((N[e1] - 6, Xs[idx[e1, :N[e1]]], N[e2] - 6, Xs[idx[e2, :N[e2]]]))
    return out


def grad_f(v1, P1, v2, P2, x, handle=None):
    L1 = O.limit_point(v1, P1, x[0], x[1], handle)
    L2 = O.limit_point(v2, P2, x[2], x[3], handle)
    r = L2[0] - L1[0]
    g = 2 * np.array([-L1[1] @ r, -L1[2] @ r, L2[1] @ r, L2[2] @ r])
    scale = 2 * np.linalg.norm(r) * max(np.abs(L1[1:3]).max(), np.abs(L2[1:3]).max())
    return g, scale


def kkt_violation(g, x, scale, tol=1e-12):
    """largest violation of the KKT conditions of min f over the two
    triangles at x, relative to the gradient's scale: stationary along the
    free directions, nonnegative multipliers of the active constraints"""
    worst = 0.0
    for e in range(2):
        v, w = x[2 * e], x[2 * e + 1]
        gv, gw = g[2 * e], g[2 * e + 1]
        a0, a1, a2 = v <= tol, w <= tol, v + w >= 1 - tol
        if a0 and a1:
            viol = [-gv, -gw]
        elif a0 and a2:
            viol = [gw, gw - gv]           # lambda_u = -gw, lambda_v = gv - gw
        elif a1 and a2:
            viol = [gv, gv - gw]
        elif a0:
            viol = [abs(gw), -gv]
        elif a1:
            viol = [abs(gv), -gw]
        elif a2:
            viol = [abs(gv - gw), (gv + gw) / 2]
        else:
            viol = [abs(gv), abs(gw)]
        worst = max(worst, max(viol) / scale)
    return worst


def brute_min(v1, P1, v2, P2, x=None, radius=0.15, n=20, handle=None):
    """the smallest sampled f over both triangles (n + 1 points per edge) and
    the smallest within `radius` (barycentric, per element) of x"""
    g = np.array([(i / n, j / n) for i in range(n + 1) for j in range(n + 1 - i)])
    L1 = np.array([O.limit_point(v1, P1, a, b, handle)[0] for a, b in g])
    L2 = np.array([O.limit_point(v2, P2, a, b, handle)[0] for a, b in g])
    D = ((L1[:, None, :] - L2[None, :, :]) ** 2).sum(-1)
    if x is None:
        return D.min(), None
    near1 = np.linalg.norm(g - x[0:2], axis=1) <= radius
    near2 = np.linalg.norm(g - x[2:4], axis=1) <= radius
    return D.min(), D[np.ix_(near1, near2)].min() if near1.any() and near2.any() else np.inf


# ----------------------------------------------------------- limit surface

@pytest.fixture(scope="module")
def coarse():
    X, T = O.icosphere(1)
    X = imperfect(X, 0.04, seed=5)
    o = O.Oracle(X, T)
    idx, tri, N = o.patches()
    o.close()
    return X, tri, idx, N


def limit_brute_force_error(coarse, kind, handle=None):
    """the oracle's limit points at dyadic parameters of an element against
    the exact limit positions of the fine vertices after 5 generic Loop steps"""
    X, tri, idx, N = coarse
    e = int(np.argmax(N == 11)) if kind == "valence 5" else int(np.argmax(N == 12))
    n = int(N[e])
    P = X[idx[e, :n]]
    m = 5
    Xf, Tf = X.copy(), np.array(tri)
    faces = {e: np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])}
    for _ in range(m):
        Xf, Tf = _loop_subdivide(Xf, Tf)
        nf = {}
        for f, p in faces.items():
            a, b, c = p
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nf[4 * f], nf[4 * f + 1] = np.array([a, ab, ca]), np.array([b, bc, ab])
            nf[4 * f + 2], nf[4 * f + 3] = np.array([c, ca, bc]), np.array([ab, bc, ca])
        faces = nf
    vertex_at = {}
    for f, p in faces.items():
        for k in range(3):
            vertex_at[tuple(np.round(p[k] * 2 ** m).astype(int))] = Tf[f][k]
    pts = [(5, 1), (1, 6), (3, 3), (17, 2), (2, 24), (9, 9), (1, 1), (3, 0), (0, 12), (16, 16), (30, 1), (1, 0),
           (20, 3), (4, 21)]
    worst = 0.0
    for i, j in pts:
        L = O.limit_point(n - 6, P, i / 2 ** m, j / 2 ** m, handle)[0]
        worst = max(worst, np.linalg.norm(L - _limit(Xf, Tf, vertex_at[(i, j)])))
    return worst


@pytest.mark.parametrize("kind", ["valence 5", "regular"])
def test_oracle_limit_points_equal_brute_force(coarse, kind):
    assert limit_brute_force_error(coarse, kind) < 1e-13


def derivative_fd_error(valence, handle=None, seed=0):
    """relative error of the oracle's first and second derivatives against
    central differences of its points and first derivatives, inside one
    sub-patch (level and sub-triangle unchanged by the offsets)"""
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(60):
        X = rng.normal(size=(valence + 6, 3))
        xi, eta = rng.random(2)
        if xi + eta > 1:
            xi, eta = 1 - xi, 1 - eta
        s = xi + eta
        if valence != 6:
            n = int(np.floor(1 - np.log2(s)))
            sc = 2.0 ** (n - 1)
            a, b = xi * sc, eta * sc
            # keep away from the sub-patch borders (level and k)
            if min(abs(a - 0.5), abs(b - 0.5), abs(a + b - 0.5), abs(a + b - 1), s - 2.0 ** -n) < 0.02 * 2.0 ** -n \
                    or s < 0.05:
                continue
        h = 1e-6 * min(s, 1 - s, xi + 1e-3, eta + 1e-3)
        L = O.limit_point(valence, X, xi, eta, handle)
        Lxp, Lxm = O.limit_point(valence, X, xi + h, eta, handle), O.limit_point(valence, X, xi - h, eta, handle)
        Lep, Lem = O.limit_point(valence, X, xi, eta + h, handle), O.limit_point(valence, X, xi, eta - h, handle)
        d1 = np.array([(Lxp[0] - Lxm[0]) / (2 * h), (Lep[0] - Lem[0]) / (2 * h)])
        d2 = np.array([(Lxp[1] - Lxm[1]) / (2 * h), (Lep[2] - Lem[2]) / (2 * h), (Lxp[2] - Lxm[2]) / (2 * h)])
        worst = max(worst, np.abs(d1 - L[1:3]).max() / np.abs(L[1:3]).max(),
                    np.abs(d2 - L[3:6]).max() / max(np.abs(L[3:6]).max(), 1e-3 * np.abs(L[1:3]).max()))
    return worst


@pytest.mark.parametrize("valence", [3, 4, 5, 6])
def test_oracle_limit_derivatives_by_finite_differences(valence):
    assert derivative_fd_error(valence) < 1e-6


@pytest.mark.parametrize("valence", [3, 4, 5, 6])
def test_oracle_limit_surface_equals_the_library_weights(valence):
    """points to rounding everywhere; tangents wherever they are well
    conditioned: near an irregular vertex the parametric tangents vanish
    like (2 lambda)^n (lambda the subdominant eigenvalue: 1/4, 3/8, 0.452 for
    valence 3, 4, 5) while rounding grows like 2^n, so they are compared for
    xi + eta >= 1e-3 (valence 3, 4) and 1e-5 (valence 5)"""
    rng = np.random.default_rng(valence)
    for _ in range(300):
        X = rng.normal(size=(valence + 6, 3))
        xi, eta = rng.random(2)
        if xi + eta > 1:
            xi, eta = 1 - xi, 1 - eta
        if rng.random() < 0.3:
            sc = 10.0 ** rng.uniform(-9, 0)
            xi, eta = xi * sc, eta * sc
        o = O.limit_point(valence, X, xi, eta)
        w = S.subdiv_weights(valence, xi, eta) @ X
        assert np.abs(o[0] - w[0]).max() < 1e-14 * (1 + np.abs(X).max())
        if xi + eta >= {3: 1e-3, 4: 1e-3, 5: 1e-5, 6: 0.0}[valence]:
            assert np.abs(o[1:3] - w[1:3]).max() < 1e-9 * np.abs(w[1:3]).max()


# ---------------------------------------------------------- closest points

def planar_errors(handle=None, library=False, n=200, seed=5):
    """relative d2 error against the triangle distance over separate pairs,
    the number of crossing pairs and their largest d2 (patch size about 1)"""
    rng = np.random.default_rng(seed)
    worst, crossings, cross_d2 = 0.0, 0, 0.0
    for _ in range(n):
        P1 = planar_patch(rng.normal(size=3), rng.normal(size=3), rng.normal(size=3))
        P2 = planar_patch(rng.normal(size=3) * 1.5, rng.normal(size=3), rng.normal(size=3))
        T1, T2 = P1[[0, 1, 6]], P2[[0, 1, 6]]
        r = S.smooth_closest(6, P1, 6, P2) if library else O.smooth_closest(6, P1, 6, P2, handle=handle)
        if triangles_intersect(T1, T2):
            crossings += 1
            cross_d2 = max(cross_d2, r["d2"])
            continue
        t = O.tri_distance(np.concatenate([T1.reshape(-1), T2.reshape(-1)]))
        worst = max(worst, abs(r["d2"] - t[0]) / t[0])
    return worst, crossings, cross_d2


@pytest.mark.parametrize("who", ["oracle", "library"])
def test_planar_patches_give_the_triangle_distance(who):
    """a regular patch of a planar lattice has the flat triangle as its limit
    surface, so the smooth distance is the triangle-triangle distance
    (ShellShellContact::compute) of separate triangles, and zero for
    crossing ones (which the triangle test, built from edge-edge and
    vertex-face distances, cannot see)"""
    worst, crossings, cross_d2 = planar_errors(library=(who == "library"))
    assert crossings > 0 and cross_d2 < 1e-24
    assert worst < 1e-11


def curved_errors(sq, handle=None, n=30, seed=1):
    """over n facing pairs of the pancaked sphere: the KKT violation at the
    search's minimiser, how far its d2 lies above the sampled minimum near
    it (a local minimum, not a saddle: <= 0) and the fraction of pairs whose
    d2 reaches the global sampled minimum (two far elements can have several
    local minima, at different vertices; a search from the centroids finds
    one of them)"""
    over, kkt, hits = 0.0, 0.0, 0
    for v1, P1, v2, P2 in facing_pairs(sq, n, seed):
        r = O.smooth_closest(v1, P1, v2, P2, handle=handle)
        glob, near = brute_min(v1, P1, v2, P2, r["x"], handle=handle)
        over = max(over, (r["d2"] - near) / r["d2"])
        hits += r["d2"] <= glob * (1 + 1e-12)
        g, scale = grad_f(v1, P1, v2, P2, r["x"], handle)
        kkt = max(kkt, kkt_violation(g, r["x"], scale))
    return over, kkt, hits / n


def test_curved_closest_points_are_the_constrained_minimum(squashed):
    over, kkt, global_share = curved_errors(squashed)
    assert kkt < 1e-8 and over <= 1e-12 and global_share >= 0.8


def test_library_search_equals_the_oracle(squashed):
    """the library's host search (the code the contact kernels run) and the
    oracle's restatement: the same minimiser and distance from the centroid
    start and from a perturbed warm start"""
    rng = np.random.default_rng(2)
    iters = []
    for v1, P1, v2, P2 in facing_pairs(squashed, 150, 4):
        a = S.smooth_closest(v1, P1, v2, P2)
        b = O.smooth_closest(v1, P1, v2, P2)
        iters.append(a["iterations"])
        assert np.abs(a["x"] - b["x"]).max() < 1e-10 and abs(a["d2"] - b["d2"]) <= 1e-12 * b["d2"]
        x0 = np.clip(b["x"] + rng.normal(size=4) * 0.02, 0, 1)
        for e in range(2):
            if x0[2 * e] + x0[2 * e + 1] > 1:
                x0[2 * e: 2 * e + 2] *= (1 - 1e-12) / (x0[2 * e] + x0[2 * e + 1])
        aw = S.smooth_closest(v1, P1, v2, P2, x0)
        bw = O.smooth_closest(v1, P1, v2, P2, x0)
        assert np.abs(aw["x"] - bw["x"]).max() < 1e-10 and abs(aw["d2"] - bw["d2"]) <= 1e-12 * bw["d2"]
        assert abs(bw["d2"] - b["d2"]) <= 1e-9 * b["d2"]
    assert max(iters) < 40


# ---------------------------------------------------------- planted slips

SLIPS = {
    "edge point weights": ("  auto edge = [&](int a, int b, int o1, int o2) { return (c[a] + c[b]) * (3.0 / 8.0) + "
                           "(c[o1] + c[o2]) * (1.0 / 8.0); };",
                           "  auto edge = [&](int a, int b, int o1, int o2) { return (c[a] + c[b]) * (5.0 / 16.0) + "
                           "(c[o1] + c[o2]) * (3.0 / 16.0); };"),
    "P_0 index": ("  const int sel[3][12] = {{2, 0, N + 3, 1, N, N + 8, N + 2, N + 1, N + 4, N + 7, N + 6, N + 9},\n"
                  "                          {N + 9",
                  "  const int sel[3][12] = {{2, 0, N + 3, 1, N, N + 7, N + 2, N + 1, N + 4, N + 8, N + 6, N + 9},\n"
                  "                          {N + 9"),
    "k = 1 derivative sign": ("jfac = -std::pow(2.0, n); }", "jfac = std::pow(2.0, n); }"),
    "vertex multiplier sign": ("      else if (act[e] == 5) { lam[2] = -gw; lam[0] = gv - gw; }",
                               "      else if (act[e] == 5) { lam[2] = gw; lam[0] = gv - gw; }"),
    "edge stationarity": ("      else if (act[e] == 4) lam[2] = -0.5 * (gv + gw);",
                          "      else if (act[e] == 4) lam[2] = 0.5 * (gv + gw);"),
}


@pytest.mark.parametrize("slip", list(SLIPS))
def test_smooth_pins_catch_planted_slips(slip, coarse, squashed):
    from oracle_mutants import build
    old, new = SLIPS[slip]
    L = build(old, new, "smooth" + str(abs(hash(slip)) % 1000))
    caught = limit_brute_force_error(coarse, "valence 5", L) > 1e-9
    caught |= max(derivative_fd_error(v, L) for v in (3, 5)) > 1e-5
    if not caught:
        worst, _, cross_d2 = planar_errors(handle=L)
        caught |= worst > 1e-9 or cross_d2 > 1e-20
    if not caught:
        over, kkt, global_share = curved_errors(squashed, handle=L)
        caught |= over > 1e-9 or kkt > 1e-6 or global_share < 0.8
    assert caught
