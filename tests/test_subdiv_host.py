"""Host primitives of the C-ABI against the oracle and the mathematics (no
GPU needed): the Loop limit-surface evaluation of SubdivEvaluation
(SubdivEvaluation.C:14-187) at arbitrary barycentric points, pinned by brute
force (repeated generic Loop subdivision and the exact limit-point mask at
dyadic points, for an element at a valence-5 vertex and a regular one),
partition of unity, its own finite differences and continuity across the
sub-patch levels; and Friction::apply (Friction.C:4-60) against the oracle's
restatement on random cases and every branch."""
import numpy as np
import pytest

import oracle as O
import shell_amd as S
from conftest import imperfect


def _loop_subdivide(X, T):
    """one generic Loop step; faces e -> 4e..4e+3 = (a,ab,ca), (b,bc,ab), (c,ca,bc), (ab,bc,ca)"""
    nv = len(X)
    opp = {}
    for t in T:
        for k in range(3):
            opp[(t[k], t[(k + 1) % 3])] = t[(k + 2) % 3]
    nb = [set() for _ in range(nv)]
    for (a, b) in opp:
        nb[a].add(b)
    newX = []
    for v in range(nv):
        N = len(nb[v])
        beta = (5 / 8 - (3 / 8 + np.cos(2 * np.pi / N) / 4) ** 2) / N
        newX.append((1 - N * beta) * X[v] + beta * sum(X[q] for q in nb[v]))
    edges = {}
    for (a, b), c in opp.items():
        key = (min(a, b), max(a, b))
        if key in edges:
            continue
        newX.append(3 / 8 * (X[a] + X[b]) + 1 / 8 * (X[c] + X[opp[(b, a)]]))
        edges[key] = nv + len(edges)
    mid = lambda a, b: edges[(min(a, b), max(a, b))]
    T2 = []
    for a, b, c in T:
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        T2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    return np.array(newX), np.array(T2)


def _limit(X, T, v):
    """exact Loop limit position of vertex v (limit mask of its valence)"""
    nb = set()
    for t in T:
        for k in range(3):
            if t[k] == v:
                nb.add(t[(k + 1) % 3])
    N = len(nb)
    beta = (5 / 8 - (3 / 8 + np.cos(2 * np.pi / N) / 4) ** 2) / N
    chi = 1 / (3 / (8 * beta) + N)
    return (1 - N * chi) * X[v] + chi * sum(X[q] for q in nb)


@pytest.fixture(scope="module")
def coarse():
    X, T = O.icosphere(1)
    X = imperfect(X, 0.04, seed=5)
    o = O.Oracle(X, T)
    idx, tri, N = o.patches()
    o.close()
    return X, tri, idx, N


@pytest.mark.parametrize("valence", [3, 4, 5, 6])
def test_centroid_equals_the_element_basis(valence):
    w = S.subdiv_weights(valence, 1 / 3, 1 / 3)
    rb = S.raw_elem_basis(valence)
    assert np.allclose(w, rb[:3], rtol=0, atol=1e-14)


@pytest.mark.parametrize("valence", [3, 4, 5, 6])
def test_partition_of_unity_and_finite_differences(valence):
    rng = np.random.default_rng(valence)
    for _ in range(40):
        xi, eta = rng.random(2) * 0.5
        w = S.subdiv_weights(valence, xi, eta)
        assert abs(w[0].sum() - 1) < 1e-13 and np.all(np.abs(w[1:].sum(axis=1)) < 1e-11)
        h = 1e-7 * (xi + eta)
        n = int(np.floor(1 - np.log2(xi + eta)))
        # stay inside one sub-patch level for the differences
        if np.floor(1 - np.log2(xi + eta + 2 * h)) != n or np.floor(1 - np.log2(xi + eta - 2 * h)) != n:
            continue
        dx = (S.subdiv_weights(valence, xi + h, eta)[0] - S.subdiv_weights(valence, xi - h, eta)[0]) / (2 * h)
        de = (S.subdiv_weights(valence, xi, eta + h)[0] - S.subdiv_weights(valence, xi, eta - h)[0]) / (2 * h)
        scale = np.abs(w[1:]).max()
        assert np.abs(dx - w[1]).max() < 1e-5 * scale and np.abs(de - w[2]).max() < 1e-5 * scale


def test_continuous_across_sub_patch_levels():
    """across xi + eta = 2^-m the evaluation switches to one more
    subdivision; the surface and its tangents are continuous there"""
    for valence in (3, 5):
        for m in (1, 2, 5):
            s = 2.0 ** -m
            for frac in (0.2, 0.5, 0.9):
                a = S.subdiv_weights(valence, frac * s * (1 + 1e-12), (1 - frac) * s * (1 + 1e-12))
                b = S.subdiv_weights(valence, frac * s * (1 - 1e-12), (1 - frac) * s * (1 - 1e-12))
                assert np.abs(a[0] - b[0]).max() < 1e-10
                assert np.abs(a[1:] - b[1:]).max() < 1e-7 * 2.0 ** m


@pytest.mark.parametrize("kind", ["valence 5", "regular"])
def test_limit_points_at_dyadic_parameters_equal_brute_force(coarse, kind):
    """after m generic Loop steps, the fine vertex at the dyadic point (xi,
    eta) of an element has the exact limit position the evaluation gives at
    that point (one point per sub-patch level and per sub-triangle k)"""
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
            nf[4 * f] = np.array([a, ab, ca])
            nf[4 * f + 1] = np.array([b, bc, ab])
            nf[4 * f + 2] = np.array([c, ca, bc])
            nf[4 * f + 3] = np.array([ab, bc, ca])
        faces = nf
    vertex_at = {}
    for f, p in faces.items():
        for k in range(3):
            vertex_at[tuple(np.round(p[k] * 2 ** m).astype(int))] = Tf[f][k]
    pts = [(5, 1), (1, 6), (3, 3), (17, 2), (2, 24), (9, 9), (1, 1), (3, 0), (0, 12), (16, 16), (30, 1), (1, 0)]
    worst = 0
    for i, j in pts:
        xi, eta = i / 2 ** m, j / 2 ** m
        w = S.subdiv_weights(n - 6, xi, eta)
        L = _limit(Xf, Tf, vertex_at[(i, j)])
        worst = max(worst, np.linalg.norm(w[0] @ P - L))
    assert worst < 1e-13


def _oracle_friction(E, fc, fn, n, v, vn, v_stop, ks, mu_s, mu_d, damping, stick, so, cp):
    io = np.zeros(24)
    io[0], io[1:4], io[4], io[5:8], io[8:11], io[11] = E, fc, fn, n, v, vn
    io[12], io[13], io[14], io[15], io[16], io[17] = v_stop, ks, mu_s, mu_d, damping, stick
    io[18:21], io[21:24] = so, cp
    out = O.friction_apply(io)
    return out[0], out[1:4], int(out[17]), out[18:21]


def test_friction_apply_equals_the_oracle():
    """Friction::apply through the C-ABI: every output bit-identical to the
    oracle's restatement over random cases that visit stick creation,
    sticking, break-away and sliding (both compiled with one rounding per
    operation)"""
    rng = np.random.default_rng(11)
    branches = set()
    for _ in range(4000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v = rng.normal(size=3) * 10.0 ** rng.uniform(-5, -1)
        vn = float(v @ n)
        args = (float(rng.uniform(0, 1e-3)), rng.normal(size=3) * 1e-4, float(rng.uniform(0, 0.02)), n, v, vn,
                1e-3, float(rng.uniform(0.5, 50)), float(rng.choice([0.0, 0.3])), float(rng.uniform(0, 0.5)),
                float(rng.choice([0.0, 0.01])), int(rng.integers(0, 2)), rng.normal(size=3) * 1e-3,
                rng.normal(size=3) * 1e-3)
        a = S.friction_apply(*args)
        b = _oracle_friction(*args)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2] and np.array_equal(a[3], b[3])
        branches.add((args[11], a[2]))
    assert branches == {(0, 0), (0, 1), (1, 0), (1, 1)}
