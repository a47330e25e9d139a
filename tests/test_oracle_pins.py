"""Pins of the CPU FP64 oracle against what the mathematics and the
reference's own data fix (the reference ships no tests or golden vectors;
its only data are icosasphere3.off and icosasphere4.off).  Each test is
chosen so that a plausible slip (a wrong sign, a swapped index, a dropped
term) fails it."""
import hashlib

import numpy as np
import pytest

import oracle as O
from conftest import imperfect

H = 0.01
PCR = 2 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))

# SHA-256 of the reference's mesh files icosasphere3.off / icosasphere4.off
OFF_SHA = {3: "461df12da867577434b5f4db6d261d4ac66749fa1e3666df9a07814fd7a45033",
           4: "7a0dbb5a8c29168e4c0feac280debc9304cf20065e52390aeb5129e304025622"}


def off_text(level):
    X, T = O.icosphere(level, round_g=False)
    lines = ["OFF", f"{len(X)} {len(T)} 0"] + ["%g %g %g" % tuple(x) for x in X] + \
            ["3 %d %d %d" % tuple(t) for t in T]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("level", [3, 4])
def test_icosphere_reproduces_reference_off(level):
    """the generator writes the reference's mesh files byte for byte"""
    assert hashlib.sha256(off_text(level).encode()).hexdigest() == OFF_SHA[level]
    X, T = O.icosphere(level)
    assert len(X) == 10 * 4 ** level + 2 and len(T) == 20 * 4 ** level


def test_box_splines_partition_and_vertex_masks():
    rng = np.random.default_rng(0)
    for _ in range(50):
        v, w = rng.random(2)
        if v + w > 1:
            v, w = 1 - v, 1 - w
        b = O.stam_basis(v, w)
        assert abs(b[0].sum() - 1) < 1e-14
        assert np.all(np.abs(b[1:].sum(axis=1)) < 1e-13)
        assert np.all(b[0] >= -1e-16)
    # corners of the triangle: Loop's regular limit mask 1/2 centre, 1/12 ring
    for (v, w), centre in (((0, 0), 3), ((1, 0), 6), ((0, 1), 7)):
        b = O.stam_basis(v, w)[0]
        assert abs(b[centre] - 0.5) < 1e-15
        assert np.sum(np.abs(b - 1 / 12) < 1e-15) == 6


def test_box_splines_refine_with_loop_masks():
    """B(x) = S^T B'(fine x): the box splines are refinable with Loop's edge
    (3/8, 3/8, 1/8, 1/8) and regular vertex (5/8, 1/16) masks"""
    S = O.subdiv_rows(6)
    cvi = [3, 6, 2, 0, 1, 4, 7, 10, 9, 5, 11, 8]
    for xi, eta in ((1 / 3, 1 / 3), (0.4, 0.3), (0.45, 0.2), (0.26, 0.3), (0.5, 0.25)):
        coarse = O.stam_basis(xi, eta)[:, cvi]
        fine = O.stam_basis(1 - 2 * xi, 1 - 2 * eta)
        fac = np.array([1, -2, -2, 4, 4, 4])[:, None]
        np.testing.assert_allclose(fac * (fine @ S), coarse, atol=2e-15)


@pytest.mark.parametrize("valence", [3, 4, 5, 6])
def test_patch_basis_partition_of_unity(valence):
    b = O.raw_elem_basis(valence)
    assert abs(b[0].sum() - 1) < 1e-14
    assert np.all(np.abs(b[1:].sum(axis=1)) < 1e-13)


# ---------------------------------------------------------------- limit surface by brute force

def _loop_subdivide(X, T):
    """one step of Loop subdivision on a closed mesh (generic: any valence)"""
    nv = len(X)
    edges = {}
    opp = {}
    for t in T:
        for k in range(3):
            a, b, c = t[k], t[(k + 1) % 3], t[(k + 2) % 3]
            opp[(a, b)] = c
    nb = [set() for _ in range(nv)]
    for (a, b) in opp:
        nb[a].add(b)
    newX = []
    for v in range(nv):
        N = len(nb[v])
        beta = (5 / 8 - (3 / 8 + np.cos(2 * np.pi / N) / 4) ** 2) / N
        newX.append((1 - N * beta) * X[v] + beta * sum(X[q] for q in nb[v]))
    idx = nv
    for (a, b), c in opp.items():
        key = (min(a, b), max(a, b))
        if key in edges:
            continue
        d = opp[(b, a)]
        newX.append(3 / 8 * (X[a] + X[b]) + 1 / 8 * (X[c] + X[d]))
        edges[key] = idx
        idx += 1
    mid = lambda a, b: edges[(min(a, b), max(a, b))]
    T2 = []
    for a, b, c in T:
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        T2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    return np.array(newX), np.array(T2), nb


def _limit_point(X, nb, v):
    N = len(nb[v])
    beta = (5 / 8 - (3 / 8 + np.cos(2 * np.pi / N) / 4) ** 2) / N
    chi = 1 / (3 / (8 * beta) + N)
    return (1 - N * chi) * X[v] + chi * sum(X[q] for q in nb[v])


def test_patch_basis_matches_repeated_subdivision(icosphere):
    """value and tangents at the element centroid from the oracle's patch
    basis versus k levels of generic Loop subdivision (the centroid stays the
    centroid of the central sub-triangle): errors shrink by about 4 per
    level, for an element at a valence-5 vertex and a regular one"""
    X0, T0 = icosphere(1)
    X0 = imperfect(X0, 0.03, seed=1)
    o = O.Oracle(X0, T0)
    idx, tri, N = o.patches()
    tab = o.elem_tables()
    for e in (int(np.argmax(N == 11)), int(np.argmax(N == 12))):
        n = N[e]
        rb = O.raw_elem_basis(n - 6)
        P = X0[idx[e, :n]]
        val, b1, b2 = rb[0] @ P, rb[1] @ P, rb[2] @ P
        assert np.allclose(val, tab[e, 98:101], atol=1e-14)
        X, T = X0.copy(), np.array(tri)
        face = e
        params = np.array([[0, 0], [1, 0], [0, 1]], dtype=float)
        errs = []
        for level in range(1, 6):
            X, T, nb = _loop_subdivide(X, T)
            face = 4 * face + 3
            params = np.array([(params[0] + params[1]) / 2, (params[1] + params[2]) / 2,
                               (params[2] + params[0]) / 2])
            if level < 3:
                continue
            # limit points of the three corners of the nested central triangle
            nbs = [set() for _ in range(len(X))]
            for t in T:
                for k in range(3):
                    nbs[t[k]].add(t[(k + 1) % 3])
            L = np.array([_limit_point(X, nbs, v) for v in T[face]])
            M = np.column_stack([params[1] - params[0], params[2] - params[0]])
            D = np.column_stack([L[1] - L[0], L[2] - L[0]])
            grad = D @ np.linalg.inv(M)
            errs.append((L.mean(axis=0), grad[:, 0], grad[:, 1]))
        (c3, g13, g23), (c4, g14, g24), (c5, g15, g25) = errs
        # corner average: O(4^-k); 3-point tangent fit: O(2^-k)
        e_val = [np.linalg.norm(c - val) for c in (c3, c4, c5)]
        e_tan = [np.linalg.norm(g - b1) for g in (g13, g14, g15)]
        assert 3.0 < e_val[0] / e_val[1] < 5.0 and 3.0 < e_val[1] / e_val[2] < 5.0, e_val
        assert e_tan[2] < e_tan[1] < e_tan[0], e_tan
        # Richardson extrapolation from levels 4 and 5; the central triangle
        # turns by 180 degrees each level, so the first-order tangent error
        # alternates in sign
        assert np.linalg.norm((4 * c5 - c4) / 3 - val) < 2e-6
        assert np.linalg.norm((2 * g15 + g14) / 3 - b1) < min(1e-4, e_tan[0] / 5)
        assert np.linalg.norm((2 * g25 + g24) / 3 - b2) < 1e-4


# ------------------------------------------------------------------------- forces

@pytest.fixture(scope="module")
def l3(icosphere):
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=3)
    return X, T, O.Oracle(X, T, pressure_final=-2 * PCR)


def test_reference_frame_and_mass(l3):
    X, T, o = l3
    tab = o.elem_tables()
    A1, A2 = tab[:, 72:75], tab[:, 75:78]
    # reinit_mod: A1 is horizontal (east), A2 in the meridian plane, A1 . A2 = 0
    assert np.abs(A1[:, 2]).max() < 1e-12
    assert np.abs((A1 * A2).sum(1)).max() < 1e-12
    # JxW = |A1 x A2| / 2 before and after the reparametrisation
    np.testing.assert_allclose(0.5 * np.linalg.norm(np.cross(A1, A2), axis=1), tab[:, 93], rtol=1e-12)
    # outward normal
    xyz = tab[:, 98:101]
    assert np.all((np.cross(A1, A2) * xyz).sum(1) > 0)
    Minv, C = o.mass()
    assert abs((1 / Minv[0::3]).sum() - H * tab[:, 93].sum()) < 1e-14
    np.testing.assert_allclose(C, 0.05 / Minv, rtol=1e-14)


def test_internal_force_invariants(l3):
    X, T, o = l3
    n = 3 * len(X)
    assert np.abs(o.internal_force(np.zeros(n))).max() == 0
    assert np.abs(o.internal_force(np.tile([0.01, -0.02, 0.03], len(X)))).max() < 1e-17
    th = 0.05
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    assert np.abs(o.internal_force((X @ R.T - X).reshape(-1))).max() < 1e-16
    rng = np.random.default_rng(0)
    u = rng.normal(size=n) * 1e-3
    f = o.internal_force(u)
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-18 * 0 + 1e-17 * np.abs(f).max() * len(X)
    # f = -dE/du (central difference of E_tens + E_bend)
    d = rng.normal(size=n) * 1e-3
    eps = 1e-4

    def energy(x):
        _, _, en, _ = o.forces(x, None, 0.0)
        return en[0] + en[1]
    dE = (energy(u + eps * d) - energy(u - eps * d)) / (2 * eps)
    assert abs(dE + (f * d).sum()) < 1e-7 * abs(dE)


def test_membrane_and_bending_split(icosphere):
    """bending energy scales with h^3 and membrane with h under the same
    displacement (K ~ h, D ~ h^3)"""
    X, T = icosphere(3)
    rng = np.random.default_rng(2)
    u = rng.normal(size=3 * len(X)) * 1e-4
    e = []
    for h in (0.01, 0.02):
        o = O.Oracle(X, T, thickness=h)
        e.append(o.forces(u, None, 0.0)[2])
    assert abs(e[1][0] / e[0][0] - 2) < 1e-12
    assert abs(e[1][1] / e[0][1] - 8) < 1e-12


def test_orthotropic_reduces_to_isotropic_pressurised_sphere(icosphere):
    """E1 = E2, G12 = E/(2(1+nu)) and an inflating pressure: the uniform radial
    displacement that balances the load agrees with the membrane closed form
    w = p R^2 (1-nu) / (2 E h) (R = radius of the limit surface)"""
    X, T = icosphere(4)
    E, nu, p = 1.0, 0.3, 1e-5
    o = O.Oracle(X, T, E1=E, E2=E, G12=E / (2 * (1 + nu)), nu=nu, pressure_final=p)
    tab = o.elem_tables()
    R = np.linalg.norm(tab[:, 98:101], axis=1).mean()
    nhat = (X / np.linalg.norm(X, axis=1)[:, None]).reshape(-1)
    f0 = o.forces(np.zeros(3 * len(X)), None, 0.0)[0]
    w1 = 1e-4
    f1 = o.forces(w1 * nhat, None, 0.0)[0]
    k = (f1 - f0) / w1
    w = -np.dot(f0, k) / np.dot(k, k)
    w_closed = p * R ** 2 * (1 - nu) / (2 * E * H)
    assert abs(w - w_closed) < 0.02 * w_closed


# ---------------------------------------------------------------- narrow phase / contact

def _dense_tri_distance(P0, P1, n=40):
    g = np.array([(i / n, j / n) for i in range(n + 1) for j in range(n + 1 - i)])
    A = P0[0] + g[:, :1] * (P0[1] - P0[0]) + g[:, 1:] * (P0[2] - P0[0])
    B = P1[0] + g[:, :1] * (P1[1] - P1[0]) + g[:, 1:] * (P1[2] - P1[0])
    d = ((A[:, None, :] - B[None, :, :]) ** 2).sum(-1)
    return d.min()


def _seg_hits_tri(p, q, T):
    """Moller-Trumbore: does segment pq cross triangle T"""
    e1, e2, d = T[1] - T[0], T[2] - T[0], q - p
    h = np.cross(d, e2)
    a = e1 @ h
    if abs(a) < 1e-14:
        return False
    s = p - T[0]
    u = (s @ h) / a
    qq = np.cross(s, e1)
    v = (d @ qq) / a
    t = (e2 @ qq) / a
    return 0 <= u <= 1 and 0 <= v and u + v <= 1 and 0 <= t <= 1


def _intersect(A, B):
    return any(_seg_hits_tri(X[i], X[(i + 1) % 3], Y) for X, Y in ((A, B), (B, A)) for i in range(3))


def test_triangle_distance_against_dense_sampling():
    """for disjoint triangles the minimum over 9 edge-edge and 6 vertex-face
    distances is the triangle distance (interpenetrating pairs are the
    reference's CONTACT_PROBLEM case and are excluded)"""
    rng = np.random.default_rng(4)
    tested = 0
    for k in range(120):
        P = rng.normal(size=(2, 3, 3))
        if k % 3 == 1:
            P[1] = P[0][[1, 2, 0]] + rng.normal(size=(3, 3)) * 0.3
        if _intersect(P[0], P[1]):
            continue
        tested += 1
        out = O.tri_distance(P.reshape(-1))
        d2, dc, v, w = out[0], out[1:4], out[4:6], out[6:8]
        # the reported closest points are on the triangles and dc joins them
        q0 = P[0][0] + v[0] * (P[0][1] - P[0][0]) + w[0] * (P[0][2] - P[0][0])
        q1 = P[1][0] + v[1] * (P[1][1] - P[1][0]) + w[1] * (P[1][2] - P[1][0])
        assert min(v.min(), w.min(), (1 - v - w).min()) > -1e-12
        np.testing.assert_allclose(q1 - q0, dc, atol=1e-12)
        assert abs(d2 - dc @ dc) < 1e-12
        dense = _dense_tri_distance(P[0], P[1])
        assert d2 <= dense + 1e-12          # never above a sampled distance
        assert dense - d2 < 0.05 * (np.abs(P).max() ** 2)  # and close to it
    assert tested >= 60


def test_segment_and_point_triangle_closed_forms():
    # parallel segments: distance of the lines
    out = O.segment_distance([0, 0, 0, 1, 0, 0, 0.5, 1, 0, 1.5, 1, 0])
    assert abs(np.linalg.norm(out[:3]) - 1) < 1e-15
    # crossing skew segments
    out = O.segment_distance([-1, 0, 0, 1, 0, 0, 0, -1, 2, 0, 1, 2])
    np.testing.assert_allclose(out[:3], [0, 0, 2])
    np.testing.assert_allclose(out[3:], [0.5, 0.5])
    # point above the interior of a triangle
    out = O.point_triangle([0.2, 0.3, 1.5, 0, 0, 0, 1, 0, 0, 0, 1, 0])
    np.testing.assert_allclose(out, [0, 0, -1.5, 0.2, 0.3], atol=1e-15)
    # point beyond vertex b
    out = O.point_triangle([2, -1, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0])
    np.testing.assert_allclose(out, [-1, 1, 0, 1, 0], atol=1e-15)


def test_contact_force_is_energy_derivative():
    k, h = 0.7, 0.01
    for d in (0.002, 0.005, 0.0099):
        eps = 1e-9
        dE = (O.contact_energy(k, h, d + eps) - O.contact_energy(k, h, d - eps)) / (2 * eps)
        assert abs(-dE - O.contact_force(k, h, d)) < 1e-6 * abs(O.contact_force(k, h, d))
    assert O.contact_force(k, h, h) == 0 and O.contact_energy(k, h, h) == 0


def _fric(f_normal, n, v, stick=0, origin=(0, 0, 0), cp=(0, 0, 0), mu_s=0.3, mu_d=0.2, ks=10.0,
          damping=0.0, v_stop=1e-3):
    n = np.asarray(n, float)
    v = np.asarray(v, float)
    io = np.zeros(24)
    io[4] = f_normal
    io[5:8] = n
    io[8:11] = v
    io[11] = v @ n
    io[12], io[13], io[14], io[15], io[16] = v_stop, ks, mu_s, mu_d, damping
    io[17] = stick
    io[18:21] = origin
    io[21:24] = cp
    out = O.friction_apply(io)
    return out[0], out[1:4], out[17], out[18:21]


def test_friction_branches():
    # slip: tangential speed above v_stop -> Coulomb, opposite to v_t, |f| = mu_d f_n
    E, f, stick, org = _fric(2.0, [0, 0, 1], [0.1, 0, 0.5])
    assert stick == 0 and E == 0 and np.allclose(org, 0)
    np.testing.assert_allclose(f, [-0.2 * 2.0 * 0.1 / (0.1 + 1e-11), 0, 0])
    # stick creation: spring origin so that the force is continuous with slip
    vt = np.array([5e-4, 0, 0])
    E, f, stick, org = _fric(2.0, [0, 0, 1], vt, cp=[1, 2, 3])
    assert stick == 1
    np.testing.assert_allclose(f, -0.2 * 2.0 * vt / (5e-4 + 1e-11), rtol=1e-12)
    np.testing.assert_allclose(E, 0.5 * 10 * ((0.2 * 2.0 / (10 * (5e-4 + 1e-11))) * 5e-4) ** 2, rtol=1e-12)
    # sticking spring, normal component of the spring removed
    E, f, stick, org = _fric(2.0, [0, 0, 1], [0, 0, 0], stick=1, origin=[0.01, 0, 0.5], cp=[0.02, 0, 0])
    assert stick == 1
    np.testing.assert_allclose(f, [-10 * 0.01, 0, 0], atol=1e-15)
    # break-away: spring force beyond mu_s f_n -> slip with the dynamic force
    E, f, stick, org = _fric(0.1, [0, 0, 1], [1e-4, 0, 0], stick=1, origin=[0, 0, 0], cp=[1.0, 0, 0])
    assert stick == 0 and E == 0
    np.testing.assert_allclose(f, [-0.2 * 0.1 * 1e-4 / (1e-4 + 1e-11), 0, 0])


def test_broad_phase_equals_brute_force(icosphere):
    """linked cells at a fresh rebuild give exactly the O(ne^2) set
    {e1 < e2, not topologically close, boxes within d_max, d2 <= d_max^2}"""
    X, T = icosphere(3)
    o = O.Oracle(X, T, contact=1, cell_size_mult=1.0)
    u = X.copy()
    u[:, 2] *= 0.006
    u = (u - X).reshape(-1)
    o.set_state(u)
    n = o.rebuild(force=True)
    e1, e2 = o.pairs()
    _, tri, _ = o.patches()
    P = X + u.reshape(-1, 3)
    bmin = P[tri].min(1)
    bmax = P[tri].max(1)
    avg = (bmax - bmin).max(1)
    avg = sum(avg[i:i + 256].sum() for i in range(0, len(avg), 256)) / len(avg)
    cell = 1.0 * avg
    tol = 0.02 * cell
    dmax = H + tol
    start, ids = o.close_sets()
    close = set()
    for e in range(len(T)):
        for c in ids[start[e]:start[e + 1]]:
            close.add((e, int(c)))
    cand = []
    ne = len(T)
    for a in range(ne):
        b = np.arange(a + 1, ne)
        apart = np.any((bmin[a] > bmax[b] + dmax) | (bmin[b] > bmax[a] + dmax), axis=1)
        for bb in b[~apart]:
            if (a, int(bb)) in close:
                continue
            d2 = O.tri_distance(np.concatenate([P[tri[a]].reshape(-1), P[tri[bb]].reshape(-1)]))[0]
            if d2 <= dmax * dmax:
                cand.append((a, int(bb)))
    assert n == len(cand) > 0
    assert list(zip(e1.tolist(), e2.tolist())) == cand


def test_energy_conservation_undamped(icosphere):
    """central difference (beta = 0), no damping, no load: the total energy of
    a small free vibration stays constant to O(dt^2)"""
    X, T = icosphere(3)
    o = O.Oracle(X, T, damping_alpha=0.0, dt=0.05)
    rng = np.random.default_rng(1)
    u0 = rng.normal(size=3 * len(X)) * 1e-5
    o.set_state(u0)
    o.init_dynamics()
    E = []
    for _ in range(40):
        r = o.step(25)
        E.append(r["e_kin"] + r["e_tens"] + r["e_bend"])
    E = np.array(E)
    assert (E.max() - E.min()) / E.mean() < 2e-2


def test_enclosed_volume_is_the_polyhedral_volume(icosphere):
    """compute_enclosed_volume (ShellSystem.C:2505-2541): |sum of det(x0,x1,x2)|/6
    over the control triangles minus half the thickness layer; the first term
    equals the polyhedron volume from the divergence theorem, close to 4 pi/3"""
    X, T = icosphere(3)
    h = 0.01
    o = O.Oracle(X, T, thickness=h, dt=1e-6, dt_max=1e-6)
    o.set_state(np.zeros(X.size))
    o.init_dynamics()
    r = o.step(1)
    a, b, c = X[T[:, 0]], X[T[:, 1]], X[T[:, 2]]
    poly = np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0
    # the same volume as a sum over faces of (centroid . area vector) / 3
    cen = (a + b + c) / 3.0
    poly2 = np.sum(np.einsum("ij,ij->i", cen, 0.5 * np.cross(b - a, c - a))) / 3.0
    assert abs(poly - poly2) < 1e-12
    assert abs(r["volume"] + r["area"] * h / 2 - poly) < 1e-12 * poly
    assert 0.98 * 4 * np.pi / 3 < poly < 4 * np.pi / 3
    # area: one-point quadrature of the limit surface, which lies inside the
    # control net (about 1% below 4 pi at level 3)
    assert abs(r["area"] - 4 * np.pi) < 0.02 * 4 * np.pi


def test_displacement_trigger_is_safe_and_rebuilds_less(icosphere):
    """rebuild trigger 1 (twice the largest node displacement since the list's
    build reaches cell_margin x cell_size) against the reference's du_sum
    trigger, on a squashed level-3 shell whose sheets close (no friction):
    every pair within h of the current positions (brute force) is listed at
    every checkpoint, the trajectories are identical (pairs beyond h carry no
    force) and trigger 1 rebuilds less often"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=7)
    kw = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, adapt_dt_every=1, contact=1, cell_size_mult=1.0,
              mu_s=0.0, mu_d=0.0)
    Y = X.copy()
    Y[:, 2] *= 0.03
    u = (Y - X).reshape(-1)
    v = np.zeros_like(X)
    v[:, 2] = -0.02 * np.sign(X[:, 2])
    runs = {}
    for trig in (0, 1):
        o = O.Oracle(X, T, rebuild_trigger=trig, **kw)
        o.set_state(u, v.reshape(-1))
        o.init_dynamics()
        _, tri, _ = o.patches()
        start, ids = o.close_sets()
        close = set()
        for e in range(len(T)):
            for c in ids[start[e]:start[e + 1]]:
                close.add((e, int(c)))
        contacts = 0
        for k in range(8):
            r = o.step(25)
            if trig == 1 and k % 2 == 1:
                P = X + o.get_state()[0].reshape(-1, 3)
                bmin, bmax = P[tri].min(1), P[tri].max(1)
                listed = set(zip(*[a.tolist() for a in o.pairs()]))
                for a in range(len(T)):
                    b = np.arange(a + 1, len(T))
                    apart = np.any((bmin[a] > bmax[b] + H) | (bmin[b] > bmax[a] + H), axis=1)
                    for bb in b[~apart]:
                        if (a, int(bb)) in close:
                            continue
                        d2 = O.tri_distance(np.concatenate([P[tri[a]].reshape(-1), P[tri[bb]].reshape(-1)]))[0]
                        if d2 <= H * H:
                            contacts += 1
                            assert (a, int(bb)) in listed, (k, a, int(bb), d2)
        runs[trig] = (o.get_state()[0], r["n_rebuilds"], contacts)
    assert runs[1][2] > 0 and np.all(np.isfinite(runs[1][0]))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert 0 < runs[1][1] < runs[0][1], (runs[0][1], runs[1][1])


@pytest.mark.parametrize("band", ["above", "below", "clamp"])
def test_adaptive_dt_known_answer(icosphere, band):
    """the step control of advance (ShellSystem.C:1027-1047) on a motion
    whose local-error estimate is known in closed form: a rigid translation
    v0 under mass-proportional damping, from the consistent acceleration
    a0 = -c v0 (c = M^-1 C per DOF, no elastic or external force).  One step
    of the scheme gives v_pred = v0 + (1 - gamma) dt a0 and a1 = -c v_pred, so
    ||a1 - a0||_inf = max c^2 (1 - gamma) dt |v0| and the estimate
    rel = dt^2 |beta - 1/6| ||a1 - a0||_inf / h grows as dt^3.  Hence: above
    the band the attempt is rejected once, and the retry at
    dt1 = dt (sqrt(min max) / rel)^(1/3) lands on the band's centre and is
    accepted; below the band the step is accepted at once and dt grows to the
    same formula.  A wrong power of dt, exponent, factor or the target would
    change the rejection count or dt1.  With contact, the velocity clamp of
    the step (dt <= h / (20 v_max)) is checked the same way."""
    X, T = icosphere(2)
    alpha, beta, gamma, dt0 = 0.5, 0.0, 0.5, 0.1
    v0 = np.zeros_like(X)
    v0[:, 2] = 0.01                       # dt |v0| < h / 2: no large-increment problem
    v0 = v0.reshape(-1)
    probe = O.Oracle(X, T, damping_alpha=alpha)
    Minv, C = probe.mass()
    probe.close()
    c = Minv * C
    rel0 = dt0 ** 2 * abs(beta - 1 / 6) * np.max(c ** 2 * (1 - gamma) * dt0 * np.abs(v0)) / H
    if band == "clamp":
        # with contact the step is also bounded by dt v_max 2 * 10 <= h (contact_resolution
        # 10, v_max = max |v_pred|); with gamma = 1, v_pred = v0 whatever dt, so the first
        # attempt is rejected and the retry at h / (20 max |v0|) is accepted with that dt
        o = O.Oracle(X, T, damping_alpha=alpha, newmark_beta=beta, newmark_gamma=1.0, dt=dt0,
                     adapt_dt_every=1, rel_loc_err_min=rel0 / 20, rel_loc_err_max=rel0 / 10, contact=1)
        o.set_state(np.zeros_like(v0), v0, -c * v0, 0.0)
        r = o.step(1)
        o.close()
        dt1 = H / (20 * np.max(np.abs(v0)))
        assert r["n_rejected"] == 1 and r["problem"] == 0
        assert abs(r["t"] - dt1) <= 1e-12 * dt1 and abs(r["dt"] - dt1) <= 1e-12 * dt1
        return
    lo, hi = (rel0 / 20, rel0 / 10) if band == "above" else (10 * rel0, 20 * rel0)
    o = O.Oracle(X, T, damping_alpha=alpha, newmark_beta=beta, newmark_gamma=gamma, dt=dt0,
                 adapt_dt_every=1, rel_loc_err_min=lo, rel_loc_err_max=hi)
    o.set_state(np.zeros_like(v0), v0, -c * v0, 0.0)
    r = o.step(1)
    dt1 = dt0 * (np.sqrt(lo * hi) / rel0) ** (1 / 3)
    assert r["n_rejected"] == (1 if band == "above" else 0)
    assert abs(r["dt"] - dt1) <= 1e-9 * dt1
    assert abs(r["t"] - (dt1 if band == "above" else dt0)) <= 1e-9 * dt0
    assert r["problem"] == 0
    o.close()


def test_rigid_free_residual_is_the_least_squares_remainder(icosphere):
    """the deformation residual (the force minus its least-squares fit by the
    rigid-body inertial forces m_i (t + w x x_i) at the force-evaluation
    positions; an extension with no counterpart in the reference) equals the
    remainder of a dense least-squares solve of the same force, on a strongly
    imperfect shell under pressure whose force has a rigid part (about 1 %
    of its norm, from the discrete pressure resultant)"""
    X, T = icosphere(2)
    X = imperfect(X, 0.2, seed=3)
    dt = 0.01
    o = O.Oracle(X, T, pressure_final=-1.5 * PCR, dt=dt)
    u = np.random.default_rng(0).normal(size=X.size) * 1e-3
    z = np.zeros(X.size)
    o.set_state(u, z, z, 0.0)
    r = o.step(1)                          # v = a = 0: the force is evaluated at u, t = dt
    f, fext, _, _ = o.forces(u, t=dt)
    Minv, _ = o.mass()
    o.close()
    m = 1 / Minv.reshape(-1, 3)[:, 0]
    x = X + u.reshape(-1, 3)
    B = np.zeros((X.size, 6))
    for i in range(len(X)):
        xi = x[i]
        cross = np.array([[0, -xi[2], xi[1]], [xi[2], 0, -xi[0]], [-xi[1], xi[0], 0]])
        B[3 * i:3 * i + 3, :3] = m[i] * np.eye(3)
        B[3 * i:3 * i + 3, 3:] = -m[i] * cross   # w x x_i = -[x_i]x w
    q, *_ = np.linalg.lstsq(B, f, rcond=None)
    res = np.linalg.norm(f - B @ q)
    fn = np.linalg.norm(f)
    assert fn - res > 1e-5 * fn            # the fit removes a material rigid part
    assert abs(r["fext_norm"] - np.linalg.norm(fext)) <= 1e-12 * r["fext_norm"]
    assert abs(r["residual_deform"] * r["fext_norm"] - res) <= 1e-12 * fn


def test_close_sets_are_the_cutoff_closure(icosphere):
    """close sets (build_topological_neighborhood, ShellSystem.C:684-741) by
    their defining properties, on a coarse mesh whose cutoff (1 + margin) h
    reaches several rings out (so the search must iterate): each holds every element sharing a node and
    not the element itself; every other member lies within the cutoff of the
    element (triangle distance on the undeformed mesh) and shares a node
    with another member; and no element that shares a node with a member and
    lies within the cutoff is missing"""
    X, T = icosphere(2)
    h, margin = 0.4, 0.4
    o = O.Oracle(X, T, thickness=h, neighbor_margin=margin, contact=1)
    start, ids = o.close_sets()
    o.close()
    cutoff = (1 + margin) * h
    node_elems = [[] for _ in range(len(X))]
    for e, t in enumerate(T):
        for n in t:
            node_elems[n].append(e)
    nbr = [set(x for n in t for x in node_elems[n]) - {e} for e, t in enumerate(T)]

    def dist(a, b):
        return np.sqrt(max(O.tri_distance(np.concatenate([X[T[a]], X[T[b]]]).reshape(-1))[0], 0.0))

    beyond = 0
    for e in range(len(T)):
        close = set(ids[start[e]:start[e + 1]].tolist())
        assert e not in close and nbr[e] <= close
        for c in close - nbr[e]:
            beyond += 1
            assert dist(e, c) < cutoff
            assert nbr[c] & (close | {e})
        frontier = set().union(*(nbr[c] for c in close)) - close - {e}
        for c in frontier:
            assert dist(e, c) >= cutoff
    assert beyond > 0                      # the closure reaches past the one-ring


def _damping(X, T, a, b, g, handle=None, **base):
    o = O.Oracle(X, T, lib_handle=handle, damping_alpha=a, damping_beta=b, damping_gamma=g, **base)
    Minv, C = o.mass()
    o.close()
    return Minv, C


def _beta_damping_brute(X, T, h, nu, E1, E2, beta):
    """the stiffness-proportional damping of compute_mass_damping
    (ShellSystem.C:828-833) by brute force: every element adds
    beta * 2 (K + D / (hmin hmax)) / 3 to each node of its one-ring patch
    (the vertices of all triangles that touch one of its three vertices),
    hmin / hmax its shortest and longest edge (libMesh Elem::hmin/hmax of a
    triangle: the extreme vertex distances), K = h / (1 - nu nu21),
    D = h^3 / (12 (1 - nu nu21)), nu21 = E2 nu / E1 (ShellSystem.C:473-476)"""
    nu21 = E2 * nu / E1
    K = h / (1 - nu * nu21)
    D = h ** 3 / (12 * (1 - nu * nu21))
    node_tris = [[] for _ in range(len(X))]
    for e, t in enumerate(T):
        for n in t:
            node_tris[n].append(e)
    C = np.zeros(len(X))
    for e, t in enumerate(T):
        patch = set()
        for n in t:
            for f in node_tris[n]:
                patch.update(T[f].tolist())
        edges = [np.linalg.norm(X[t[i]] - X[t[j]]) for i, j in ((0, 1), (0, 2), (1, 2))]
        stiff = 2 * (K + D / (min(edges) * max(edges))) / 3
        for n in patch:
            C[n] += beta * stiff
    return np.repeat(C, 3)


def test_damping_alpha_beta_gamma(icosphere):
    """compute_mass_damping (ShellSystem.C:762-852): the damping vector is
    alpha times the lumped mass (rho h != 1, so a damping formed from the
    unscaled element mass fails) plus the beta term formed by brute force
    (an independent patch-incidence sum) plus gamma per DOF, and linear in
    (alpha, beta, gamma); an imperfect mesh, so that the edge lengths differ
    from element to element"""
    X, T = icosphere(2)
    X = imperfect(X, 0.05, seed=3)
    base = dict(rho=2.5, thickness=0.02, nu=0.3, E1=1.0, E2=0.5)
    Minv, C = _damping(X, T, 0.3, 0.0, 0.0, **base)
    assert np.allclose(C * Minv, 0.3, rtol=1e-14, atol=0)
    _, Cg = _damping(X, T, 0.0, 0.0, 0.7, **base)
    assert np.array_equal(Cg, np.full_like(Cg, 0.7))
    _, Cb = _damping(X, T, 0.0, 0.2, 0.0, **base)
    brute = _beta_damping_brute(X, T, 0.02, 0.3, 1.0, 0.5, 0.2)
    assert np.allclose(Cb, brute, rtol=1e-13, atol=0)
    _, Call = _damping(X, T, 0.3, 0.2, 0.7, **base)
    assert np.allclose(Call, C + Cb + Cg, rtol=1e-14, atol=0)


@pytest.mark.parametrize("slip", ["hmin*hmin", "hmax only", "K without 2/3"])
def test_damping_beta_pin_catches_planted_slips(icosphere, slip):
    """the brute-force beta pin fails for an oracle copy with a plausible
    slip in the beta term (planted in a copy, on the CPU)"""
    from oracle_mutants import build
    old = "const double avg_stiff = 2 * (s.K + s.D / (avg_growth * hmin * hmax)) / 3;"
    new = {"hmin*hmin": "const double avg_stiff = 2 * (s.K + s.D / (avg_growth * hmin * hmin)) / 3;",
           "hmax only": "const double avg_stiff = 2 * (s.K + s.D / (avg_growth * hmax * hmax)) / 3;",
           "K without 2/3": "const double avg_stiff = (s.K + s.D / (avg_growth * hmin * hmax));"}[slip]
    L = build(old, new, slip.replace(" ", "_").replace("*", "x").replace("/", "_"))
    X, T = icosphere(2)
    X = imperfect(X, 0.05, seed=3)
    _, Cb = _damping(X, T, 0.0, 0.2, 0.0, handle=L, rho=2.5, thickness=0.02, nu=0.3, E1=1.0, E2=0.5)
    brute = _beta_damping_brute(X, T, 0.02, 0.3, 1.0, 0.5, 0.2)
    assert not np.allclose(Cb, brute, rtol=1e-6, atol=0)


def _torsion_energy_ratio(icosphere, level, handle=None, E1=1.0, E2=0.5, nu=0.3, G12=0.35, h=0.01, eps=1e-3):
    """E_tens of the torsion u = eps z (-y, x, 0) over its closed form
    1/2 K G12 (1 - nu nu') * int gamma^2 dA with gamma = -eps sin^2(theta)
    (int over the unit sphere: 32 pi / 15 eps^2), nu' = E1 nu / E2 (the
    reference's H22, ShellSystem.C:1199, 1208-1209), K = h / (1 - nu nu21),
    nu21 = E2 nu / E1 (ShellSystem.C:473-476)"""
    X, T = icosphere(level)
    o = O.Oracle(X, T, lib_handle=handle, E1=E1, E2=E2, nu=nu, G12=G12, thickness=h)
    u = np.zeros_like(X)
    u[:, 0] = -eps * X[:, 2] * X[:, 1]
    u[:, 1] = eps * X[:, 2] * X[:, 0]
    en = o.forces(u.reshape(-1))[2]
    o.close()
    K = h / (1 - nu * (E2 * nu / E1))
    closed = 0.5 * K * G12 * (1 - nu * (E1 * nu / E2)) * 32 * np.pi / 15 * eps ** 2
    return en[0] / closed


@pytest.mark.parametrize("E1,E2,nu", [(1.0, 0.5, 0.3), (0.8, 1.2, 0.25)])
def test_torsion_pins_the_orthotropic_shear_stiffness(icosphere, E1, E2, nu):
    """a torsion about the polar axis, u = eps z (-y, x, 0), is pure in-plane
    shear in the lat/long frame of reinit_mod (A1 east, A2 north), with
    engineering shear gamma = -eps sin^2(theta) on the unit sphere and no
    stretch to first order.  Its membrane energy is therefore the shear term
    of the reference's orthotropic H alone (H22, ShellSystem.C:1208-1209):
      E_tens = 1/2 K G12 (1 - nu nu') int gamma^2 dA,  nu' = E1 nu / E2.
    The discrete energy converges at second order in the mesh size (3.3 %
    low at level 3, 0.84 % at level 4); its Richardson extrapolation from
    levels 3 and 4 meets the closed form to 0.1 %.  With E2/E1 = 0.5,
    nu nu' = 0.18: a dropped (1 - nu nu') is 22 % off, nu' taken as
    nu21 = E2 nu / E1 is 16 % off."""
    r3 = _torsion_energy_ratio(icosphere, 3, E1=E1, E2=E2, nu=nu)
    r4 = _torsion_energy_ratio(icosphere, 4, E1=E1, E2=E2, nu=nu)
    assert 0.95 < r3 < r4 < 1.0
    assert abs(r4 + (r4 - r3) / 3 - 1) < 1e-3


@pytest.mark.parametrize("slip", ["drop (1 - nu nu')", "nu' -> nu21", "G12 -> E2/2(1+nu)"])
def test_torsion_pin_catches_planted_slips(icosphere, slip):
    """the torsion pin fails for oracle copies with a slip in the shear
    stiffness (planted in a copy, on the CPU)"""
    from oracle_mutants import build
    h8 = "H[8] = G12 * ((1 - nu * nu2) * g[1] * g[0] + (1 - nu * nu2) * ((1 + nu) / (1 - nu)) * g[2] * g[2]);"
    if slip == "drop (1 - nu nu')":
        old, new = h8, "H[8] = G12 * (g[1] * g[0] + ((1 + nu) / (1 - nu)) * g[2] * g[2]);"
    elif slip == "nu' -> nu21":
        old, new = "const double nu2 = (E1 * nu) / E2; /* ShellSystem.C:1199 */", "const double nu2 = (E2 * nu) / E1;"
    else:
        old, new = h8, h8.replace("H[8] = G12 *", "H[8] = (E2 / (2 * (1 + nu))) *")
    L = build(old, new, "torsion_%d" % ["drop (1 - nu nu')", "nu' -> nu21", "G12 -> E2/2(1+nu)"].index(slip))
    r3 = _torsion_energy_ratio(icosphere, 3, handle=L)
    r4 = _torsion_energy_ratio(icosphere, 4, handle=L)
    assert abs(r4 + (r4 - r3) / 3 - 1) > 0.05


@pytest.mark.parametrize("E1,E2", [(1.0, 0.5), (0.7, 1.3), (1.0, 1.0)])
def test_dilation_energies_closed_form(icosphere, E1, E2):
    """a uniform dilation u = eps X maps every tangent and second derivative
    to (1 + eps) times itself (the basis reproduces X), so the membrane
    strain is the equibiaxial Green strain E = eps + eps^2 / 2 times the
    reference metric and a3 is unchanged.  In the element's orthonormal
    frame the reference's H (ShellSystem.C:1199-1213) is [[E1, nu E1, .],
    [nu E1, E2, .], [., ., G12 ...]], so with K = h / (1 - nu nu21),
    nu21 = E2 nu / E1 (update_derived, ShellSystem.C:421-487):
      E_tens = 1/2 K E^2 (E1 + E2 + 2 nu E1) * area   (exactly),
      E_bend = 1/2 D eps^2 (E1 + E2 + 2 nu E1) / R^2 * area, D = h^3 / 12 ...
    with R = 1 up to the limit surface's deviation from the unit sphere."""
    X, T = icosphere(3)
    h, nu = 0.01, 0.3
    o = O.Oracle(X, T, E1=E1, E2=E2, nu=nu, G12=0.35, thickness=h)
    area = o.forces(np.zeros(X.size))[3].sum()
    c = (E1 + E2 + 2 * nu * E1) / (1 - nu * (E2 * nu / E1))
    for eps in (1e-3, 2e-2):
        en = o.forces((eps * X).reshape(-1))[2]
        G = eps + eps * eps / 2
        assert abs(en[0] - 0.5 * h * G ** 2 * c * area) <= 1e-12 * en[0]
        assert abs(en[1] - 0.5 * h ** 3 / 12 * eps ** 2 * c * area) <= 0.03 * en[1]
    o.close()


def test_volume_constraint_is_a_uniform_pressure_at_init(icosphere):
    """the volume term of compute_external_force (ShellSystem.C:1467-1585)
    is the pressure kappa (V0 r / V - 1) on the same follower load path as
    the applied pressure: right after initialisation V = V0, so a shell with
    bulk modulus kappa and target ratio r loads exactly like one under the
    pressure kappa (r - 1)"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=5)
    u = np.random.default_rng(2).normal(size=X.size) * 1e-3
    z = np.zeros(X.size)
    kappa, r = 0.05, 0.9
    out = []
    for kw in (dict(bulk_modulus=kappa, volume_final_ratio=r), dict(pressure_final=kappa * (r - 1))):
        o = O.Oracle(X, T, **kw)
        o.set_state(u, z, z, 0.0)
        o.init_dynamics()
        out.append(o.forces(u, t=0.0)[1])
        o.close()
    assert np.linalg.norm(out[0]) > 0
    assert np.linalg.norm(out[0] - out[1]) <= 1e-13 * np.linalg.norm(out[1])


def _ellipsoid_integrals(a, b, c, n=1600):
    """integrated mean curvature, Gaussian curvature and area of the
    ellipsoid with semi-axes a, b, c by dense midpoint quadrature of its
    (theta, phi) parametrisation (outward normal: mean curvature < 0)"""
    th = (np.arange(n) + 0.5) * np.pi / n
    ph = (np.arange(2 * n) + 0.5) * np.pi / n
    T, P = np.meshgrid(th, ph, indexing="ij")
    st, ct, sp, cp = np.sin(T), np.cos(T), np.sin(P), np.cos(P)
    xt = np.stack([a * ct * cp, b * ct * sp, -c * st])
    xp = np.stack([-a * st * sp, b * st * cp, 0 * T])
    xtt = np.stack([-a * st * cp, -b * st * sp, -c * ct])
    xpp = np.stack([-a * st * cp, -b * st * sp, 0 * T])
    xtp = np.stack([-a * ct * sp, b * ct * cp, 0 * T])
    E, F, G = (xt * xt).sum(0), (xt * xp).sum(0), (xp * xp).sum(0)
    nv = np.cross(xt, xp, axis=0)
    J = np.sqrt((nv ** 2).sum(0))
    nv = nv / J
    e, f, g = (xtt * nv).sum(0), (xtp * nv).sum(0), (xpp * nv).sum(0)
    H = (e * G - 2 * f * F + g * E) / (2 * (E * G - F * F))
    K = (e * g - f * f) / (E * G - F * F)
    dA = J * (np.pi / n) ** 2
    return (H * dA).sum(), (K * dA).sum(), dA.sum()


def _measured(icosphere, level, axes, handle=None):
    """area, integrated Gaussian and mean curvature measured by the oracle
    (measure_properties) on the icosphere mapped by diag(axes)"""
    X, T = icosphere(level)
    o = O.Oracle(X, T, lib_handle=handle)
    o.set_state((X * np.array(axes) - X).reshape(-1))
    o.init_dynamics()
    o.measure()
    p = o.properties()
    o.close()
    return p[2], p[14], p[15]


@pytest.mark.parametrize("axes", [(1.0, 1.0, 1.0), (1.3, 0.8, 1.1)])
def test_curvature_measurement_closed_forms(icosphere, axes):
    """measure_properties (ShellSystem.C:2728-2788) on the limit surface of
    an icosphere mapped by diag(a, b, c) (Loop surfaces are affine
    invariant: the limit of the mapped net is the mapped limit, which tends
    to the ellipsoid): the integrated Gaussian curvature is 4 pi
    (Gauss-Bonnet), the integrated mean curvature and the area are the
    ellipsoid's (dense parametric quadrature; -4 pi and 4 pi on the unit
    sphere).  The discrete values converge at second order; the Richardson
    extrapolation of levels 3 and 4 meets the closed forms to 2e-4.  The
    ellipsoid's off-diagonal metric (F != 0 in the element frames) brings in
    every term of both formulas."""
    H, K, A = _ellipsoid_integrals(*axes)
    r3 = _measured(icosphere, 3, axes)
    r4 = _measured(icosphere, 4, axes)
    ext = [r4[i] + (r4[i] - r3[i]) / 3 for i in range(3)]
    assert abs(K - 4 * np.pi) < 1e-5
    assert abs(ext[0] - A) < 2e-4 * A
    assert abs(ext[1] - K) < 2e-4 * K
    assert abs(ext[2] - H) < 2e-4 * abs(H)


def _shape_operator_curvatures(m):
    """per element, K and H from the eigenvalues of the shape operator
    I^-1 II (numpy), times the element weight jac / 2"""
    out = []
    for r in m:
        I = np.array([[r[0], r[2]], [r[2], r[1]]])
        II = np.array([[r[3], r[5]], [r[5], r[4]]])
        lam = np.linalg.eigvals(np.linalg.solve(I, II)).real
        w = r[6] * 0.5
        out.append((w * lam[0] * lam[1], w * 0.5 * (lam[0] + lam[1])))
    return np.array(out)


def _sheared_elem_measure(handle=None):
    X, T = O.icosphere(3)
    M = np.array([[1.3, 0.5, 0.0], [-0.3, 0.8, 0.4], [0.2, -0.4, 1.1]])
    o = O.Oracle(X, T, lib_handle=handle)
    o.forces((X @ M.T - X).reshape(-1))
    m = o.elem_measure()
    o.close()
    return m


def test_curvature_per_element_is_the_shape_operator():
    """each element's Gaussian and mean curvature (measure_properties,
    ShellSystem.C:2746-2751) equal the determinant and half trace of the
    shape operator I^-1 II formed from the same fundamental forms by an
    eigen-decomposition, on a sheared shell whose element frames are far
    from orthogonal (ff1_2 != 0)"""
    m = _sheared_elem_measure()
    assert np.max(np.abs(m[:, 2]) / np.sqrt(m[:, 0] * m[:, 1])) > 0.2
    ref = _shape_operator_curvatures(m)
    assert np.allclose(m[:, 7], ref[:, 0], rtol=1e-12, atol=0)
    assert np.allclose(m[:, 8], ref[:, 1], rtol=1e-12, atol=0)
    assert np.allclose(m[:, 9], 0.5 * m[:, 6], rtol=1e-15, atol=0)


@pytest.mark.parametrize("slip", ["mean without 1/2", "mean without cross term", "gauss without ff2_2^2"])
def test_curvature_pins_catch_planted_slips(slip):
    from oracle_mutants import build
    kg = "kG = w * invdet * (ff2[0] * ff2[1] - ff2[2] * ff2[2]);"
    km = "kM = w * invdet * 0.5 * (ff2[0] * ff1[1] + ff2[1] * ff1[0] - 2 * ff2[2] * ff1[2]);"
    old, new = {
        "mean without 1/2": (km, km.replace("0.5 * ", "")),
        "mean without cross term": (km, km.replace(" - 2 * ff2[2] * ff1[2]", "")),
        "gauss without ff2_2^2": (kg, kg.replace(" - ff2[2] * ff2[2]", "")),
    }[slip]
    L = build(old, new, "curv_%d" % ["mean without 1/2", "mean without cross term",
                                     "gauss without ff2_2^2"].index(slip))
    m = _sheared_elem_measure(L)
    ref = _shape_operator_curvatures(m)
    assert not (np.allclose(m[:, 7], ref[:, 0], rtol=1e-6) and np.allclose(m[:, 8], ref[:, 1], rtol=1e-6))


@pytest.mark.parametrize("adapt", [0, 1])
def test_moving_averages_are_dt_weighted_means(icosphere, adapt):
    """print_properties (ShellSystem.C:2849-2898) with MovingAverage
    (MovingAverage.h:25-35): each column is the dt-weighted mean over the
    accepted steps since the last read (numpy's weighted average of the
    per-step values here), the dt column the mean step, and a read resets"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=8)
    o = O.Oracle(X, T, pressure_final=-2 * PCR, pressure_ramp_time=3.0, dt=0.1, damping_alpha=0.05,
                 adapt_dt_every=adapt, newmark_beta=0.25 if adapt else 0.0)
    o.set_state(np.zeros(X.size))
    o.init_dynamics()
    o.properties()                           # start clean
    dts, et, eb, ek, t0 = [], [], [], [], 0.0
    for _ in range(25):
        r = o.step(1)
        dts.append(r["t"] - t0)
        t0 = r["t"]
        et.append(r["e_tens"]); eb.append(r["e_bend"]); ek.append(r["e_kin"])
    p = o.properties()
    assert p[1] == t0
    assert abs(p[0] - np.mean(dts)) <= 1e-14 * np.mean(dts)
    for col, vals in ((9, et), (10, eb), (8, ek)):
        assert abs(p[col] - np.average(vals, weights=dts)) <= 1e-12 * abs(np.average(vals, weights=dts))
    if adapt:
        assert np.std(dts) > 0                # the weights differ
    again = o.properties()                    # reset: no samples left
    assert again[0] == 0
    o.close()


# ---------------------------------------------------------------- cavity, gravity, rotation damping

def _stiff_density(E1, E2, nu, h):
    """the Hertz force density of update_derived: half the mean Young's
    modulus over (1 - nu nu21) h (ShellSystem.C:477-483)"""
    return 0.5 * 0.5 * (E1 + E2) / ((1 - nu * (E2 * nu / E1)) * h)


def _mean_vertex_area(T, area, nv):
    """per node, the mean area of the elements that have it as a vertex"""
    s = np.zeros(nv)
    c = np.zeros(nv)
    for k in range(3):
        np.add.at(s, T[:, k], area)
        np.add.at(c, T[:, k], 1)
    return s / c


def _cavity_sphere(handle=None, **kw):
    X, T = O.icosphere(3, round_g=False)
    base = dict(E1=1.0, E2=0.5, nu=0.3, G12=0.35, thickness=H, cavity_contact=1, cavity=2,
                cavity_R0=(0.99, 0.99, 0.99))
    base.update(kw)
    return X, T, O.Oracle(X, T, lib_handle=handle, **base)


def test_cavity_sphere_in_sphere_closed_form():
    """a unit sphere dilated by (1 + eps) inside a spherical cavity of radius
    R: every node penetrates by depth = 1 + eps - (R - h/2) (the effective
    semi-axis of update_derived, ShellSystem.C:424-435, and the isoparameter
    of ShellSystem.h:336-365 reduce to this for a sphere), the element areas
    scale by (1 + eps)^2, and the node is pushed back along its radius by
    stiff_density * (mean area of its elements) * depth; its energy over
    that area is stiff_density depth^2 / 2 (ShellSystem.C:1615-1717)"""
    X, T, o = _cavity_sphere()
    sd = _stiff_density(1.0, 0.5, 0.3, H)
    R_eff = 0.99 - H / 2
    area0 = o.forces(np.zeros(X.size))[3]
    abar = _mean_vertex_area(T, area0, len(X))
    for eps in (0.0, -0.004, 0.003):
        fext = o.forces((eps * X).reshape(-1))[1].reshape(-1, 3)
        depth = 1 + eps - R_eff
        expect = -(sd * (1 + eps) ** 2 * abar * depth)[:, None] * X
        assert np.abs(fext - expect).max() <= 1e-12 * np.abs(expect).max()
        _, ecav, _ = o.node_cavity()
        np.testing.assert_allclose(ecav, 0.5 * sd * depth ** 2, rtol=1e-12)
    o.close()


def test_cavity_box_and_cylinder_normals():
    """special cases of the box and cylinder normals (ShellSystem.C:1635-1666):
    a shell pushed against a face of a box is pushed back along the face
    normal only, by stiff_density * mean area * depth, the depth being the
    overshoot x - R_eff scaled by mean(R_eff) / R_eff (the isoparameter's
    depth, ShellSystem.C:1621); pushed against the wall of a cylinder (axis
    x) with equal radial semi-axes, along the radius in the (y, z) plane;
    pushed against the cylinder's end cap, along x"""
    sd = _stiff_density(1.0, 0.5, 0.3, H)
    for cav, R0, shift, axis in ((0, (1.0, 1.2, 1.2), 0.012, 0), (1, (1.2, 1.006, 1.006), 0.012, 1),
                                 (1, (1.0, 1.2, 1.2), 0.012, 0)):
        X, T, o = _cavity_sphere(cavity=cav, cavity_R0=R0)
        u = np.zeros_like(X)
        u[:, axis] = shift
        fext, area = (lambda r: (r[1].reshape(-1, 3), r[3]))(o.forces(u.reshape(-1)))
        abar = _mean_vertex_area(T, area, len(X))
        p = X + u
        R_eff = np.array(R0) - H / 2
        if cav == 1 and axis == 1:
            rho = np.hypot(p[:, 1], p[:, 2])
            hit = rho > R_eff[1]
            n = np.zeros_like(p)
            n[:, 1:] = p[:, 1:] / rho[:, None]
            depth = (rho - R_eff[1]) * (R_eff.mean() / R_eff[1])
        else:
            hit = p[:, axis] > R_eff[axis]
            n = np.zeros_like(p)
            n[:, axis] = 1
            depth = (p[:, axis] - R_eff[axis]) * (R_eff.mean() / R_eff[axis])
        assert 5 < hit.sum() < len(X) // 4
        expect = np.where(hit[:, None], -(sd * abar * depth)[:, None] * n, 0.0)
        assert np.abs(fext - expect).max() <= 1e-12 * np.abs(expect).max()
        o.close()


def test_cavity_damping_and_friction_special_cases():
    """with a uniform radial velocity c the normal force grows by
    cavity_damping * k * c (ShellSystem.C:1709-1712); a tangential velocity
    above friction_stop_velocity adds the dynamic friction -mu_d f_n v_t /
    (|v_t| + 1e-8 v_stop) formed from the undamped normal force f_n = k depth
    (Friction.C:4-60 called before the damping term, ShellSystem.C:1688-1712)"""
    X, T, o = _cavity_sphere(cavity_damping=0.7, mu_d_cavity=0.4, friction_stop_velocity=1e-3)
    sd = _stiff_density(1.0, 0.5, 0.3, H)
    R_eff = 0.99 - H / 2
    z = np.zeros(X.size)
    area = o.forces(z)[3]
    k = sd * _mean_vertex_area(T, area, len(X))
    depth = 1 - R_eff
    c, w = 0.02, 0.05
    vt = w * np.cross([0.0, 0.0, 1.0], X)              # rotation about z
    v = c * X + vt
    fext = o.forces(z, v.reshape(-1))[1].reshape(-1, 3)
    avt = np.linalg.norm(vt, axis=1)
    fric = -(0.4 * k * depth / (avt + 1e-8 * 1e-3))[:, None] * vt
    expect = -((k * depth + 0.7 * k * c))[:, None] * X + fric
    fast = avt > 1e-3
    assert fast.sum() > len(X) // 2
    assert np.abs(fext[fast] - expect[fast]).max() <= 1e-12 * np.abs(expect).max()
    o.close()


def test_gravity_and_rotation_damping_closed_forms(icosphere):
    """gravity loads each node with -g m_n along its direction
    (ShellSystem.C:1587-1597); for a rigid rotation v = w0 x p the angular
    velocity I^-1 L of ShellSystem.C:1727-1770 is w0 itself, so the rotation
    damping force is -c (w0 x p) at every node, also on an imperfect shell
    whose inertia tensor is full"""
    X, T = icosphere(3)
    X = imperfect(X, 0.05, seed=11) * np.array([1.0, 1.3, 0.8])
    u = np.random.default_rng(4).normal(size=X.size) * 1e-3
    p = X + u.reshape(-1, 3)
    g, cr = 0.37, 0.8
    w0 = np.array([0.3, -0.7, 0.45])
    o = O.Oracle(X, T, gravity=g, gravity_dir=1, damping_rotation=cr)
    Minv, _ = o.mass()
    m = 1 / Minv[0::3]
    fext = o.forces(u, np.cross(w0, p).reshape(-1))[1].reshape(-1, 3)
    expect = -cr * np.cross(w0, p)
    expect[:, 1] -= g * m
    assert np.abs(fext - expect).max() <= 1e-12 * np.abs(expect).max()
    o.close()


def test_cavity_time_functions_and_measures():
    """R(t) = R(0) exp(rate t) at the force time; the properties of one
    sample: packing density area h / (4/3 pi R^3), cavity pressure
    (sum of the normal forces) / (4 pi R^2) (Knud Thomsen's surface is exact
    for a sphere), the count of nodes in contact and the cavity energy
    (ShellSystem.C:440-470, 2818-2843)"""
    X, T, o = _cavity_sphere(cavity_rate=-0.5)
    X2, T2, o2 = _cavity_sphere(cavity_R0=tuple([0.99 * np.exp(-0.5 * 0.03)] * 3))
    z = np.zeros(X.size)
    f1 = o.forces(z, None, 0.03)[1]
    f2 = o2.forces(z, None, 0.0)[1]
    assert np.abs(f1 - f2).max() <= 1e-15 * np.abs(f2).max()
    o2.close()
    o.close()
    X, T, o = _cavity_sphere()
    o.set_state(z, z, z, 0.0)
    o.init_dynamics()
    o.measure()
    _, fext, _, area = o.forces(z, None, 0.0)
    fn = np.linalg.norm(fext.reshape(-1, 3), axis=1)
    _, ecav, _ = o.node_cavity()
    abar = _mean_vertex_area(T, area, len(X))
    R = 0.99
    pr = o.properties()
    assert abs(pr[4] - area.sum() * H / (4 / 3 * np.pi * R ** 3)) <= 1e-13 * pr[4]
    assert abs(pr[5] - fn.sum() / (4 * np.pi * R ** 2)) <= 1e-12 * pr[5]
    assert pr[6] == len(X)
    assert abs(pr[11] - (ecav * abar).sum()) <= 1e-12 * pr[11]
    o.close()


@pytest.mark.parametrize("slip", ["R instead of R_eff", "area sum", "damping sign"])
def test_cavity_pins_catch_planted_slips(slip):
    from oracle_mutants import build
    old, new = {
        "R instead of R_eff": ("const double h_half = 0.5 * P.thickness;", "const double h_half = 0.0;"),
        "area sum": ("  return a / s.node_elems[n].size();", "  return a;"),
        "damping sign": ("f_normal += P.cavity_damping * stiffness * abs_v_normal;",
                         "f_normal -= P.cavity_damping * stiffness * abs_v_normal;"),
    }[slip]
    L = build(old, new, "cav" + str(abs(hash(slip)) % 1000))
    if slip == "damping sign":
        X, T, o = _cavity_sphere(handle=L, cavity_damping=0.7)
        sd = _stiff_density(1.0, 0.5, 0.3, H)
        z = np.zeros(X.size)
        area = o.forces(z)[3]
        k = sd * _mean_vertex_area(T, area, len(X))
        fext = o.forces(z, (0.02 * X).reshape(-1))[1].reshape(-1, 3)
        expect = -(k * (1 - 0.99 + H / 2) + 0.7 * k * 0.02)[:, None] * X
        assert np.abs(fext - expect).max() > 1e-6 * np.abs(expect).max()
        o.close()
        return
    X, T, o = _cavity_sphere(handle=L)
    sd = _stiff_density(1.0, 0.5, 0.3, H)
    area0 = o.forces(np.zeros(X.size))[3]
    abar = _mean_vertex_area(T, area0, len(X))
    fext = o.forces(np.zeros(X.size))[1].reshape(-1, 3)
    expect = -(sd * abar * (1 - 0.99 + H / 2))[:, None] * X
    assert np.abs(fext - expect).max() > 1e-6 * np.abs(expect).max()
    o.close()


@pytest.mark.parametrize("slip", ["inertia off-diagonal sign", "gravity on the mass inverse"])
def test_gravity_rotation_pins_catch_planted_slips(icosphere, slip):
    from oracle_mutants import build
    old, new = {
        "inertia off-diagonal sign": ("      I[1] -= mi * ri(0) * ri(1);", "      I[1] += mi * ri(0) * ri(1);"),
        "gravity on the mass inverse": ("      s.fext[dof] -= P.gravity / s.Minv[dof];",
                                        "      s.fext[dof] -= P.gravity * s.Minv[dof];"),
    }[slip]
    L = build(old, new, "rot" + str(abs(hash(slip)) % 1000))
    X, T = icosphere(3)
    X = imperfect(X, 0.05, seed=11) * np.array([1.0, 1.3, 0.8])
    u = np.random.default_rng(4).normal(size=X.size) * 1e-3
    p = X + u.reshape(-1, 3)
    w0 = np.array([0.3, -0.7, 0.45])
    o = O.Oracle(X, T, lib_handle=L, gravity=0.37, gravity_dir=1, damping_rotation=0.8)
    Minv, _ = o.mass()
    fext = o.forces(u, np.cross(w0, p).reshape(-1))[1].reshape(-1, 3)
    expect = -0.8 * np.cross(w0, p)
    expect[:, 1] -= 0.37 / Minv[0::3]
    assert np.abs(fext - expect).max() > 1e-6 * np.abs(expect).max()
    o.close()


# ---------------------------------------------------------------- growth

def _grown(X, T, handle=None, **kw):
    base = dict(E1=1.0, E2=0.5, nu=0.3, G12=0.35, thickness=H)
    base.update(kw)
    return O.Oracle(X, T, lib_handle=handle, **base)


@pytest.mark.parametrize("g1,g2", [(1.1, 1.1), (1.08, 0.95), (0.9, 1.2)])
def test_local_growth_energy_closed_form(g1, g2):
    """growth factors g1 along the east tangent a1 and g2 along the north
    tangent a2 (growth_base local, thickness_dir 2, ShellSystem.C:2339-2415)
    make the unloaded sphere's membrane strain the Green strain of the
    inverse growth, eps_k = (1/g_k^2 - 1)/2, in the element's orthogonal
    frame, so that E_tens = 1/2 K (E1 eps1^2 + 2 nu E1 eps1 eps2 + E2 eps2^2)
    * area exactly, K = h/(1 - nu nu21); and the mass follows det G = g1 g2
    (grow_mass, ShellSystem.C:795-811).  For g1 = g2 = g the dilated sphere
    u = (g - 1) X is free of membrane strain, and its bending energy is that
    of the ungrown sphere dilated by 1/g - 1."""
    X, T = O.icosphere(3, round_g=False)
    o = _grown(X, T, growth_g0=(g1, g2))
    z = np.zeros(X.size)
    _, _, en, area = o.forces(z)
    e1, e2 = 0.5 * (1 / g1 ** 2 - 1), 0.5 * (1 / g2 ** 2 - 1)
    K = H / (1 - 0.3 * (0.5 * 0.3 / 1.0))
    closed = 0.5 * K * (1.0 * e1 ** 2 + 2 * 0.3 * 1.0 * e1 * e2 + 0.5 * e2 ** 2) * area.sum()
    assert abs(en[0] - closed) <= 1e-11 * closed
    Minv, _ = o.mass()
    assert abs((1 / Minv[0::3]).sum() - H * g1 * g2 * area.sum()) <= 1e-12 * H * area.sum()
    if g1 == g2:
        en_g = o.forces(((g1 - 1) * X).reshape(-1))[2]
        p = _grown(X, T)
        en_p = p.forces(((1 / g1 - 1) * X).reshape(-1))[2]
        assert abs(en_g[0]) <= 1e-12 * en_p[0]
        assert abs(en_g[1] - en_p[1]) <= 1e-10 * en_p[1]
        p.close()
    o.close()


def _rot(axis, ang):
    c, s = np.cos(ang), np.sin(ang)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    R = np.eye(3)
    R[i, i] = R[j, j] = c
    R[i, j], R[j, i] = -s, s
    return R


def test_cartesian_and_cylindrical_growth_symmetries():
    """cartesian growth diag(g1, g2, 1) (thickness_dir 2) on a mesh turned a
    quarter about z equals diag(g2, g1, 1) on the original; cylindrical
    growth about x (axial g1, hoop g2; thickness_dir 1) is unchanged by a
    turn of the mesh about x, for an isotropic material (whose energy does
    not see the lat/long frame).  Energies at the same strained state (the
    displacement turned with the mesh) agree to rounding."""
    X, T = O.icosphere(3)
    X = imperfect(X, 0.02, seed=4)
    u = np.random.default_rng(5).normal(size=X.shape) * 1e-3
    iso = dict(E1=1.0, E2=1.0, nu=0.3, G12=1.0 / 2.6)
    R = _rot(2, np.pi / 2)
    a = _grown(X, T, growth_base=1, growth_g0=(1.1, 0.93), **iso)
    b = _grown(X @ R.T, T, growth_base=1, growth_g0=(0.93, 1.1), **iso)
    ea, eb = a.forces(u.reshape(-1))[2], b.forces((u @ R.T).reshape(-1))[2]
    assert np.allclose(ea[:2], eb[:2], rtol=1e-10, atol=0)
    R = _rot(0, 0.61)
    a = _grown(X, T, growth_base=2, thickness_dir=1, growth_g0=(1.07, 0.9), **iso)
    b = _grown(X @ R.T, T, growth_base=2, thickness_dir=1, growth_g0=(1.07, 0.9), **iso)
    ea, eb = a.forces(u.reshape(-1))[2], b.forces((u @ R.T).reshape(-1))[2]
    assert np.allclose(ea[:2], eb[:2], rtol=1e-10, atol=0)
    # and the anisotropic growth is seen: swapping the factors changes the energy
    c = _grown(X, T, growth_base=2, thickness_dir=1, growth_g0=(0.9, 1.07), **iso)
    assert abs(c.forces(u.reshape(-1))[2][0] - ea[0]) > 1e-3 * ea[0]


def test_dynamic_growth_steps_with_the_previous_strains(icosphere):
    """dynamic growth (apply_growth, ShellSystem.C:2317-2336, in
    prepare_next_step before the step's force evaluation): each step adds
    dt (metric_growth_rate alpha + curvature_growth_rate beta) to G, with the
    strains of the previous force evaluation"""
    X, T = icosphere(3)
    X = imperfect(X, 1e-3, seed=6)
    o = _grown(X, T, metric_growth_rate=0.8, curvature_growth_rate=30.0, pressure_final=-2e-4, dt=0.05)
    o.set_state(np.zeros(X.size))
    o.init_dynamics()
    G0, _, al, be = o.elem_growth()
    assert np.array_equal(G0, np.tile([1.0, 1.0, 0.0], (len(T), 1)))
    for _ in range(3):
        o.step(1)
        G1, GV, al1, be1 = o.elem_growth()
        expect = (G0 + al * (0.05 * 0.8)) + be * (0.05 * 30.0)
        assert np.array_equal(G1, expect)
        # GV is the Voigt form of the inverse of [[G0, G2/2], [G2/2, G1]]
        det = G1[:, 0] * G1[:, 1] - 0.25 * G1[:, 2] ** 2
        gi00, gi11, gi01 = G1[:, 1] / det, G1[:, 0] / det, -0.5 * G1[:, 2] / det
        assert np.allclose(GV[:, 0, 0], gi00 ** 2, rtol=1e-14) and np.allclose(GV[:, 2, 2], gi01 ** 2 + gi00 * gi11,
                                                                               rtol=1e-13)
        G0, al, be = G1, al1, be1
    assert np.abs(G0 - np.array([1.0, 1.0, 0.0])).max() > 1e-6
    o.close()


def _growth_energy_geometric(o, X, T, base, td, g1, g2, E1=1.0, E2=0.5, nu=0.3, G12=0.35):
    """membrane energy of the unloaded shell under static growth, formed in
    each element's orthonormal (east, north) frame without the reference's
    basis transforms: the inverse growth F^-1 = B diag(1/g) B^T of
    growth_base B (local: the element frame; cartesian; cylindrical about
    x), projected on the tangent plane (the in-plane block G of the inverse
    growth tensor), the Green strain eps = (G^T G - I)/2, and the orthotropic
    energy density 1/2 K (E1 e11^2 + 2 nu E1 e11 e22 + E2 e22^2 +
    G12 (1 - nu nu') gamma^2), gamma = 2 e12, nu' = E1 nu / E2 (the shear
    term as the torsion pin fixes it)"""
    tab = o.elem_tables()
    A1, A2, xyz, JxW = tab[:, 72:75], tab[:, 75:78], tab[:, 98:101], tab[:, 93]
    e = A1 / np.linalg.norm(A1, axis=1)[:, None]
    n = A2 / np.linalg.norm(A2, axis=1)[:, None]
    N = np.cross(e, n)
    gd = [1 if td < 1 else 0, 2 if td < 2 else 1, td]
    K = H / (1 - nu * (E2 * nu / E1))
    total = 0.0
    for k in range(len(T)):
        if base == 0:
            B = np.column_stack([e[k], n[k], N[k]])
        elif base == 1:
            B = np.eye(3)
        else:
            r = np.hypot(xyz[k, 1], xyz[k, 2])
            c, s_ = xyz[k, 1] / r, xyz[k, 2] / r
            B = np.array([[1, 0, 0], [0, c, -s_], [0, s_, c]])
        d = np.ones(3)
        d[gd[0]], d[gd[1]] = 1 / g1, 1 / g2
        Fi = B @ np.diag(d) @ B.T
        G = np.array([[e[k] @ Fi @ e[k], e[k] @ Fi @ n[k]], [n[k] @ Fi @ e[k], n[k] @ Fi @ n[k]]])
        eps = 0.5 * (G.T @ G - np.eye(2))
        total += JxW[k] * 0.5 * K * (E1 * eps[0, 0] ** 2 + 2 * nu * E1 * eps[0, 0] * eps[1, 1] + E2 * eps[1, 1] ** 2 +
                                     G12 * (1 - nu * E1 * nu / E2) * (2 * eps[0, 1]) ** 2)
    return total


GROWTH_CASES = [(0, 2, 1.08, 0.95), (1, 2, 1.1, 0.93), (1, 0, 1.05, 1.0), (2, 1, 1.07, 0.9), (2, 0, 0.92, 1.06)]


@pytest.mark.parametrize("base,td,g1,g2", GROWTH_CASES)
def test_growth_energy_geometric_closed_form(base, td, g1, g2):
    """static growth in every growth_base and thickness_dir
    (ShellSystem.C:2339-2440, with growth_dir from thickness_dir, :246-250):
    the unloaded shell's membrane energy equals the energy of the projected
    inverse growth formed geometrically in each element's orthonormal frame
    (_growth_energy_geometric), to rounding"""
    X, T = O.icosphere(3)
    X = imperfect(X, 0.02, seed=4)
    o = _grown(X, T, growth_base=base, thickness_dir=td, growth_g0=(g1, g2))
    en = o.forces(np.zeros(X.size))[2]
    closed = _growth_energy_geometric(o, X, T, base, td, g1, g2)
    assert abs(en[0] - closed) <= 1e-12 * closed
    o.close()


@pytest.mark.parametrize("slip", ["growth_dir swapped", "GV transposed", "mass without det", "B^-1 left out"])
def test_growth_pins_catch_planted_slips(slip):
    from oracle_mutants import build
    old, new = {
        "growth_dir swapped": ("      Ginv[s.growth_dir[0] * 4] = 1 / g[0];\n      Ginv[s.growth_dir[1] * 4] = 1 / g[1];",
                               "      Ginv[s.growth_dir[1] * 4] = 1 / g[0];\n      Ginv[s.growth_dir[0] * 4] = 1 / g[1];"),
        "GV transposed": ("    GV[1] = gi(1, 0) * gi(1, 0);", "    GV[1] = gi(0, 1) * gi(0, 1);"),
        "mass without det": ("      JxW = c.JxW * det;", "      JxW = c.JxW;"),
        "B^-1 left out": ("      matmul3(Bi, A, BA);", "      matmul3(B, A, BA);"),
    }[slip]
    L = build(old, new, "grow" + str(abs(hash(slip)) % 1000))
    X, T = O.icosphere(3)
    X = imperfect(X, 0.02, seed=4)
    caught = False
    for base, td, g1, g2 in GROWTH_CASES:
        o = _grown(X, T, handle=L, growth_base=base, thickness_dir=td, growth_g0=(g1, g2))
        _, _, en, area = o.forces(np.zeros(X.size))
        closed = _growth_energy_geometric(o, X, T, base, td, g1, g2)
        Minv, _ = o.mass()
        caught |= abs(en[0] - closed) > 1e-9 * closed
        caught |= abs((1 / Minv[0::3]).sum() - H * g1 * g2 * area.sum()) > 1e-9 * H * area.sum()
        o.close()
    assert caught
