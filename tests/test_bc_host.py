"""Open meshes on the CPU (no GPU needed): the ghost layer the library adds
to an open mesh (libMesh's add_boundary_ghosts, which the reference's
prepare_subdiv_mesh calls, shell.C:228-230), the oracle's BoundaryCache list
(ShellSystem.C:603-651) and its boundary conditions
(apply_boundary_conditions, ShellSystem.C:2138-2292), pinned by closed
forms of plate and shell theory:

- a square plate under a uniform load, pinned on all four edges (Navier's
  series: w_max = 0.00406 q a^4 / D) and clamped (w_max = 0.00126 q a^4 / D,
  Timoshenko & Woinowsky-Krieger, Theory of Plates and Shells, tables 8 and
  35), approached from above and below as the mesh is refined (the ghost
  conditions are first order in the element size);
- the weak conditions tending to the strong ones as their springs stiffen;
- an open tube with free ends under internal pressure p: the membrane
  solution w = p r^2 / (E h), axial strain -nu p r / (E h).
"""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import plate, tube

E, NU, RHO = 1.0, 0.3, 1.0
ISO = dict(E1=E, E2=E, nu=NU, G12=E / (2 * (1 + NU)), rho=RHO)


def with_ghosts(X, T):
    Xg, Tg, ng = S.mesh_add_ghosts(X, T)
    return Xg, Tg, ng


def boundary_edges_of(T):
    """directed boundary edges (a, b) of the real triangles and their third nodes"""
    d = {}
    for t in T:
        for k in range(3):
            d[(t[k], t[(k + 1) % 3])] = t[(k + 2) % 3]
    return {e: c for e, c in d.items() if (e[1], e[0]) not in d}


@pytest.mark.parametrize("mesh", ["plate", "tube"])
def test_ghost_layer(mesh):
    X, T = plate(5) if mesh == "plate" else tube(12, 3)
    Xg, Tg, ng = with_ghosts(X, T)
    nv, ne = len(X), len(T)
    # the real nodes and triangles keep their numbers
    assert np.array_equal(Xg[:nv], X) and np.array_equal(Tg[:ne], T) and len(Tg) == ne + ng
    be = boundary_edges_of(T)
    nreal = np.bincount(T.ravel(), minlength=nv)
    bverts = sorted({a for a, _ in be})
    # one ghost triangle per boundary edge, then 4 - r per boundary vertex of r triangles
    assert ng == len(be) + sum(4 - nreal[v] for v in bverts)
    # every real node has valence 6 or its closed-mesh valence; boundary vertices 6
    val = np.bincount(Tg.ravel(), minlength=len(Xg))
    assert all(val[v] == 6 for v in bverts)
    # consistently oriented: each directed edge once, the real nodes' rings closed
    d = {}
    for t in Tg:
        for k in range(3):
            e = (t[k], t[(k + 1) % 3])
            assert e not in d
            d[e] = 1
    for v in range(nv):
        out = {b for (a, b) in d if a == v}
        inn = {a for (a, b) in d if b == v}
        assert out == inn
    # the ghost of a boundary edge a -> b of (a, b, c) is c mirrored through the edge's midpoint
    ghost_of = {}
    for t in Tg[ne:]:
        for k in range(3):
            if (t[(k + 1) % 3], t[k]) in be:
                ghost_of[(t[(k + 1) % 3], t[k])] = t[(k + 2) % 3]
    assert set(ghost_of) == set(be)
    for (a, b), c in be.items():
        assert np.allclose(Xg[ghost_of[(a, b)]], X[a] + X[b] - X[c], rtol=0, atol=1e-15)


def test_ghost_layer_of_closed_and_unsupported_meshes():
    X, T = S.mesh_icosphere(2)
    Xg, Tg, ng = with_ghosts(X, T)
    assert ng == 0 and np.array_equal(Tg, T)
    # a boundary vertex with four triangles (a reflex corner) takes no valence-6 ring
    X, T = plate(2)
    keep = [t for t in T if not (set(t) <= {4, 5, 7, 8})]       # cut one square out of a corner
    with pytest.raises(S.ShellError):
        S.mesh_add_ghosts(X, np.array(keep, dtype=np.int32))


def test_oracle_boundary_edges():
    X, T = plate(4)
    Xg, Tg, ng = with_ghosts(X, T)
    o = oracle.Oracle(Xg, Tg, ghost_elems=ng, bc_every=1)
    be = o.boundary_edges()
    ref = boundary_edges_of(T)
    assert len(be) == len(ref) == 16
    for n1, n2, n3, n4 in be:
        # the edge n3 -> n2 is a real triangle's boundary edge, n1 its third node, n4 the ghost
        assert ref[(n3, n2)] == n1 and n4 >= len(X)
        assert np.allclose(Xg[n4], Xg[n2] + Xg[n3] - Xg[n1], rtol=0, atol=1e-15)
    # no list without conditions acting (bc_every 0) or without coordinates (bc_coords 0)
    assert len(oracle.Oracle(Xg, Tg, ghost_elems=ng, bc_every=0).boundary_edges()) == 0
    assert len(oracle.Oracle(Xg, Tg, ghost_elems=ng, bc_every=1, bc_coords=0).boundary_edges()) == 0
    o.close()


def settle(o, nv_real, comp, tol=1e-6, chunk=100, limit=400):
    """steps to rest; the largest |u_comp| of the real nodes"""
    o.init_dynamics()
    last = None
    for _ in range(limit):
        o.step(chunk)
        u = o.get_state()[0].reshape(-1, 3)[:nv_real]
        w = np.abs(u[:, comp]).max()
        assert np.isfinite(w)
        if last is not None and abs(w - last) <= tol * w:
            return w, u
        last = w
    raise AssertionError("no rest")


def plate_deflection(n, bc, mult=1.0, cfl=1.5):
    """max deflection of the unit plate (h 0.05) under its own weight, over
    the closed form's value"""
    X, T = plate(n)
    Xg, Tg, ng = with_ghosts(X, T)
    h, g = 0.05, 3e-4
    kw = dict(ISO, thickness=h, gravity=g, gravity_dir=1, damping_alpha=0.6, ghost_elems=ng, shell_boundary=bc,
              bc_every=1, bc_stiff_mult=mult, dt=1.0)
    o = oracle.Oracle(Xg, Tg, **kw)
    om = o.cfl_omega_max()
    if bc in (1, 3):
        # the springs on the light ghost nodes: sqrt(k / m)
        Minv = o.mass()[0]
        om = max(om, np.sqrt(mult * E * h / (1 - NU ** 2) * Minv.max()))
    o.close()
    kw["dt"] = cfl / om
    o = oracle.Oracle(Xg, Tg, **kw)
    w, u = settle(o, len(X), 1, tol=1e-6 if cfl > 1 else 1e-5, chunk=100 if cfl > 1 else 1000)
    be = o.boundary_edges()
    o.close()
    D = E * h ** 3 / (12 * (1 - NU ** 2))
    q = RHO * h * g
    return w / ((0.00406 if bc in (2, 3) else 0.00126) * q / D), u, be


def test_pinned_and_clamped_plates_approach_the_closed_forms():
    r12, u, be = plate_deflection(12, 2)
    r16, _, _ = plate_deflection(16, 2)
    # pinned: the edge nodes at rest, ghosts mirror n1 exactly; from above, the error shrinking
    assert np.abs(u[be[:, 1:3].ravel()]).max() == 0
    assert 1.0 < r16 < r12 < 1.2 and (r12 - 1) / (r16 - 1) > 1.2
    c12, u, be = plate_deflection(12, 0)
    c16, _, _ = plate_deflection(16, 0)
    assert np.abs(u[be[:, :3].ravel()]).max() == 0
    assert 0.85 < c12 < c16 < 1.0 and (1 - c12) / (1 - c16) > 1.2
    # the two differ by their closed forms' factor 3.2 (and a bit more on a coarse mesh)
    assert 3.0 < r16 * 0.00406 / (c16 * 0.00126) < 4.2


@pytest.mark.parametrize("weak,strong", [(3, 2), (1, 0)])
def test_weak_conditions_tend_to_the_strong_ones(weak, strong):
    rs, _, _ = plate_deflection(8, strong)
    r1, _, _ = plate_deflection(8, weak, mult=1.0, cfl=0.5)
    r10, _, _ = plate_deflection(8, weak, mult=10.0, cfl=0.5)
    # springs are softer than constraints: more deflection, less as they stiffen
    assert r1 > r10 > rs
    assert (r10 - rs) < 0.5 * (r1 - rs)


@pytest.mark.parametrize("bc", [4, 5])
def test_free_tube_under_pressure_is_the_membrane_solution(bc):
    X, T = tube(24, 8, r=0.5, length=1.0)
    Xg, Tg, ng = with_ghosts(X, T)
    h, p = 0.02, 1e-5
    kw = dict(ISO, thickness=h, pressure_final=p, damping_alpha=2.0, ghost_elems=ng, shell_boundary=bc, bc_every=1,
              dt=1.0)
    o = oracle.Oracle(Xg, Tg, **kw)
    om = o.cfl_omega_max()
    o.close()
    kw["dt"] = (0.3 if bc == 5 else 1.5) / om
    o = oracle.Oracle(Xg, Tg, **kw)
    _, u = settle(o, len(X), 0, tol=1e-7)
    o.close()
    rad = (u[:, 0] * X[:, 0] + u[:, 1] * X[:, 1]) / 0.5
    w0 = p * 0.5 ** 2 / (E * h)
    assert abs(rad.mean() / w0 - 1) < 0.02 and (rad.max() - rad.min()) < 0.01 * w0
    ez = np.polyfit(X[:, 2], u[:, 2], 1)[0]
    assert abs(ez / (-NU * p * 0.5 / (E * h)) - 1) < 0.02
