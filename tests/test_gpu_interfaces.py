"""The reference's other named interfaces through the C-ABI on a GPU
context: SubdivEvaluation::evaluate / evaluate_deriv at arbitrary points
(limit surface of the current positions X + u: the host weights applied to
the patch, and at the quadrature point the same point and tangents the
element tables hold), and ShellShellContact(eid1, eid2)::compute /
compute_on_mesh against the oracle's triangle distance."""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_subdiv_evaluate(icosphere):
    X, T = icosphere(3)
    X = imperfect(X, 1e-3, seed=51)
    g = S.ShellSystem(X, T)
    idx, tri = g.patches()
    rng = np.random.default_rng(3)
    u = rng.normal(size=X.shape) * 1e-2
    g.set_state(u.reshape(-1))
    Y = X + u
    o = oracle.Oracle(X, T)
    oidx, otri, N = o.patches()
    say("case subdiv evaluate L3")
    for e in list(range(0, 1280, 97)) + [int(np.argmax(N == 11))]:
        n = int((idx[e] >= 0).sum())
        for xi, eta in ((1 / 3, 1 / 3), (0.05, 0.02), (0.6, 0.3), (0.1, 0.7), (0.3, 0.2), (0.0, 0.0)):
            p, d1, d2 = S.subdiv_evaluate(g, e, xi, eta)
            w = S.subdiv_weights(n - 6, xi, eta)
            P = Y[idx[e, :n]]
            assert np.allclose(p, w[0] @ P, rtol=0, atol=1e-14)
            assert np.allclose(d1, w[1] @ P, rtol=0, atol=1e-12) and np.allclose(d2, w[2] @ P, rtol=0, atol=1e-12)
    # at the quadrature point of the undeformed shell: the tables' point and tangents
    g.set_state(np.zeros(X.size))
    tab = o.elem_tables()
    for e in range(0, 1280, 61):
        p, d1, d2 = S.subdiv_evaluate(g, e, 1 / 3, 1 / 3)
        assert np.allclose(p, tab[e, 98:101], rtol=0, atol=1e-13)
    with pytest.raises(S.ShellError):
        S.subdiv_evaluate(g, 0, 0.8, 0.4)
    g.close()
    o.close()


def test_contact_compute(icosphere):
    X, T = icosphere(3)
    g = S.ShellSystem(X, T, contact=1, cell_size_mult=1.0)
    _, tri = g.patches()
    Y = X.copy()
    Y[:, 2] *= 0.01
    u = Y - X
    g.set_state(u.reshape(-1))
    Y = X + u                                  # the positions the library forms
    rng = np.random.default_rng(4)
    say("case contact compute")
    for _ in range(200):
        e1, e2 = rng.integers(0, len(T), 2)
        r = S.contact_compute(g, int(e1), int(e2))
        ref = oracle.tri_distance(np.concatenate([Y[tri[e1]], Y[tri[e2]]]).reshape(-1))
        assert r["d2"] == ref[0] and np.array_equal(r["dc"], ref[1:4])
        assert r["v"][0] == ref[4] and r["v"][1] == ref[5] and r["w"][0] == ref[6] and r["w"][1] == ref[7]
        A = Y[tri[e1]]
        pt = (1 - ref[4] - ref[6]) * A[0] + ref[4] * A[1] + ref[6] * A[2] + 0.5 * ref[1:4]
        assert np.allclose(r["point"], pt, rtol=0, atol=1e-15)
        assert r["problem"] == (6 if ref[0] <= 0 else 0)
        m = S.contact_compute(g, int(e1), int(e2), on_mesh=True)
        refm = oracle.tri_distance(np.concatenate([X[tri[e1]], X[tri[e2]]]).reshape(-1))
        assert m["d2"] == refm[0] and m["problem"] == 0
    # two neighbours share a vertex: distance 0, a problem for compute() only
    e2 = next(k for k in range(1, len(T)) if set(tri[k]) & set(tri[0]))
    assert S.contact_compute(g, 0, e2)["problem"] == 6
    assert S.contact_compute(g, 0, e2, on_mesh=True)["problem"] == 0
    g.close()
