"""Committed golden fixtures (tests/golden/golden_L3.npz, written by
tests/golden/make_golden.py from the CPU oracle): the oracle must keep
reproducing them (CPU), and the HIP path is held to them (GPU): pair lists,
rebuild and rejection counts bit-exact, forces and states within the FP64
bar."""
import os
import sys

import numpy as np
import pytest

import oracle
import shell_amd as S

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as MG  # noqa: E402

G = np.load(os.path.join(HERE, "golden", "golden_L3.npz"))


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_oracle_reproduces_golden():
    fresh = MG.make()
    for k in G.files:
        if G[k].dtype.kind in "iu":
            assert np.array_equal(fresh[k], G[k]), k
        else:
            assert np.allclose(fresh[k], G[k], rtol=1e-13, atol=0), k


@pytest.mark.gpu
def test_gpu_against_golden():
    X, T = G["X"], G["T"]
    X0 = oracle.icosphere(3)[0]
    g = S.ShellSystem(X, T, **MG.FORCE_KW)
    f, fe, en, _ = g.forces(G["u"], t=0.5)
    print("forces rel", relerr(f, G["f"]), flush=True)
    assert relerr(f, G["f"]) < 1e-10 and relerr(fe, G["fext"]) < 1e-10
    assert np.allclose(en[:2], G["en"], rtol=1e-10)
    g.close()
    g = S.ShellSystem(X0, T, **MG.PAIR_KW)
    g.set_state(MG.squash(X0, 0.006))
    g.compute_shell_shell_contact_candidates(force=True)
    e1, e2 = g.pairs()
    assert np.array_equal(e1, G["pe1"]) and np.array_equal(e2, G["pe2"])
    g.close()
    g = S.ShellSystem(X, T, **MG.TRAJ_KW)
    g.set_state(G["traj_u0"])
    g.apply_initial_conditions()
    g.advance(100)
    print("trajectory rel", relerr(g.get_state()[0], G["traj_u"]), flush=True)
    assert relerr(g.get_state()[0], G["traj_u"]) < 1e-10
    g.close()
    g = S.ShellSystem(X, T, **MG.DYN_KW)
    g.set_state(G["dyn_u0"], G["dyn_v0"])
    g.apply_initial_conditions()
    r = g.advance(100)
    print("contact dynamics rel", relerr(g.get_state()[0], G["dyn_u"]), r["n_rebuilds"], r["n_rejected"], flush=True)
    assert [r["n_rebuilds"], r["n_rejected"]] == list(G["dyn_rep"][:2])
    assert relerr(g.get_state()[0], G["dyn_u"]) < 1e-9
    g.close()
