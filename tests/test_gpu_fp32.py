"""FP32 split-precision force path against the CPU FP64 oracle.

The FP32 path forms metric and curvature changes in FP32 from the
displacement gradients relative to the reference surface, keeps the
predictor, the volume and every reduction in FP64, and gathers FP32 slot
contributions.  Its error against FP64 is measured and reported here; the
bars are loose on purpose (single precision carries a relative rounding of
about 6e-8 per operation): forces within 1e-4 of the FP64 oracle, and a
trajectory within 1e-3 after 200 steps.
"""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01
PCR = 2 * 1.0 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("level", [3, 4])
def test_fp32_forces(icosphere, level):
    say("case fp32 forces level", level)
    X, T = icosphere(level)
    X = imperfect(X, 0.01 * H, seed=20 + level)
    kw = dict(pressure_final=-2 * PCR)
    g = S.ShellSystem(X, T, **kw)
    o = oracle.Oracle(X, T, **kw)
    rng = np.random.default_rng(level)
    for scale in (1e-5, 1e-3, 3e-2):
        u = rng.normal(size=3 * len(X)) * scale
        f32, fe32, en32, a32 = g.forces(u, t=0.5, precision=S.FP32_SPLIT)
        f64, fe64, en64, a64 = g.forces(u, t=0.5, precision=S.FP64)
        fo, feo, eno, ao = o.forces(u, t=0.5)
        e32, e64 = relerr(f32, fo), relerr(f64, fo)
        say(f"  scale {scale:g}: |f32-f|/|f| {e32:.2e}  |f64-f|/|f| {e64:.2e}  "
            f"E_tens {abs(en32[0] - eno[0]) / abs(eno[0]):.2e}  E_bend {abs(en32[1] - eno[1]) / abs(eno[1]):.2e}")
        assert e64 < 1e-10
        assert e32 < 1e-4
        assert relerr(fe32, feo) < 1e-4
        assert abs(en32[0] - eno[0]) <= 1e-4 * abs(eno[0]) and abs(en32[1] - eno[1]) <= 1e-4 * abs(eno[1])
    g.close()


def test_fp32_trajectory(icosphere):
    say("case fp32 trajectory L3")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=3)
    kw = dict(pressure_final=-2 * PCR, pressure_ramp_time=5.0, dt=0.1, damping_alpha=0.05)
    g = S.ShellSystem(X, T, **kw)
    o = oracle.Oracle(X, T, **kw)
    rng = np.random.default_rng(5)
    u0 = rng.normal(size=3 * len(X)) * 1e-4
    g.set_state(u0)
    o.set_state(u0)
    g.apply_initial_conditions(precision=S.FP32_SPLIT)
    o.init_dynamics()
    for k in range(4):
        g.advance(50, precision=S.FP32_SPLIT)
        o.step(50)
        ug = g.get_state()[0]
        uo = o.get_state()[0]
        say(f"  steps {50 * (k + 1)} |u32-u|/|u| {relerr(ug, uo):.2e}")
        assert relerr(ug, uo) < 1e-3
    g.close()


def test_fp32_equilibrium(icosphere):
    """deflated L3 shell relaxed in the FP32 split mode (FP32 forces until
    the residual of the deformation is within 1000x the tolerance, then
    FP64 to 1e-6): the equilibrium it reaches against the FP64 path's, and
    most steps taken in FP32"""
    say("case fp32 equilibrium L3")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=2024)
    kw = dict(bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=60.0, dt=0.1175, dt_max=0.1175,
              damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.3, mu_d=0.3, cell_margin=0.1)
    out = {}
    for prec in (S.FP64, S.FP32_SPLIT):
        g = S.ShellSystem(X, T, **kw)
        g.set_state(np.zeros(3 * len(X)))
        g.apply_initial_conditions(precision=prec)
        r = g.advance(30000, tol=1e-6, stop_on_converged=True, precision=prec)
        out[prec] = (g.get_state()[0], r)
        say(f"  precision {prec}: converged at {r['converged_step']} residual {r['residual_deform']:.2e} "
            f"E_bend {r['e_bend']:.8e} V {r['volume']:.8f}")
        g.close()
    (u64, r64), (u32, r32) = out[S.FP64], out[S.FP32_SPLIT]
    say(f"  FP32 steps {r32['n_steps_fp32']} of {r32['n_steps']}")
    assert r64["converged_step"] > 0 and r32["converged_step"] > 0
    assert r64["n_steps_fp32"] == 0 and r32["n_steps_fp32"] > r32["n_steps"] // 2
    say(f"  |u32-u64|/|u64| {relerr(u32, u64):.2e}")
    assert relerr(u32, u64) < 1e-3
    assert abs(r32["e_bend"] - r64["e_bend"]) <= 1e-3 * abs(r64["e_bend"])
