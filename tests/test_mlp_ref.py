"""Pins of the CPU reference of the warm-start MLP (oracle/mlp_ref.py): its
analytic gradient against central finite differences in FP64, and its
bfloat16 rounding against values fixed by the format."""
import numpy as np

from oracle import mlp_ref as R


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = R.bf16(x)
    # 8 significand bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    width = [4, 7, 5, 3]
    n = sum(width[l] * width[l + 1] + width[l + 1] for l in range(len(width) - 1))
    p = rng.normal(size=n) * 0.5
    x = rng.normal(size=(6, 4))
    t = rng.normal(size=(6, 3))
    loss, g = R.loss_and_grad(p, width, x, t, bf16_ops=False)
    eps = 1e-6
    for i in rng.choice(n, 25, replace=False):
        q = p.copy()
        q[i] += eps
        lp = R.loss_and_grad(q, width, x, t, bf16_ops=False)[0]
        q[i] -= 2 * eps
        lm = R.loss_and_grad(q, width, x, t, bf16_ops=False)[0]
        assert abs((lp - lm) / (2 * eps) - g[i]) <= 1e-7 * max(1.0, abs(g[i]))
