"""The warm-start MLP's kernels on the GPU: matrix-core GEMM (BF16 operands,
FP32 accumulation) against the CPU reference with the same operand
rounding, forward pass and parameter gradients of the 6-256-256-256-64
network against oracle/mlp_ref.py, and training."""
import numpy as np
import pytest

from oracle import mlp_ref as R
import shell_amd as S
from shell_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("M,N,K,ta,tb", [(16, 16, 16, 0, 0), (37, 21, 45, 0, 0), (64, 256, 6, 0, 0),
                                         (256, 64, 128, 1, 0), (128, 100, 64, 0, 1), (5, 300, 33, 1, 1)])
def test_mfma_gemm_matches_reference(M, N, K, ta, tb):
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    C = SG.gemm(A, B, ta=bool(ta), tb=bool(tb))
    Ar = R.bf16(A.T if ta else A)
    Br = R.bf16(B.T if tb else B)
    ref = Ar @ Br
    err = np.max(np.abs(C - ref)) / np.max(np.abs(ref))
    say(f"  gemm {M}x{N}x{K} ta={ta} tb={tb}: max rel err {err:.2e}")
    assert err < 1e-5


def test_mlp_forward_and_gradient_match_reference():
    width = [6, 256, 256, 256, 64]
    net = SG.MLP(width, seed=5)
    rng = np.random.default_rng(9)
    x = rng.normal(size=(96, 6)).astype(np.float32)
    t = rng.normal(size=(96, 64)).astype(np.float32)
    p = net.params().astype(np.float64)
    y = net.forward(x)
    acts, _ = R.forward(p, width, x)
    fe = np.linalg.norm(y - acts[-1]) / np.linalg.norm(acts[-1])
    loss, g = net.gradient(x, t)
    lr, gr = R.loss_and_grad(p, width, x, t)
    ge = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say(f"  forward rel err {fe:.2e}, loss {loss:.6f} vs {lr:.6f}, gradient rel err {ge:.2e}")
    assert fe < 1e-4
    assert abs(loss - lr) <= 1e-4 * lr
    assert ge < 1e-3
    # and within bfloat16 distance of the exact FP64 gradient
    g64 = R.loss_and_grad(p, width, x, t, bf16_ops=False)[1]
    say(f"  gradient vs exact FP64: {np.linalg.norm(g - g64) / np.linalg.norm(g64):.2e}")
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2
    net.close()


def test_mlp_learns_a_smooth_map():
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(1024, 6)).astype(np.float32)
    M = rng.normal(size=(6, 8))
    t = np.sin(x @ M).astype(np.float32)
    net = SG.MLP([6, 64, 64, 8], seed=3)
    l0 = net.gradient(x[:256], t[:256])[0]
    l1 = net.train(x, t, batch=128, epochs=300, lr=3e-3)
    say(f"  loss {l0:.4f} -> {l1:.5f}")
    assert l1 < l0 / 50
    net.close()


def test_mlp_train_refuses_a_minibatch_larger_than_the_set():
    """batch > n would run no optimizer step: refused with SHX_E_ARG, the
    parameters untouched"""
    net = SG.MLP([6, 32, 8], seed=5)
    p0 = net.params()
    x = np.zeros((10, 6), np.float32)
    t = np.zeros((10, 8), np.float32)
    with pytest.raises(S.ShellError, match="minibatch larger"):
        net.train(x, t, batch=16, epochs=3)
    assert np.array_equal(net.params(), p0)
    net.close()
