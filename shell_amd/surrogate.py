"""Learned warm start of the explicit relaxation.

A small MLP maps a shell's load and material parameters to the principal-
component coefficients of its post-buckled displacement field; the predicted
field seeds the damped explicit relaxation, which remains the judge: a
seeded shell that interpenetrates (CONTACT_PROBLEM at the start) or whose
residual grows instead of falling restarts from the undeformed state with
the load ramp (the plain start).

The network is trained and evaluated by the library's own kernels (BF16
matrix-core GEMMs with FP32 accumulation, Adam) through the shx_mlp_* entry
points; the principal components are computed on the host (numpy SVD of the
training displacements).
"""
import ctypes

import numpy as np

from . import _D, lib, ShellError

_F = ctypes.POINTER(ctypes.c_float)
_declared = False


def _declare():
    global _declared
    if _declared:
        return
    L = lib()
    V = ctypes.c_void_p
    L.shx_mlp_create.argtypes = [ctypes.POINTER(V), ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_uint64,
                                 ctypes.c_int]
    L.shx_mlp_destroy.argtypes = [V]
    L.shx_mlp_last_error.argtypes = [V]
    L.shx_mlp_last_error.restype = ctypes.c_char_p
    L.shx_mlp_num_params.argtypes = [V]
    L.shx_mlp_get_params.argtypes = [V, _F]
    L.shx_mlp_set_params.argtypes = [V, _F]
    L.shx_mlp_forward.argtypes = [V, ctypes.c_int, _F, _F]
    L.shx_mlp_gradient.argtypes = [V, ctypes.c_int, _F, _F, _F, ctypes.POINTER(ctypes.c_double)]
    L.shx_mlp_train.argtypes = [V, ctypes.c_int, _F, _F, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.shx_mlp_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, _F, ctypes.c_int, _F]
    _declared = True


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(_F)


def gemm(A, B, ta=False, tb=False):
    """C = op(A) op(B) on the matrix cores (BF16 operands, FP32 accumulation)"""
    _declare()
    A, B = _f32(A), _f32(B)
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    C = np.zeros((M, N), dtype=np.float32)
    r = lib().shx_mlp_gemm(M, N, K, fp(A), int(ta), fp(B), int(tb), fp(C))
    if r != 0:
        raise ShellError(f"shx_mlp_gemm: {r}")
    return C


class MLP:
    """width[0] -> ... -> width[-1], GELU on hidden layers"""

    def __init__(self, width, seed=1, device=0):
        _declare()
        self.width = list(width)
        w = (ctypes.c_int * len(width))(*width)
        h = ctypes.c_void_p()
        r = lib().shx_mlp_create(ctypes.byref(h), len(width) - 1, w, seed, device)
        if r != 0:
            raise ShellError(f"shx_mlp_create: {r}: {lib().shx_mlp_last_error(None).decode()}")
        self.h = h
        self.nparam = lib().shx_mlp_num_params(h)

    def _check(self, r, what):
        if r != 0:
            raise ShellError(f"{what}: {r}: {lib().shx_mlp_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            lib().shx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._check(lib().shx_mlp_get_params(self.h, fp(p)), "shx_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        self._check(lib().shx_mlp_set_params(self.h, fp(p)), "shx_mlp_set_params")

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((len(x), self.width[-1]), dtype=np.float32)
        self._check(lib().shx_mlp_forward(self.h, len(x), fp(x), fp(y)), "shx_mlp_forward")
        return y

    def gradient(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(lib().shx_mlp_gradient(self.h, len(x), fp(x), fp(t), fp(g), ctypes.byref(loss)),
                    "shx_mlp_gradient")
        return loss.value, g

    def train(self, x, t, batch=128, epochs=100, lr=1e-3, seed=7):
        x, t = _f32(x), _f32(t)
        loss = ctypes.c_double()
        self._check(lib().shx_mlp_train(self.h, len(x), fp(x), fp(t), batch, epochs, lr, seed, ctypes.byref(loss)),
                    "shx_mlp_train")
        return loss.value


def pca(U, k):
    """principal components of the rows of U (samples x DOF): mean, the
    first k right singular vectors (k x DOF, orthonormal rows), the
    coefficients of every sample and the fraction of the variance kept"""
    U = np.asarray(U, np.float64)
    umean = U.mean(0)
    _, s, vt = np.linalg.svd(U - umean, full_matrices=False)
    k = min(k, vt.shape[0])
    basis = vt[:k]
    return umean, basis, (U - umean) @ basis.T, float(np.sum(s[:k] ** 2) / np.sum(s ** 2))


class Surrogate:
    """parameters -> principal-component coefficients of the equilibrium
    displacement -> displacement field.  Inputs and coefficients are
    standardised (zero mean, unit variance over the training set) around the
    network."""

    def __init__(self, umean, basis, hidden=(256, 256, 256), seed=1, device=0):
        self.umean = np.asarray(umean, np.float64)
        self.basis = np.asarray(basis, np.float64)
        self.k = self.basis.shape[0]
        self.hidden = tuple(hidden)
        self.seed = seed
        self.device = device
        self.net = None

    def fit(self, P, C, epochs=400, batch=128, lr=1e-3):
        """P: samples x parameters; C: samples x k coefficients.  Returns the
        final training loss (standardised units)"""
        P = np.asarray(P, np.float64)
        C = np.asarray(C, np.float64)
        self.pmean, self.pstd = P.mean(0), P.std(0) + 1e-12
        self.cmean, self.cstd = C.mean(0), C.std(0) + 1e-12
        X = (P - self.pmean) / self.pstd
        T = (C - self.cmean) / self.cstd
        if self.net is None:
            self.net = MLP([P.shape[1], *self.hidden, self.k], seed=self.seed, device=self.device)
        return self.net.train(X, T, batch=min(batch, len(X)), epochs=epochs, lr=lr)

    def coefficients(self, P):
        X = (np.asarray(P, np.float64) - self.pmean) / self.pstd
        return self.net.forward(X).astype(np.float64) * self.cstd + self.cmean

    def predict(self, P):
        return self.umean + self.coefficients(P) @ self.basis
