"""Synthetic workloads of the explicit shell path (host-side input
preparation: meshes, parameters, imperfections).

Meshes are icospheres built by recursive midpoint subdivision of the
icosahedron with projection to R = 1, face order and 6-significant-digit
coordinates identical to the reference's icosasphere3.off and
icosasphere4.off (the generator reproduces those files byte for byte).
"""
import numpy as np

H_DEFAULT = 0.01


def icosphere(level, round_g=True):
    t = (1.0 + np.sqrt(5.0)) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    X = []

    def add(p):
        p = np.asarray(p, dtype=np.float64)
        X.append(p / np.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]))
        return len(X) - 1
    for p in V:
        add(p)
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                cache[key] = add((X[a] + X[b]) / 2.0)
            return cache[key]
        F2 = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    X = np.array(X)
    if round_g:
        X = np.vectorize(lambda x: float("%g" % x))(X)
    return X, np.array(F, dtype=np.int32)


def radial_imperfection(X, std, seed):
    """seeded Gaussian radial perturbation of every vertex (standard deviation std)"""
    rng = np.random.default_rng(seed)
    n = X / np.linalg.norm(X, axis=1)[:, None]
    return X + n * rng.normal(size=(len(X), 1)) * std


def p_critical(E1, h, nu, R=1.0):
    """classical buckling pressure of a sphere, 2 E h^2 / (R^2 sqrt(3(1-nu^2)))"""
    return 2 * E1 * h * h / (R * R * np.sqrt(3 * (1 - nu * nu)))
