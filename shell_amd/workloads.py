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


# level-3/4/5 mix of the batched sweep and the deflation range; the ramp is
# given in time units so that shells with different stable steps see the
# same loading history
SWEEP_LEVELS = (3, 4, 5)
SWEEP_MIX = (2, 2, 1)
SWEEP_RAMP_TIME = 280.0


def sweep_counts(n_shells, mix=SWEEP_MIX):
    """shell counts per level for a mix ratio (largest remainder)"""
    w = np.asarray(mix, dtype=float) / sum(mix)
    c = np.floor(w * n_shells).astype(int)
    rest = n_shells - c.sum()
    order = np.argsort(-(w * n_shells - c))
    c[order[:rest]] += 1
    return [int(x) for x in c]


def batch_sweep(n_shells=512, seed=7, levels=SWEEP_LEVELS, mix=SWEEP_MIX, loading="volume"):
    """Seeded batch of independent post-buckling problems.

    Per shell: icosphere level by the mix (2:2:1 of 3/4/5 by default),
    E2/E1 ~ U[0.2, 1], h/R ~ logU[0.005, 0.02], mu ~ U[0, 0.5], Gaussian
    radial imperfection 0.01 h with its own seed, self-contact with friction
    (contact lists rebuilt on the displacement trigger), mass damping 0.05.
    loading "volume": deflation (bulk modulus 0.05, enclosed volume of the
    force configuration) to V/V0 ~ U[0.85, 0.92] over SWEEP_RAMP_TIME;
    loading "pressure": external pressure ramped from 0 to p/p_cr ~
    U[1.5, 2.5] over SWEEP_RAMP_TIME (the factors come from a second stream
    seeded by seed + 1, so the other draws are those of the volume loading).
    Returns a list of dicts with keys level, X, T, params (dt left to the
    caller: it depends on the CFL bound of each shell)."""
    if loading not in ("volume", "pressure"):
        raise ValueError(f"unknown loading {loading!r}")
    rng = np.random.default_rng(seed)
    prng = np.random.default_rng(seed + 1)
    counts = sweep_counts(n_shells, mix)
    lv = np.concatenate([[L] * c for L, c in zip(levels, counts)])
    meshes = {L: icosphere(L) for L in set(lv.tolist())}
    out = []
    for k, L in enumerate(lv):
        e2 = rng.uniform(0.2, 1.0)
        h = float(np.exp(rng.uniform(np.log(0.005), np.log(0.02))))
        mu = rng.uniform(0.0, 0.5)
        ratio = rng.uniform(0.85, 0.92)
        iseed = int(rng.integers(1 << 30))
        X, T = meshes[int(L)]
        params = dict(E1=1.0, E2=e2, nu=0.3, G12=0.35, thickness=h, rho=1.0, damping_alpha=0.05,
                      bulk_modulus=0.05, volume_final_ratio=ratio, volume_ramp_time=SWEEP_RAMP_TIME,
                      volume_at_prediction=1, contact=1, mu_s=mu, mu_d=mu, cell_size_mult=4.0,
                      cell_margin=0.1, rebuild_trigger=1)
        if loading == "pressure":
            f = prng.uniform(1.5, 2.5)
            params.update(bulk_modulus=0.0, volume_final_ratio=1.0, volume_ramp_time=0.0, volume_at_prediction=0,
                          pressure_final=-f * p_critical(1.0, h, 0.3), pressure_ramp_time=SWEEP_RAMP_TIME)
        out.append(dict(level=int(L), X=radial_imperfection(X, 0.01 * h, iseed), T=T, params=params,
                        imperfection_seed=iseed))
    return out
