"""Shared builders for the parity tests: seeded structures and a sparse-GP
state (sparse set from a structure, Kuu + jitter Cholesky, alpha) fed to
both the oracle and the GPU path."""
import numpy as np


def fcc(a, reps, sigma, seed, nos=1):
    from flare_amd.workloads import FCC_BASIS, lattice
    cell, pos = lattice(FCC_BASIS, a, reps)
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(scale=sigma, size=pos.shape)
    species = rng.integers(0, nos, len(pos)).astype(np.int32)
    return cell, species, pos


def random_cell(n, seed, nos=2, scale=7.0, skew=0.3):
    rng = np.random.default_rng(seed)
    cell = np.eye(3) * scale + rng.uniform(-skew, skew, (3, 3))
    pos = rng.uniform(0, 1, (n, 3)) @ cell
    species = rng.integers(0, nos, n).astype(np.int32)
    return cell, species, pos


def ndp_kuu(X, norms, sigma, power):
    """Kuu of one species block in FP64 (normalized_dot_product.cpp:28-89)."""
    M = len(X)
    K = np.zeros((M, M))
    ok = norms >= 1e-8
    Xh = np.zeros_like(X)
    Xh[ok] = X[ok] / norms[ok, None]
    K = sigma ** 2 * (Xh @ Xh.T) ** power
    K[~ok, :] = 0
    K[:, ~ok] = 0
    return K


def model_state(desc_values, picks_by_type, sigma, power, seed, jitter=1e-8, alpha_scale=0.1):
    """Sparse rows picked per species from a descriptor container with
    .descriptors / .descriptor_norms lists; returns X list, norms list,
    alpha, L_inv."""
    rng = np.random.default_rng(seed)
    X = [desc_values.descriptors[s][np.asarray(p, int)] for s, p in enumerate(picks_by_type)]
    N = [desc_values.descriptor_norms[s][np.asarray(p, int)] for s, p in enumerate(picks_by_type)]
    M = sum(len(p) for p in picks_by_type)
    Kuu = np.zeros((M, M))
    o = 0
    for x, nr in zip(X, N):
        m = len(x)
        if m:
            Kuu[o:o + m, o:o + m] = ndp_kuu(x, nr, sigma, power)
        o += m
    L = np.linalg.cholesky(Kuu + jitter * np.eye(M))
    from scipy.linalg import solve_triangular
    Linv = solve_triangular(L, np.eye(M), lower=True)
    alpha = rng.normal(scale=alpha_scale, size=M)
    return X, N, alpha, Linv


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    if den == 0:
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / den)


def var_err(a, b, sig2):
    """Error of local variances sigma^2 - |L_inv k|^2 relative to sigma^2,
    the size of the cancelling terms (a conditioning-aware bound: the
    variances themselves can be many orders of magnitude smaller)."""
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / sig2 if np.size(a) else 0.0
