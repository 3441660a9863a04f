"""Pins the oracle's NormalizedDotProduct / SparseGP path
(normalized_dot_product.cpp, sparse_gp.cpp, lammps_descriptor.cpp,
pair_flare.cpp): forces and stress against finite differences
(ctests/test_kernels.cpp:140-406), sum rules (:35-75), closed forms with
one sparse environment, empty environments (tests/get_sgp.py:41-56), the
literal envs_struc path against the per-atom adjoint, the mapped-beta
identity of tests/test_lammps.py:33-120, the local-variance identities, and
DTC/SOR diagonals against full-matrix numpy."""
import numpy as np
import pytest

from helpers import model_state, random_cell, rel_err

NOS, NMAX, LMAX = 2, 4, 3
CUT = np.array([[4.2, 3.8], [3.8, 4.0]])
SIG, PW = 1.7, 2.0


def build(oracle, cell, species, pos, rc=4.2, dervs=True):
    nb = oracle.Neighbors(cell, species, pos, rc, "literal")
    return nb, oracle.B2Values(nb, NOS, NMAX, LMAX, CUT, with_dervs=dervs)


@pytest.fixture(scope="module")
def system(oracle):
    cell, species, pos = random_cell(22, 31, nos=2, scale=7.0, skew=0.4)
    nb, d = build(oracle, cell, species, pos)
    picks = [list(range(min(4, d.n_clusters_by_type[s]))) for s in range(NOS)]
    X, N, alpha, Linv = model_state(d, picks, SIG, PW, seed=5)
    M = len(alpha)
    Sigma = np.diag(np.random.default_rng(6).uniform(0.01, 0.02, M))
    Rinv = np.linalg.cholesky(Sigma)
    m = oracle.Model(SIG, PW, X, N, alpha, L_inv=Linv, Kuu_inverse=Linv.T @ Linv, Sigma=Sigma,
                     R_inv=Rinv)
    return dict(cell=cell, species=species, pos=pos, nb=nb, d=d, m=m, X=X, N=N, alpha=alpha,
                Linv=Linv, Sigma=Sigma, Rinv=Rinv)


def test_forces_fd(oracle, system):
    cell, species, pos, m = system["cell"], system["species"], system["pos"], system["m"]
    mean, _ = m.predict(system["d"], "mean")
    n = len(pos)
    F = mean[1:1 + 3 * n].reshape(n, 3)
    h = 1e-5
    for a in (0, 7, 15):
        for k in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, k] += h
            pm[a, k] -= h
            ep = m.predict(build(oracle, cell, species, pp)[1], "mean")[0][0]
            em = m.predict(build(oracle, cell, species, pm)[1], "mean")[0][0]
            assert F[a, k] == pytest.approx(-(ep - em) / (2 * h), abs=2e-6)


def test_stress_fd(oracle, system):
    cell, species, pos, m = system["cell"], system["species"], system["pos"], system["m"]
    mean, _ = m.predict(system["d"], "mean")
    S = mean[-6:]
    V = abs(np.linalg.det(cell))
    h = 1e-5
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (a, b) in enumerate(comps):
        es = []
        for dl in (h, -h):
            c, p = cell.copy(), pos.copy()
            # ctests/test_kernels.cpp strain: cell(:, a) += cell(:, b) delta.
            c[:, a] += c[:, b] * dl
            p[:, a] += p[:, b] * dl
            es.append(m.predict(build(oracle, c, species, p)[1], "mean")[0][0])
        assert S[q] * V == pytest.approx(-(es[0] - es[1]) / (2 * h), abs=2e-6)


def test_sum_rules(oracle, system):
    m, d = system["m"], system["d"]
    n = len(system["pos"])
    mean, _ = m.predict(d, "mean")
    assert np.abs(mean[1:1 + 3 * n].reshape(n, 3).sum(0)).max() < 1e-12
    K = m.envs_struc(d)
    # Energy column = sum over clusters of envs_envs (ctests/test_kernels.cpp:55-75).
    Kc = m.envs_envs(d)
    np.testing.assert_allclose(K[:, 0], Kc.sum(0), rtol=1e-12)
    np.testing.assert_allclose(mean, K.T @ system["alpha"], rtol=1e-12, atol=1e-13)


def test_literal_equals_scalable(oracle, system):
    m, nb, d = system["m"], system["nb"], system["d"]
    mean, var = m.predict(d, "local")
    mean2, var2, le = m.predict_scalable(nb, NOS, NMAX, LMAX, CUT)
    assert rel_err(mean2, mean) < 1e-12
    order = np.concatenate(d.atom_indices)
    assert rel_err(var2[order], var) < 1e-12
    assert le.sum() == pytest.approx(mean[0], rel=1e-13)


def test_single_sparse_env_closed_form(oracle, system):
    d = system["d"]
    x = d.descriptors[0][1:2]
    xn = d.descriptor_norms[0][1:2]
    a = np.array([0.37])
    m = oracle.Model(SIG, PW, [x, np.zeros((0, d.n_descriptors))], [xn, np.zeros(0)], a)
    mean, _ = m.predict(d, "mean")
    B, Bn = d.descriptors[0], d.descriptor_norms[0]
    expect = SIG ** 2 * a[0] * np.sum(((B @ x[0]) / (Bn * xn[0])) ** PW)
    assert mean[0] == pytest.approx(expect, rel=1e-13)


def test_empty_environment(oracle):
    # tests/get_sgp.py:41-56: an isolated atom has no neighbours.
    cell = np.eye(3) * 30.0
    pos = np.array([[0, 0, 0], [1, 1, 1], [15, 15, 15]], dtype=float)
    species = np.array([0, 1, 1], np.int32)
    nb, d = build(oracle, cell, species, pos)
    X = [d.descriptors[0][:1], d.descriptors[1][:2]]
    N = [d.descriptor_norms[0][:1], d.descriptor_norms[1][:2]]
    assert N[1][1] == 0.0  # the isolated atom
    from helpers import ndp_kuu
    Kuu = np.zeros((3, 3))
    Kuu[:1, :1] = ndp_kuu(X[0], N[0], SIG, PW)
    Kuu[1:, 1:] = ndp_kuu(X[1], N[1], SIG, PW)
    Linv = np.linalg.inv(np.linalg.cholesky(Kuu + 1e-8 * np.eye(3)))
    m = oracle.Model(SIG, PW, X, N, np.array([0.3, -0.2, 0.9]), L_inv=Linv)
    mean, var = m.predict(d, "local")
    assert np.all(mean[1 + 3 * 2:1 + 3 * 3] == 0)  # no force on the isolated atom
    assert var[2] == 0.0  # sparse_gp.cpp:133-135: empty environment, zero variance


def test_local_variance_identities(oracle, system):
    m, d = system["m"], system["d"]
    _, var = m.predict(d, "local")
    # Atoms that are sparse environments have jitter-level variance.
    for s, off in ((0, 0), (1, d.n_clusters_by_type[0])):
        for r in range(min(4, d.n_clusters_by_type[s])):
            assert abs(var[off + r]) < 1e-6
    assert np.all(var <= SIG ** 2 + 1e-12)
    # With no same-species sparse environment, variance = sigma^2.
    X = [system["X"][0], np.zeros((0, d.n_descriptors))]
    N = [system["N"][0], np.zeros(0)]
    M0 = len(X[0])
    mm = oracle.Model(SIG, PW, X, N, system["alpha"][:M0], L_inv=system["Linv"][:M0, :M0])
    _, v2 = mm.predict(d, "local")
    np.testing.assert_allclose(v2[d.n_clusters_by_type[0]:], SIG ** 2, rtol=1e-12)


def test_dtc_and_sor_diagonals(oracle, system):
    m, d = system["m"], system["d"]
    K = m.envs_struc(d)
    kself = m.self_kernel(d)
    mean, vd = m.predict(d, "dtc")
    Kinv = system["Linv"].T @ system["Linv"]
    expect = kself - np.einsum("ic,ij,jc->c", K, Kinv, K) + np.einsum("ic,ij,jc->c", K,
                                                                     system["Sigma"], K)
    assert rel_err(vd, expect) < 1e-10
    _, vs = m.predict(d, "sor")
    V = K.T @ system["Rinv"]
    assert rel_err(vs, np.sum(V * V, axis=1)) < 1e-12
    # Energy entry of the self kernel = sum_ij k(B_i, B_j) (closed form).
    E = 0
    for s in range(NOS):
        B, Bn = d.descriptors[s], d.descriptor_norms[s]
        ok = Bn >= 1e-8
        Bh = B[ok] / Bn[ok, None]
        E += SIG ** 2 * np.sum((Bh @ Bh.T) ** PW)
    assert kself[0] == pytest.approx(E, rel=1e-12)
    # DTC variances are non-negative up to jitter (ctests/test_sparse_gp.cpp:90-100).
    assert np.all(vd > -1e-6)


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_mapped_beta_identity(oracle, system, power):
    """pair_style flare on the written coefficient file reproduces the
    sparse-GP energies and forces (tests/test_lammps.py:33-120)."""
    d, nb = system["d"], system["nb"]
    X, N, alpha = system["X"], system["N"], system["alpha"]
    m = oracle.Model(SIG, power, X, N, alpha)
    mean, _ = m.predict(d, "mean")
    text = m.file_text(NOS, NMAX, LMAX, CUT)
    e, F, vir = oracle.pair_flare(text, nb)
    n = len(system["pos"])
    # Cutoffs are written with 2 decimals (b2.cpp:66-73); these are exact.
    assert e.sum() == pytest.approx(mean[0], rel=1e-10)
    assert rel_err(F.reshape(-1), mean[1:1 + 3 * n]) < 1e-10
    V = abs(np.linalg.det(system["cell"]))
    assert rel_err(vir / V, mean[-6:]) < 1e-10


def test_coefficient_file_grammar(oracle, system):
    m = system["m"]
    text = m.file_text(NOS, NMAX, LMAX, CUT, contributor="Jane")
    lines = text.split("\n")
    assert lines[0] == "DATE: Thu Jan  1 00:00:00 1970 CONTRIBUTOR: Jane"
    assert lines[1] == "2 NormalizedDotProduct"
    assert lines[2] == "chebyshev"
    nd = system["d"].n_descriptors
    bs = nd * (nd + 1) // 2
    assert lines[3] == f"2 {NMAX} {LMAX} {bs}"
    assert lines[4] == "quadratic"
    assert lines[5] == "4.20 3.80 3.80 4.00 "
    body = lines[6:]
    # 5 values per line, counter not reset at the species boundary
    # (sparse_gp.cpp:1233-1260): species 0 ends mid-line.
    vals = " ".join(body).split()
    assert len(vals) == 2 * bs
    assert all(len(l.split()) <= 5 for l in body)
    beta = m.mapping()
    np.testing.assert_array_equal(np.array(vals, dtype=float), beta.reshape(-1))
    for l in body:
        for tok in l.split(" ")[1:]:
            pass
    first = body[0]
    assert first.startswith("  ") or first.startswith(" -")


def _pair_kernel_sum(d1, d2, power):
    """sum over same-species atom pairs i of structure 1, j of structure 2 of
    k(B_i, B_j) = sigma^2 (B_i . B_j / |B_i||B_j|)^p, empty environments
    skipped (normalized_dot_product.cpp:28-89 summed over both structures)."""
    tot = 0.0
    for s in range(NOS):
        A, An = d1.descriptors[s], d1.descriptor_norms[s]
        B, Bn = d2.descriptors[s], d2.descriptor_norms[s]
        a_ok, b_ok = An >= 1e-8, Bn >= 1e-8
        if not a_ok.any() or not b_ok.any():
            continue
        Ah = A[a_ok] / An[a_ok, None]
        Bh = B[b_ok] / Bn[b_ok, None]
        tot += np.sum(SIG ** 2 * (Ah @ Bh.T) ** power)
    return tot


def _mixed_second_difference(oracle, displaced, power, h):
    """d^2 K(t1, t2) / dt1 dt2 at 0 for K(t1, t2) = sum_ij k(B_i(t1), B_j(t2)),
    where displaced(t) returns (cell, species, positions) moved by t along
    one coordinate: central differences at steps h and 2h combined by
    Richardson extrapolation (error O(h^4))."""
    def D(step):
        ds = {t: build(oracle, *displaced(t), dervs=False)[1] for t in (step, -step)}
        k = lambda a, b: _pair_kernel_sum(ds[a], ds[b], power)
        return (k(step, step) - k(step, -step) - k(-step, step) + k(-step, -step)) / (4 * step ** 2)
    return (4 * D(h) - D(2 * h)) / 3


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_self_kernel_forces_mixed_fd(oracle, system, power):
    """The force entries of self_kernel_struc (normalized_dot_product.cpp:
    450-586) are the variances of the forces under the structure-structure
    kernel: entry (a, m) = d^2 K / dx_am dx'_am with both copies displaced
    independently (the ctests/test_kernels.cpp:140-406 recipe applied to
    the second derivative)."""
    cell, species, pos = system["cell"], system["species"], system["pos"]
    m = oracle.Model(SIG, power, system["X"], system["N"], system["alpha"])
    kself = m.self_kernel(system["d"])
    for a, c in ((0, 0), (7, 1), (15, 2), (21, 0)):
        def moved(t, a=a, c=c):
            p = pos.copy()
            p[a, c] += t
            return cell, species, p
        fd = _mixed_second_difference(oracle, moved, power, 1e-3)
        assert kself[1 + 3 * a + c] == pytest.approx(fd, rel=1e-7, abs=1e-9)


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_self_kernel_stress_mixed_fd(oracle, system, power):
    """Stress entries: entry (m, n) = d^2 K / de de' / V^2 for the strain
    cell(:, m) += cell(:, n) e, positions likewise (the strain of the
    stress FD test above), applied to each copy independently."""
    cell, species, pos = system["cell"], system["species"], system["pos"]
    n_at = len(pos)
    V = abs(np.linalg.det(cell))
    m = oracle.Model(SIG, power, system["X"], system["N"], system["alpha"])
    kself = m.self_kernel(system["d"])
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (a, b) in enumerate(comps):
        def strained(t, a=a, b=b):
            c, p = cell.copy(), pos.copy()
            c[:, a] += c[:, b] * t
            p[:, a] += p[:, b] * t
            return c, species, p
        fd = _mixed_second_difference(oracle, strained, power, 1e-4) / V ** 2
        assert kself[1 + 3 * n_at + q] == pytest.approx(fd, rel=1e-6, abs=1e-10)
