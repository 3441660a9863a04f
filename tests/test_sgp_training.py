"""The sparse-GP training state on the host (sparse_gp.cpp:343-641,
759-1198): Kuu / Kuf bookkeeping as sparse environments and training
structures are added in either order (ctests/test_sparse_gp.cpp AddOrder),
the QR producer of alpha / Sigma / R_inv against the normal-equation closed
form, the stable likelihood against the direct O(n^3) one and its gradient
against finite differences and the direct gradient (LikeGradStable), and
set_hyperparameters (Set_Hyps).

The kernel matrices come from the oracle here (a stand-in for the GPU's
envs_struc, which the -m gpu tests exercise); everything under test is the
product's host code in flare_amd/_C_flare.py."""
import numpy as np
import pytest

from helpers import random_cell

NOS, NMAX, LMAX = 2, 4, 3
CUT = np.array([[4.2, 3.8], [3.8, 4.0]])


@pytest.fixture(scope="module")
def C():
    from flare_amd import _C_flare
    return _C_flare


class FakeStructure:
    """Stands in for _C_flare.Structure with oracle descriptors."""

    def __init__(self, oracle, cell, species, pos, seed):
        nb = oracle.Neighbors(cell, species, pos, 4.2, "literal")
        self._nb = nb
        self.descriptors = [oracle.B2Values(nb, NOS, NMAX, LMAX, CUT, with_dervs=True)]
        self.noa = len(pos)
        rng = np.random.default_rng(seed)
        self.energy = rng.normal(size=1)
        self.forces = rng.normal(size=3 * self.noa)
        self.stresses = rng.normal(size=6)
        self.form = 0
        self.descriptor_calculators = []


def oracle_kernel(C, oracle, sigma, power):
    class OracleNDP(C.NormalizedDotProduct):
        def envs_struc(self, envs, struc, hyps=None):
            X = [np.asarray(x) for x in envs.descriptors]
            N = [np.asarray(x) for x in envs.descriptor_norms]
            m = oracle.Model(self.sigma, self.power, X, N, np.zeros(envs.n_clusters))
            return m.envs_struc(struc)
    return OracleNDP(sigma, power)


@pytest.fixture(scope="module")
def strucs(oracle):
    out = []
    for k, (n, seed) in enumerate(((11, 31), (9, 32))):
        cell, species, pos = random_cell(n, seed, nos=2, scale=6.0, skew=0.3)
        out.append(FakeStructure(oracle, cell, species, pos, 100 + k))
    return out


def trained(C, oracle, strucs, sigma=1.7, power=2.0, noise=(0.1, 0.2, 0.3)):
    sgp = C.SparseGP([oracle_kernel(C, oracle, sigma, power)], *noise)
    sgp.add_training_structure(strucs[0], [-1], 0.4, 0.2, 0.3)
    sgp.add_specific_environments(strucs[0], [0, 1, 3])
    sgp.add_training_structure(strucs[1], [0, 1, 3, 5], 0.64, 0.55, 0.45)
    sgp.add_specific_environments(strucs[1], [2, 3, 4])
    sgp.update_matrices_QR()
    return sgp


def test_add_order(C, oracle, strucs):
    """Kuf and Kuu do not depend on whether labels or sparse environments
    come first (ctests/test_sparse_gp.cpp AddOrder)."""
    a = C.SparseGP([oracle_kernel(C, oracle, 1.5, 2.0)], 1.0, 2.0, 3.0)
    b = C.SparseGP([oracle_kernel(C, oracle, 1.5, 2.0)], 1.0, 2.0, 3.0)
    a.add_training_structure(strucs[0])
    a.add_all_environments(strucs[0])
    a.add_training_structure(strucs[1], [2, 4])
    a.add_specific_environments(strucs[1], [5, 1, 1])
    b.add_all_environments(strucs[0])
    b.add_specific_environments(strucs[1], [5, 1, 1])
    b.add_training_structure(strucs[0])
    b.add_training_structure(strucs[1], [2, 4])
    np.testing.assert_array_equal(a.Kuf, b.Kuf)
    np.testing.assert_array_equal(a.Kuu, b.Kuu)
    # Sparse rows are species-major; the atom listed twice is added twice.
    cd = a.sparse_descriptors[0]
    assert a.n_sparse == strucs[0].noa + 3 == cd.n_clusters
    assert a.Kuu.shape == (a.n_sparse, a.n_sparse)
    # Kuu equals the kernel of the final sparse set with itself.
    kern = a.kernels[0]
    np.testing.assert_allclose(a.Kuu, kern.envs_envs(cd, cd), rtol=1e-13, atol=1e-14)
    # Kuf columns are the envs_struc columns of each structure's labels.
    n0 = strucs[0].noa
    K0 = kern.envs_struc(cd, strucs[0].descriptors[0])
    np.testing.assert_allclose(a.Kuf[:, :1 + 3 * n0 + 6], K0, rtol=1e-13, atol=1e-14)
    K1 = kern.envs_struc(cd, strucs[1].descriptors[0])
    sel = [0, 7, 8, 9, 13, 14, 15] + list(range(1 + 3 * strucs[1].noa, K1.shape[1]))
    np.testing.assert_allclose(a.Kuf[:, 1 + 3 * n0 + 6:], K1[:, sel], rtol=1e-13, atol=1e-14)
    assert a.n_labels == (1 + 3 * n0 + 6) + (1 + 6 + 6)
    assert list(a.label_count) == [0, 1 + 3 * n0 + 6, a.n_labels]


def test_qr_matches_normal_equations(C, oracle, strucs):
    """alpha = (Kuu + jitter I + Kuf Lambda Kfu)^-1 Kuf Lambda y and
    Sigma = (Kuu + jitter I + Kuf Lambda Kfu)^-1 (the QR route of
    sparse_gp.cpp:607-641 solves exactly these normal equations)."""
    sgp = trained(C, oracle, strucs)
    M = sgp.n_sparse
    lam = sgp.noise_vector
    A = sgp.Kuu + sgp.Kuu_jitter * np.eye(M) + (sgp.Kuf * lam) @ sgp.Kuf.T
    alpha = np.linalg.solve(A, sgp.Kuf @ (lam * sgp.y))
    assert np.max(np.abs(sgp.alpha - alpha)) / np.max(np.abs(alpha)) < 1e-8
    Sig = np.linalg.inv(A)
    assert np.max(np.abs(sgp.Sigma - Sig)) / np.max(np.abs(Sig)) < 1e-8
    np.testing.assert_allclose(sgp.R_inv @ sgp.R_inv.T, sgp.Sigma, rtol=1e-12, atol=1e-14)
    assert np.allclose(np.tril(sgp.R_inv, -1), 0)  # inverse of the upper factor
    L = np.linalg.cholesky(sgp.Kuu + sgp.Kuu_jitter * np.eye(M))
    np.testing.assert_allclose(sgp.L_inv @ L, np.eye(M), atol=1e-8)
    np.testing.assert_allclose(sgp.Kuu_inverse, sgp.L_inv.T @ sgp.L_inv, rtol=1e-14)
    # Noise vector: 1 / (sigma_x rel_x)^2 per label kind (sparse_gp.cpp:518-524).
    n0 = strucs[0].noa
    assert sgp.noise_vector[0] == pytest.approx(1 / (0.1 * 0.4) ** 2)
    assert sgp.noise_vector[1] == pytest.approx(1 / (0.2 * 0.2) ** 2)
    assert sgp.noise_vector[1 + 3 * n0] == pytest.approx(1 / (0.3 * 0.3) ** 2)


def test_likelihoods_agree(C, oracle, strucs):
    """ctests/test_sparse_gp.cpp:102-118: stable, direct and gradient-path
    likelihoods agree."""
    sgp = trained(C, oracle, strucs)
    sgp.compute_likelihood()
    like1 = sgp.log_marginal_likelihood
    assert like1 == sgp.data_fit + sgp.complexity_penalty + sgp.constant_term
    sgp.compute_likelihood_stable()
    like2 = sgp.log_marginal_likelihood
    like3 = sgp.compute_likelihood_gradient(sgp.hyperparameters)
    assert like1 == pytest.approx(like2, abs=1e-6, rel=1e-9)
    assert like1 == pytest.approx(like3, abs=1e-6, rel=1e-9)


def test_likelihood_gradient_stable(C, oracle, strucs):
    """ctests/test_sparse_gp.cpp LikeGradStable: the stable gradient (with
    and without precomputed KnK) against central differences and against
    the direct gradient."""
    sgp = trained(C, oracle, strucs)
    hyps = sgp.hyperparameters.copy()
    sgp.set_hyperparameters(hyps)
    sgp.precompute_KnK()
    like = sgp.compute_likelihood_gradient_stable(True)
    g_pre = sgp.likelihood_gradient.copy()
    like_np = sgp.compute_likelihood_gradient_stable(False)
    g = sgp.likelihood_gradient.copy()
    assert like == pytest.approx(like_np, rel=1e-12)
    np.testing.assert_allclose(g_pre, g, rtol=1e-8)
    like_direct = sgp.compute_likelihood_gradient(hyps)
    g_direct = sgp.likelihood_gradient.copy()
    assert like == pytest.approx(like_direct, abs=1e-6, rel=1e-9)
    pert = 1e-6
    for i in range(len(hyps)):
        up, dn = hyps.copy(), hyps.copy()
        up[i] += pert
        dn[i] -= pert
        sgp.set_hyperparameters(up)
        lu = sgp.compute_likelihood_gradient_stable()
        sgp.set_hyperparameters(dn)
        ld = sgp.compute_likelihood_gradient_stable()
        fd = (lu - ld) / (2 * pert)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-6)
        assert g[i] == pytest.approx(g_direct[i], rel=1e-6, abs=1e-6)


def test_set_hyperparameters(C, oracle, strucs):
    """ctests/test_sparse_gp.cpp Set_Hyps: a model built at other
    hyperparameters and reset to these gives the same likelihood."""
    a = trained(C, oracle, strucs, sigma=1.5, noise=(1.0, 1.5, 3.0))
    b = trained(C, oracle, strucs, sigma=2.0, noise=(2.0, 2.5, 3.5))
    a.compute_likelihood()
    b.compute_likelihood()
    assert a.log_marginal_likelihood != b.log_marginal_likelihood
    b.set_hyperparameters([1.5, 1.0, 1.5, 3.0])
    b.compute_likelihood()
    assert a.log_marginal_likelihood == pytest.approx(b.log_marginal_likelihood, abs=1e-8)
    np.testing.assert_allclose(a.alpha, b.alpha, rtol=1e-8, atol=1e-12)
    assert b.kernels[0].sigma == 1.5 and b.force_noise == 1.5


def test_json_layout_roundtrip(C, oracle, strucs, tmp_path):
    """SparseGP::to_json / from_json (sparse_gp.cpp:1468-1479): the keys of
    the nlohmann member list (sparse_gp.h:102-108), Eigen matrices as lists
    of rows, empty ones as null; reading restores the state exactly."""
    import json
    sgp = trained(C, oracle, strucs)
    sgp.compute_likelihood_gradient_stable()
    sgp.training_structures = []  # structures need the GPU to rebuild
    path = str(tmp_path / "sgp.json")
    C.SparseGP.to_json(path, sgp)
    j = json.load(open(path))
    assert set(j) == {
        "hyperparameters", "kernels", "Kuu_kernels", "Kuf_kernels", "Kuu", "Kuf", "n_kernels",
        "Kuu_jitter", "Sigma", "Kuu_inverse", "R_inv", "L_inv", "alpha", "R_inv_diag",
        "L_diag", "sparse_descriptors", "training_structures", "sparse_indices",
        "noise_vector", "y", "label_count", "n_energy_labels", "n_force_labels",
        "n_stress_labels", "n_sparse", "n_labels", "n_strucs", "energy_noise", "force_noise",
        "stress_noise", "log_marginal_likelihood", "data_fit", "complexity_penalty",
        "trace_term", "constant_term", "likelihood_gradient"}
    assert j["kernels"][0] == {"sigma": 1.7, "sig2": 1.7 ** 2, "power": 2.0,
                               "kernel_name": "NormalizedDotProduct",
                               "kernel_hyperparameters": [1.7]}
    assert len(j["Kuu"]) == sgp.n_sparse and len(j["Kuu"][0]) == sgp.n_sparse
    cd = j["sparse_descriptors"][0]
    assert set(cd) == {"descriptors", "descriptor_norms", "cutoff_values", "n_clusters_by_type",
                       "cumulative_type_count", "n_descriptors", "n_types", "n_clusters"}
    g = C.SparseGP.from_json(path)
    for name in ("alpha", "L_inv", "Kuu_inverse", "Sigma", "R_inv", "Kuu", "Kuf", "y",
                 "noise_vector", "R_inv_diag", "L_diag", "hyperparameters"):
        np.testing.assert_array_equal(getattr(g, name), getattr(sgp, name), err_msg=name)
    for a, b in zip(g.sparse_descriptors[0].descriptors, sgp.sparse_descriptors[0].descriptors):
        np.testing.assert_array_equal(a, b)
    assert g.n_sparse == sgp.n_sparse and g.n_labels == sgp.n_labels
    assert g.log_marginal_likelihood == sgp.log_marginal_likelihood
    assert g.kernels[0].sigma == 1.7 and g.kernels[0].power == 2.0
    # An empty model writes its empty matrices as null and reads them back.
    e = C.SparseGP([C.NormalizedDotProduct(1.0, 1.0)], 1.0, 1.0, 1.0)
    C.SparseGP.to_json(path, e)
    j = json.load(open(path))
    assert j["Kuu"] is None and j["alpha"] is None
    e2 = C.SparseGP.from_json(path)
    assert e2.Kuu.shape == (0, 0) and e2.alpha.size == 0


def test_kuu_and_update_L_inverse(C, oracle, strucs):
    """Kuu from envs_envs against the kernel's closed form on the oracle's
    descriptors, sigma^2 (x_i . x_j / |x_i| |x_j|)^p within a species and 0
    across species (normalized_dot_product.cpp:28-89), and update_L_inverse
    (sparse_gp.cpp:610-618): L_inv lower triangular with
    L_inv (Kuu + jitter I) L_inv^T = I and Kuu^-1 = L_inv^T L_inv."""
    sgp = C.SparseGP([oracle_kernel(C, oracle, 1.7, 2.0)], 0.1, 0.2, 0.3)
    sgp.add_specific_environments(strucs[0], [0, 1, 3, 4])
    sgp.add_specific_environments(strucs[1], [2, 3, 5])
    sgp.update_L_inverse()
    M = sgp.n_sparse
    cd = sgp.sparse_descriptors[0]
    rows, spec = [], []
    for t in range(len(cd.descriptors)):
        X = np.asarray(cd.descriptors[t])
        rows.extend(X / np.linalg.norm(X, axis=1)[:, None])
        spec.extend([t] * len(X))
    Xh, spec = np.array(rows), np.array(spec)
    want = 1.7 ** 2 * (Xh @ Xh.T) ** 2 * (spec[:, None] == spec[None, :])
    np.testing.assert_allclose(sgp.Kuu, want, rtol=1e-13, atol=1e-14)
    A = sgp.Kuu + sgp.Kuu_jitter * np.eye(M)
    assert np.allclose(np.triu(sgp.L_inv, 1), 0)
    np.testing.assert_allclose(sgp.L_inv @ A @ sgp.L_inv.T, np.eye(M), atol=1e-8)
    np.testing.assert_allclose(sgp.Kuu_inverse, sgp.L_inv.T @ sgp.L_inv, rtol=1e-14)
