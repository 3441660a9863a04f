"""The sparse-GP training state built on the GPU path: Kuf columns from
envs_struc on the GPU, Kuu, and the QR producer of alpha / Sigma / R_inv
(sparse_gp.cpp:343-641), for the 108-atom FCC model trained on 20
Lennard-Jones frames with 10 sparse atoms each (200 sparse environments;
B2 [1, 8, 3], rc 5.0; NormalizedDotProduct sigma 2, power 2; noise 0.01 /
0.05 / 0.005). Checked against the oracle's kernel matrices and
predictions at the FP64 bar (1e-10), and through the normal-equation closed
form of alpha. Also: the reference's AddOrder and SortTest recipes
(ctests/test_sparse_gp.cpp) with the GPU kernel matrices, the SparseGP
file writers with the calculator of the first training structure, and DTC
/ SOR with a trained R_inv and with two kernels."""
import os

import numpy as np
import pytest

from helpers import rel_err, var_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def C():
    from flare_amd import _C_flare
    return _C_flare


def al108_calc(C):
    return C.B2("chebyshev", "quadratic", [0.0, 5.0], [], [1, 8, 3])


@pytest.fixture(scope="module")
def al108(C):
    """20 labelled frames, 10 sparse atoms each (seed 7)."""
    from flare_amd.workloads import lj_training_frames
    say("building the 108-atom LJ-trained model")
    calc = al108_calc(C)
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    rng = np.random.default_rng(7)
    frames, picks = lj_training_frames(20), []
    for cell, species, pos, E, F, S in frames:
        s = C.Structure(cell, species, pos, 5.0, [calc])
        s.energy = np.array([[E]])
        s.forces = F.reshape(-1)
        s.stresses = S
        sgp.add_training_structure(s)
        p = sorted(int(a) for a in rng.choice(len(pos), 10, replace=False))
        picks.append(p)
        sgp.add_specific_environments(s, p)
    sgp.update_matrices_QR()
    return dict(sgp=sgp, calc=calc, frames=frames, picks=picks)


@pytest.fixture(scope="module")
def al108_oracle(oracle, al108):
    """The oracle's sparse set and kernel matrices for the same frames."""
    X, N, descs = [], [], []
    for (cell, species, pos, *_), p in zip(al108["frames"], al108["picks"]):
        d = oracle.B2Values(oracle.Neighbors(cell, species, pos, 5.0, "literal"), 1, 8, 3,
                            np.full((1, 1), 5.0))
        descs.append(d)
        order = np.argsort(d.atom_indices[0])  # row of each atom (one species: identity)
        X.append(d.descriptors[0][order[p]])
        N.append(d.descriptor_norms[0][order[p]])
    X, N = np.concatenate(X), np.concatenate(N)
    return dict(X=X, N=N, descs=descs)


def test_sparse_set_and_kuu(C, oracle, al108, al108_oracle):
    sgp = al108["sgp"]
    assert sgp.n_sparse == 200 and sgp.n_labels == 20 * (1 + 3 * 108 + 6)
    cd = sgp.sparse_descriptors[0]
    assert rel_err(cd.descriptors[0], al108_oracle["X"]) < TOL64
    from helpers import ndp_kuu
    Kuu = ndp_kuu(al108_oracle["X"], al108_oracle["N"], 2.0, 2.0)
    assert rel_err(sgp.Kuu, Kuu) < TOL64


def test_kuf_matches_oracle(C, oracle, al108, al108_oracle):
    """Kuf columns: envs_struc (normalized_dot_product.cpp:117-214) of the
    final sparse set against every training frame."""
    sgp = al108["sgp"]
    m = oracle.Model(2.0, 2.0, [al108_oracle["X"]], [al108_oracle["N"]], np.zeros(200))
    ref = np.concatenate([m.envs_struc(d) for d in al108_oracle["descs"]], axis=1)
    assert ref.shape == sgp.Kuf.shape
    err = rel_err(sgp.Kuf, ref)
    say(f"Kuf rel err {err:.2e}")
    assert err < TOL64


def test_alpha_normal_equations(C, al108):
    """update_matrices_QR solves the least-squares problem
    min |Lambda^1/2 (y - Kfu alpha)|^2 + alpha^T (Kuu + jitter I) alpha, i.e.
    the normal equations (Kuu + jitter I + Kuf Lambda Kfu) alpha = Kuf Lambda y
    (sparse_gp.cpp:607-641). With noise weights up to 4e4 the normal matrix
    is too ill-conditioned to solve directly, so the check is the backward
    error of the optimality condition A^T (b - A alpha) = 0 for the stacked
    A = [Lambda^1/2 Kfu; L^T], b = [Lambda^1/2 y; 0], plus agreement with an
    SVD least-squares solve within its conditioning."""
    sgp = al108["sgp"]
    lam = np.sqrt(sgp.noise_vector)
    L = np.linalg.cholesky(sgp.Kuu + sgp.Kuu_jitter * np.eye(200))
    A = np.vstack([lam[:, None] * sgp.Kuf.T, L.T])
    b = np.r_[lam * sgp.y, np.zeros(200)]
    r = b - A @ sgp.alpha
    nA = np.linalg.norm(A, 2)
    back = np.linalg.norm(A.T @ r) / (nA * (nA * np.linalg.norm(sgp.alpha) + np.linalg.norm(r)))
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    cond = np.linalg.cond(A)
    say(f"backward error {back:.2e}, cond {cond:.2e}, |alpha - lstsq| / |alpha| "
        f"{np.linalg.norm(sgp.alpha - x) / np.linalg.norm(x):.2e}")
    assert back < 1e-13
    assert np.linalg.norm(sgp.alpha - x) / np.linalg.norm(x) < 1e-15 * cond * 100
    # The trained model reproduces its LJ labels to within a few noise levels.
    fit = sgp.Kuf.T @ sgp.alpha
    f = slice(1, 1 + 3 * 108)
    rmse_f = np.sqrt(np.mean((fit[f] - sgp.y[f]) ** 2))
    say(f"training force rmse {rmse_f * 1e3:.2f} meV/A")
    assert rmse_f < 0.1


def test_predictions_match_oracle(C, oracle, al108):
    """Mean, local variances, DTC and SOR of the trained model on a fresh
    frame (seed 1100) against the oracle at the FP64 bar."""
    from flare_amd.workloads import fcc_al_108
    sgp, calc = al108["sgp"], al108["calc"]
    cell, species, pos = fcc_al_108(frame=1100)
    s = C.Structure(cell, species, pos, 5.0, [calc])
    cd = sgp.sparse_descriptors[0]
    m = oracle.Model(2.0, 2.0, cd.descriptors, cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv,
                     Kuu_inverse=sgp.Kuu_inverse, Sigma=sgp.Sigma, R_inv=sgp.R_inv)
    d = oracle.B2Values(oracle.Neighbors(cell, species, pos, 5.0, "literal"), 1, 8, 3,
                        np.full((1, 1), 5.0))
    sgp.predict_local_uncertainties(s)
    mean, var = m.predict(d, "local")
    assert rel_err(s.mean_efs, mean) < TOL64
    assert var_err(s.local_uncertainties[0], var, 4.0) < TOL64
    # DTC / SOR of a trained model: the kernel matrix and self kernel match
    # the oracle at the FP64 bar; the variances are differences of terms up
    # to 1e8 times larger (K^T Kuu^-1 K with jitter-level eigenvalues), so
    # they are compared relative to the size of those terms, each evaluated
    # through the triangular factors (the reference's explicit Kuu^-1 loses
    # more digits than the variance has).
    K = m.envs_struc(d)
    kself = m.self_kernel(d)
    kern = sgp.kernels[0]
    assert rel_err(kern.envs_struc(cd, s.descriptors[0]), K) < TOL64
    assert rel_err(kern.self_kernel_struc(s.descriptors[0]), kself) < TOL64
    q1 = np.sum((sgp.L_inv @ K) ** 2, axis=0)
    q2 = np.sum((sgp.R_inv.T @ K) ** 2, axis=0)
    scale = kself + q1 + q2
    sgp.predict_DTC(s)
    err = np.max(np.abs(s.variance_efs - (kself - q1 + q2)) / scale)
    say(f"DTC error relative to its terms {err:.2e}")
    assert err < TOL64
    # SOR alone is |R_inv^T K|^2, whose entries cancel: |R_inv|^T |K| exceeds
    # |R_inv^T K| by kappa (~3e10 here), so rounding K or the products at
    # 1e-16 can move it by up to ~kappa 1e-16 of its size. The bar, 1e-7 of
    # the largest variance, sits between that and the O(1) of a wrong term
    # (well-conditioned R_inv is held to 1e-10 in the dense tests).
    sgp.predict_SOR(s)
    kappa = np.max(np.linalg.norm(np.abs(sgp.R_inv.T) @ np.abs(K), axis=0)
                   / np.sqrt(np.maximum(q2, 1e-300)))
    err = np.max(np.abs(s.variance_efs - q2)) / q2.max()
    say(f"SOR error relative to its largest value {err:.2e} (kappa {kappa:.1e})")
    assert err < 1e-7
    from flare_amd.workloads import lennard_jones
    E, F, S = lennard_jones(cell, pos)
    err_f = np.abs(s.mean_efs[1:-6] - F.reshape(-1)).max()
    say(f"held-out frame: |dF|max {err_f * 1e3:.1f} meV/A against LJ")


def test_add_order_gpu(C):
    """ctests/test_sparse_gp.cpp AddOrder with the GPU kernel matrices."""
    from flare_amd.workloads import lj_training_frames
    calc = al108_calc(C)
    (cell, species, pos, E, F, S), = lj_training_frames(1)
    s = C.Structure(cell, species, pos, 5.0, [calc])
    s.forces = F.reshape(-1)
    a = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 1.0, 2.0, 3.0)
    b = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 1.0, 2.0, 3.0)
    a.add_training_structure(s)
    a.add_all_environments(s)
    a.update_matrices_QR()
    b.add_all_environments(s)
    b.add_training_structure(s)
    b.update_matrices_QR()
    np.testing.assert_array_equal(a.Kuf, b.Kuf)
    np.testing.assert_array_equal(a.Kuu, b.Kuu)


def test_uncertain_environments_sorted(C, al108):
    """ctests/test_sparse_gp.cpp SortTest / add_uncertain_environments: the
    added clusters are those of largest local variance."""
    from flare_amd.workloads import fcc_al_108
    sgp, calc = al108["sgp"], al108["calc"]
    cell, species, pos = fcc_al_108(frame=1200)
    s = C.Structure(cell, species, pos, 5.0, [calc])
    var = sgp.compute_cluster_uncertainties(s)[0]
    order = np.argsort(-var, kind="stable")
    g = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    for st, p in zip(sgp.training_structures[:2], [[0, 5], [1, 7]]):
        g.add_training_structure(st)
        g.add_specific_environments(st, p)
    g.update_matrices_QR()
    var_g = g.compute_cluster_uncertainties(s)[0]
    top = np.argsort(-var_g, kind="stable")[:5]
    g.add_uncertain_environments(s, [5])
    assert g.n_sparse == 9
    assert g.sparse_indices[0][-1] == [int(s.descriptors[0].atom_indices[0][i]) for i in top]
    assert np.all(np.diff(var[order]) <= 0)


def test_writers_use_training_calculator(C, al108, tmp_path):
    """SparseGP.write_* (sparse_gp.cpp:1200-1466) take the descriptor
    calculator of the first training structure; the coefficient file they
    write reproduces the model's energies and forces."""
    from flare_amd import mapping
    from flare_amd.workloads import fcc_al_108
    sgp, calc = al108["sgp"], al108["calc"]
    p1, p2 = os.path.join(tmp_path, "a.flare"), os.path.join(tmp_path, "b.flare")
    sgp.write_mapping_coefficients(p1, "tester", 0)
    mapping.write_mapping_coefficients(sgp, p2, "tester", 0, calc)
    a, b = open(p1).read().split("\n"), open(p2).read().split("\n")
    assert a[1:] == b[1:]  # everything after the date line
    for name in ("write_L_inverse", "write_sparse_descriptors"):
        getattr(sgp, name)(p1, "tester")
        assert os.path.getsize(p1) > 0
    cell, species, pos = fcc_al_108(frame=1100)
    s = C.Structure(cell, species, pos, 5.0, [calc])
    sgp.predict_mean(s)
    sgp.write_mapping_coefficients(p1, "tester", 0)
    pot = mapping.MappedPotential(p1)
    e, f, v = pot.compute(cell, species, pos)
    assert np.sum(e) == pytest.approx(s.mean_efs[0], rel=1e-10)
    assert rel_err(f.reshape(-1), s.mean_efs[1:-6]) < 1e-10


def test_dtc_two_kernels(C, oracle):
    """DTC / SOR with two kernels (sparse_gp.cpp:667-729: K stacked over
    kernels, self kernels summed) against the oracle's per-kernel matrices."""
    from helpers import random_cell
    cell, species, pos = random_cell(20, 41, nos=2, scale=7.0, skew=0.3)
    cuts = [np.array([[4.2, 3.8], [3.8, 4.0]]), np.full((2, 2), 3.6)]
    calcs = [C.B2("chebyshev", "quadratic", [0.0, 4.2], [], [2, 4, 3], cuts[0]),
             C.B2("chebyshev", "quadratic", [0.0, 3.6], [], [2, 3, 2], cuts[1])]
    s = C.Structure(cell, species, pos, 4.2, calcs)
    rng = np.random.default_rng(3)
    s.energy = rng.normal(size=1)
    s.forces = rng.normal(size=60)
    s.stresses = rng.normal(size=6)
    sgp = C.SparseGP([C.NormalizedDotProduct(1.5, 2.0), C.NormalizedDotProduct(0.8, 1.0)],
                     0.1, 0.2, 0.3)
    sgp.add_training_structure(s)
    sgp.add_specific_environments(s, [0, 2, 4, 7, 9, 11])
    sgp.update_matrices_QR()
    t_cell, t_species, t_pos = random_cell(16, 42, nos=2, scale=6.5, skew=0.3)
    t = C.Structure(t_cell, t_species, t_pos, 4.2, calcs)
    nb = oracle.Neighbors(t_cell, t_species, t_pos, 4.2, "literal")
    Ks, kself, mean = [], 0, 0
    o = 0
    for k, (calc, cut, kern) in enumerate(zip(calcs, cuts, sgp.kernels)):
        d = oracle.B2Values(nb, 2, *calc.descriptor_settings[1:], cut)
        cd = sgp.sparse_descriptors[k]
        m = oracle.Model(kern.sigma, kern.power, cd.descriptors, cd.descriptor_norms,
                         sgp.alpha[o:o + cd.n_clusters])
        Ks.append(m.envs_struc(d))
        kself = kself + m.self_kernel(d)
        o += cd.n_clusters
    K = np.concatenate(Ks)
    mean = K.T @ sgp.alpha
    dtc = kself - np.einsum("ic,ij,jc->c", K, sgp.Kuu_inverse, K) + np.einsum(
        "ic,ij,jc->c", K, sgp.Sigma, K)
    V = K.T @ sgp.R_inv
    sgp.predict_DTC(t)
    assert rel_err(t.mean_efs, mean) < TOL64
    assert rel_err(t.variance_efs, dtc) < 1e-9
    sgp.predict_SOR(t)
    assert rel_err(t.variance_efs, np.sum(V * V, axis=1)) < TOL64


def test_json_roundtrip_trained(C, al108, tmp_path):
    """A trained model written with SparseGP.to_json and read back predicts
    bit for bit the same, keeps its kernel matrices, and keeps training: one
    more frame and sparse set gives the same state in both."""
    from flare_amd.workloads import lj_training_frames, fcc_al_108
    sgp, calc = al108["sgp"], al108["calc"]
    path = str(tmp_path / "al108.json")
    C.SparseGP.to_json(path, sgp)
    g = C.SparseGP.from_json(path)
    assert g.n_strucs == 20 and g.n_sparse == 200
    np.testing.assert_array_equal(g.Kuf, sgp.Kuf)
    cell, species, pos = fcc_al_108(frame=1100)
    a = C.Structure(cell, species, pos, 5.0, [calc])
    b = C.Structure(cell, species, pos, 5.0, [calc])
    sgp.predict_local_uncertainties(a)
    g.predict_local_uncertainties(b)
    np.testing.assert_array_equal(a.mean_efs, b.mean_efs)
    np.testing.assert_array_equal(a.local_uncertainties[0], b.local_uncertainties[0])
    # Continue training both.
    (cell, species, pos, E, F, S), = lj_training_frames(21)[20:]
    for model in (sgp, g):
        s = C.Structure(cell, species, pos, 5.0, [model.training_structures[0]
                                                   .descriptor_calculators[0]])
        s.energy, s.forces, s.stresses = np.array([E]), F.reshape(-1), S
        model.add_training_structure(s)
        model.add_specific_environments(s, [3, 50, 99])
        model.update_matrices_QR()
    assert rel_err(g.Kuf, sgp.Kuf) < 1e-14
    assert rel_err(g.alpha, sgp.alpha) < 1e-9


class Atoms:
    """The ASE Atoms fields the calculator reads."""

    def __init__(self, numbers, cell, positions):
        self.numbers, self.cell, self.positions = np.asarray(numbers), cell, positions


@pytest.fixture(scope="module")
def wrapper(C):
    """An SGP_Wrapper trained with update_db on three LJ frames (atomic
    number 13 coded as species 0)."""
    from flare_amd.sgp_wrapper import SGP_Wrapper
    from flare_amd.workloads import lj_training_frames
    gp = SGP_Wrapper([C.NormalizedDotProduct(2.0, 2.0)], [al108_calc(C)], 5.0, 0.01, 0.05, 0.005,
                     {13: 0}, variance_type="local", single_atom_energies={0: -1.5})
    for k, (cell, species, pos, E, F, S) in enumerate(lj_training_frames(3)):
        gp.update_db(Atoms([13] * len(pos), cell, pos), F, custom_range=[0, 9 + k, 40, 77],
                     energy=E + len(pos) * -1.5, stress=S, mode="specific", update_qr=False,
                     rel_e_noise=1.0, rel_f_noise=0.5 + k, rel_s_noise=1.0)
    gp.sparse_gp.update_matrices_QR()
    return gp


def test_wrapper_model_file(C, wrapper, tmp_path):
    """write_model / from_file (sparse_gp.py:120-320) rebuild the same model
    by replaying the training."""
    from flare_amd.sgp_wrapper import SGP_Wrapper
    from flare_amd.workloads import fcc_al_108
    path = str(tmp_path / "gp.json")
    wrapper.write_model(path)
    gp2, kernels = SGP_Wrapper.from_file(path)
    np.testing.assert_array_equal(gp2.sparse_gp.Kuf, wrapper.sparse_gp.Kuf)
    np.testing.assert_array_equal(gp2.sparse_gp.noise_vector, wrapper.sparse_gp.noise_vector)
    assert rel_err(gp2.sparse_gp.alpha, wrapper.sparse_gp.alpha) < 1e-12
    assert gp2.single_atom_energies == {0: -1.5} and gp2.species_map == {13: 0}
    from flare_amd.calculator import SGP_Calculator
    cell, species, pos = fcc_al_108(frame=1100)
    r1 = SGP_Calculator(wrapper).calculate(Atoms([13] * 108, cell, pos))
    r2 = SGP_Calculator(gp2).calculate(Atoms([13] * 108, cell, pos))
    for k in r1:
        assert rel_err(np.atleast_1d(r2[k]), np.atleast_1d(r1[k])) < 1e-12, k


def test_calculator_stress_and_energy(C, wrapper):
    """ASE stress = (1/V) dE/d(strain) (Voigt xx, yy, zz, yz, xz, xy) against
    central differences of the calculator's own energy; energy includes the
    single-atom energies; forces = -dE/dx."""
    from flare_amd.calculator import SGP_Calculator
    from flare_amd.workloads import fcc_al_108
    calc = SGP_Calculator(wrapper)
    cell, species, pos = fcc_al_108(frame=1100)
    r = dict(calc.calculate(Atoms([13] * 108, cell, pos)))
    s = C.Structure(cell, [0] * 108, pos, 5.0, wrapper.descriptor_calculators)
    wrapper.sparse_gp.predict_mean(s)
    assert r["energy"] == pytest.approx(s.mean_efs[0] + 108 * -1.5, rel=1e-14)
    V = abs(np.linalg.det(cell))
    h = 1e-5
    voigt = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (1, 2): 3, (0, 2): 4, (0, 1): 5}
    # The energy (about -500) is a sum of terms alpha_j k_ij of up to ~1e7 in
    # all (|alpha| reaches ~4e3 in this model), so its rounding is ~1e-9 and
    # central differences carry that over 2h: ~1e-4 at h = 1e-5. The strain
    # step stays that small because the quadratic cutoff is C1 and a strain
    # moves every pair (a larger step would straddle pairs at rc); the bar is
    # 1e-4 of the largest component, far below what a wrong sign or
    # component order would give.
    scale = np.abs(r["stress"]).max() * V
    for (a, b), q in voigt.items():
        es = []
        for t in (h, -h):
            c, p = cell.copy(), pos.copy()
            c[:, a] += c[:, b] * t
            p[:, a] += p[:, b] * t
            es.append(calc.calculate(Atoms([13] * 108, c, p))["energy"])
        assert abs(r["stress"][q] * V - (es[0] - es[1]) / (2 * h)) < 1e-4 * scale
    # Forces: one atom moves, so a step of 3e-4 keeps clear of the kink (the
    # nearest pair of atom 0 is 1.9e-3 from rc) while the rounding noise
    # falls to ~1e-6 and the h^2 term to ~1e-7.
    h = 3e-4
    for a, c_ in ((0, 0), (50, 2)):
        es = []
        for t in (h, -h):
            p = pos.copy()
            p[a, c_] += t
            es.append(calc.calculate(Atoms([13] * 108, cell, p))["energy"])
        assert r["forces"][a, c_] == pytest.approx(-(es[0] - es[1]) / (2 * h), abs=1e-5)
    assert r["stds"].shape == (108, 3) and np.all(r["stds"][:, 1:] == 0)


def test_build_map_std_identity(C, wrapper, tmp_path, monkeypatch):
    """build_map (calculator.py:169-185) writes the coefficient file and the
    L_inv / sparse-descriptor files (their names prefix the given one, so it
    runs in the output folder); compute flare/std/atom on those files gives
    the calculator's local stds (tests/test_lammps.py:148-260)."""
    from flare_amd.calculator import SGP_Calculator
    from flare_amd.mapping import StdAtom
    from flare_amd.workloads import fcc_al_108
    calc = SGP_Calculator(wrapper)
    monkeypatch.chdir(tmp_path)
    base = str(tmp_path / "lmp.flare")
    calc.build_map("lmp.flare", "tester")
    for f in (base, str(tmp_path / "L_inv_lmp.flare"), str(tmp_path / "sparse_desc_lmp.flare")):
        assert os.path.getsize(f) > 0
    cell, species, pos = fcc_al_108(frame=1100)
    r = calc.calculate(Atoms([13] * 108, cell, pos))
    std = StdAtom(str(tmp_path / "L_inv_lmp.flare"), str(tmp_path / "sparse_desc_lmp.flare"))
    sig = std.compute(cell, [0] * 108, pos)
    assert np.max(np.abs(sig - r["stds"][:, 0])) < 1e-9
