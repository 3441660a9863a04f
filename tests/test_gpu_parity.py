"""Parity of the HIP path (through the C ABI, via the _C_flare mirror)
with the FP64 oracle.

Bars: neighbour lists and every integer field bit-exact; FP64 drop-in
within 1e-10 relative (max |diff| / max |ref|) on descriptors, E, F and S;
local variances sigma^2 - |L_inv k|^2 within 1e-10 of sigma^2 (the size of
the cancelling terms); the chosen form (FP32 descriptors, split-BF16x3
kernel contraction, FP64 variance) within the stated error budget.
"""
import sys

import numpy as np
import pytest

from helpers import fcc, model_state, random_cell, rel_err, var_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def C():
    from flare_amd import _C_flare
    return _C_flare


def gpu_structure(C, cell, species, pos, cutoff, calc=None, form=0):
    return C.Structure(cell, species, pos, cutoff, [calc] if calc else [], form=form)


NBR_CASES = [
    ("fcc108", lambda: fcc(4.05, (3, 3, 3), 0.05, 1000), 5.0),
    ("triclinic3", lambda: random_cell(40, 4, nos=3, scale=8.0, skew=1.5), 4.5),
    ("small_cell_sweep2", lambda: random_cell(3, 7, nos=2, scale=3.0, skew=0.2), 5.0),
    ("big_list_fallback", lambda: random_cell(4, 8, nos=1, scale=3.0, skew=0.1), 18.0),
    # ~800 neighbours per atom: longer than the small-LDS list, within the large one.
    ("mid_list", lambda: random_cell(4, 9, nos=1, scale=3.0, skew=0.1), 11.0),
    ("slab", lambda: (np.diag([9.0, 9.0, 25.0]), np.zeros(20, np.int32),
                      np.random.default_rng(2).uniform(0, 1, (20, 3)) * [9, 9, 6]), 5.0),
    ("negative_positions", lambda: (np.diag([7.0, 7.5, 8.0]), np.zeros(30, np.int32),
                                    np.random.default_rng(3).uniform(-30, 30, (30, 3))), 4.0),
]


@pytest.mark.parametrize("name,make,cutoff", NBR_CASES)
def test_neighbors_bit_exact(C, oracle, name, make, cutoff):
    say("case", name)
    cell, species, pos = make()
    ref = oracle.Neighbors(cell, species, pos, cutoff, "literal")
    s = gpu_structure(C, cell, species, pos, cutoff)
    assert s.sweep == ref.sweep
    assert s.n_neighbors == ref.n_neighbors
    assert s.volume == ref.volume
    assert np.array_equal(s.wrapped_positions.view(np.int64), ref.wrapped_positions.view(np.int64))
    for f in ("neighbor_count", "cumulative_neighbor_count", "structure_indices",
              "neighbor_species"):
        np.testing.assert_array_equal(getattr(s, f), getattr(ref, f), err_msg=f)
    assert np.array_equal(s.relative_positions.view(np.int64),
                          ref.relative_positions.view(np.int64))


B2_CASES = [
    ("c1_fcc108", lambda: fcc(4.05, (3, 3, 3), 0.05, 1000), (1, 8, 3), np.full((1, 1), 5.0),
     "quadratic"),
    ("two_species_asym", lambda: random_cell(30, 12, nos=2, scale=7.5, skew=0.5), (2, 5, 4),
     np.array([[4.5, 4.0], [4.2, 3.9]]), "quadratic"),
    ("three_species_cosine", lambda: random_cell(25, 14, nos=3, scale=7.0, skew=0.3), (3, 3, 6),
     np.full((3, 3), 4.8), "cosine"),
    ("lmax10", lambda: random_cell(10, 15, nos=1, scale=6.0, skew=0.1), (1, 4, 10),
     np.full((1, 1), 4.0), "quadratic"),
]


@pytest.mark.parametrize("name,make,settings,cut,cfn", B2_CASES)
@pytest.mark.parametrize("form", [0, 1])
def test_b2_descriptors(C, oracle, name, make, settings, cut, cfn, form):
    say("case", name, "form", form)
    cell, species, pos = make()
    rc = float(cut.max())
    ref = oracle.B2Values(oracle.Neighbors(cell, species, pos, rc, "literal"), *settings, cut,
                          cutoff_function=cfn, with_dervs=False)
    calc = C.B2("chebyshev", cfn, [0.0, rc], [], list(settings), cut)
    s = gpu_structure(C, cell, species, pos, rc, calc, form=form)
    d = s.descriptors[0]
    assert d.n_descriptors == ref.n_descriptors
    assert d.n_clusters_by_type == ref.n_clusters_by_type
    assert d.n_neighbors_by_type == ref.n_neighbors_by_type
    assert d.cumulative_type_count == ref.cumulative_type_count
    for f in ("neighbor_counts", "cumulative_neighbor_counts", "atom_indices", "neighbor_indices"):
        for a, b in zip(getattr(d, f), getattr(ref, f)):
            np.testing.assert_array_equal(a, b, err_msg=f)
    for a, b in zip(d.neighbor_coordinates, ref.neighbor_coordinates):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    tol = TOL64 if form == 0 else 2e-5
    for a, b in zip(d.descriptors, ref.descriptors):
        assert rel_err(a, b) < tol
    for a, b in zip(d.descriptor_norms, ref.descriptor_norms):
        assert rel_err(a, b) < tol


def make_sgp(C, structure_for_sparse, picks, sigma, power, seed, calc):
    """A SparseGP whose sparse set comes from the GPU descriptors of a
    structure, with L_inv from Kuu + jitter and a seeded alpha."""
    kern = C.NormalizedDotProduct(sigma, power)
    sgp = C.SparseGP([kern], 0.01, 0.05, 0.005)
    d = structure_for_sparse.descriptors[0]
    atoms = []
    for s, p in enumerate(picks):
        atoms += [int(d.atom_indices[s][r]) for r in p]
    sgp.add_specific_environments(structure_for_sparse, atoms)
    sgp.update_L_inverse()
    rng = np.random.default_rng(seed)
    sgp.set_alpha(rng.normal(scale=0.1, size=sgp.n_sparse))
    return sgp


def oracle_model(oracle, sgp):
    cd = sgp.sparse_descriptors[0]
    return oracle.Model(sgp.kernels[0].sigma, sgp.kernels[0].power, cd.descriptors,
                        cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv)


PRED_CASES = [
    # 108-atom model: 108-atom FCC, B2 [1, 8, 3], rc 5, NDP sigma 2 p 2, sparse from other frames.
    ("al108", lambda: fcc(4.05, (3, 3, 3), 0.05, 1100), (1, 8, 3), np.full((1, 1), 5.0), 2.0, 2.0),
    ("two_species_asym_p1", lambda: random_cell(30, 12, nos=2, scale=7.5, skew=0.5), (2, 5, 4),
     np.array([[4.5, 4.0], [4.2, 3.9]]), 1.5, 1.0),
    ("three_species_p3", lambda: random_cell(36, 16, nos=3, scale=8.0, skew=0.6), (3, 4, 4),
     np.full((3, 3), 5.0), 1.2, 3.0),
]


@pytest.mark.parametrize("name,make,settings,cut,sigma,power", PRED_CASES)
def test_predict_fp64(C, oracle, name, make, settings, cut, sigma, power):
    say("case", name)
    cell, species, pos = make()
    rc = float(cut.max())
    calc = C.B2("chebyshev", "quadratic", [0.0, rc], [], list(settings), cut)
    # Sparse set from a second seeded structure of the same kind.
    cell2, species2, pos2 = make() if name != "al108" else fcc(4.05, (3, 3, 3), 0.05, 1000)
    if name != "al108":
        pos2 = pos2 + np.random.default_rng(99).normal(scale=0.1, size=pos2.shape)
    sp = gpu_structure(C, cell2, species2, pos2, rc, calc)
    picks = [list(range(0, n, 2)) for n in sp.descriptors[0].n_clusters_by_type]
    sgp = make_sgp(C, sp, picks, sigma, power, seed=4, calc=calc)
    s = gpu_structure(C, cell, species, pos, rc, calc)
    sgp.predict_local_uncertainties(s)
    m = oracle_model(oracle, sgp)
    ref = oracle.B2Values(oracle.Neighbors(cell, species, pos, rc, "literal"), *settings, cut)
    mean, var = m.predict(ref, "local")
    n = len(pos)
    assert abs(s.mean_efs[0] - mean[0]) / abs(mean[0]) < TOL64
    assert rel_err(s.mean_efs[1:1 + 3 * n], mean[1:1 + 3 * n]) < TOL64
    assert rel_err(s.mean_efs[-6:], mean[-6:]) < TOL64
    assert var_err(s.local_uncertainties[0], var, sigma ** 2) < TOL64
    # predict_mean alone gives the same mean.
    s2 = gpu_structure(C, cell, species, pos, rc, calc)
    sgp.predict_mean(s2)
    assert np.array_equal(s2.mean_efs, s.mean_efs)


def test_predict_empty_environment(C, oracle):
    cell = np.eye(3) * 30.0
    pos = np.array([[0, 0, 0], [1, 1, 1], [15, 15, 15]], dtype=float)
    species = np.array([0, 1, 1], np.int32)
    cut = np.full((2, 2), 4.0)
    calc = C.B2("chebyshev", "quadratic", [0.0, 4.0], [], [2, 3, 2], cut)
    s = gpu_structure(C, cell, species, pos, 4.0, calc)
    sgp = make_sgp(C, s, [[0], [0, 1]], 1.5, 2.0, seed=1, calc=calc)
    sgp.predict_local_uncertainties(s)
    m = oracle_model(oracle, sgp)
    ref = oracle.B2Values(oracle.Neighbors(cell, species, pos, 4.0), 2, 3, 2, cut)
    mean, var = m.predict(ref, "local")
    assert rel_err(s.mean_efs, mean) < TOL64
    assert np.all(s.mean_efs[7:10] == 0)
    assert s.local_uncertainties[0][2] == 0.0
    assert var_err(s.local_uncertainties[0], var, 1.5 ** 2) < TOL64


def test_predict_scalable_medium(C, oracle):
    """A 4,000-atom two-species FCC alloy, B2 [2, 12, 6] (d = 2100) at
    rc 6.0 with 400 sparse environments, against the oracle's cell-list /
    adjoint mode."""
    from flare_amd.workloads import FCC_BASIS, lattice
    cell, pos = lattice(FCC_BASIS, 4.0, (10, 10, 10))
    rng = np.random.default_rng(7)
    species = np.zeros(len(pos), np.int32)
    species[rng.permutation(len(pos))[: len(pos) // 2]] = 1
    pos = pos + rng.normal(scale=0.08, size=pos.shape)
    cut = np.full((2, 2), 6.0)
    calc = C.B2("chebyshev", "quadratic", [0.0, 6.0], [], [2, 12, 6], cut)
    s = gpu_structure(C, cell, species, pos, 6.0, calc)
    d = s.descriptors[0]
    picks = [list(rng.choice(n, 200, replace=False)) for n in d.n_clusters_by_type]
    sgp = make_sgp(C, s, picks, 2.0, 2.0, seed=4, calc=calc)
    sgp.predict_local_uncertainties(s)
    m = oracle_model(oracle, sgp)
    nb = oracle.Neighbors(cell, species, pos, 6.0, "cells")
    mean, var_atom, _ = m.predict_scalable(nb, 2, 12, 6, cut)
    n = len(pos)
    assert rel_err(s.mean_efs[1:1 + 3 * n], mean[1:1 + 3 * n]) < TOL64
    assert abs(s.mean_efs[0] - mean[0]) / abs(mean[0]) < TOL64
    assert rel_err(s.mean_efs[-6:], mean[-6:]) < TOL64
    order = np.concatenate(d.atom_indices)
    assert var_err(s.local_uncertainties[0], var_atom[order], 2.0 ** 2) < TOL64
    F = s.mean_efs[1:1 + 3 * n].reshape(n, 3)
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max() * n


# B2 settings of the chosen-form budget test: d = 2100 (adjoint rows staged
# in two passes), d = 3255 (one pass, rows not 16-byte aligned) and d = 3696
# (one pass, aligned rows).
BUDGET_CASES = [[2, 12, 6], [3, 10, 6], [2, 16, 6]]


@pytest.mark.parametrize("settings", BUDGET_CASES)
def test_chosen_form_error_budget(C, oracle, settings):
    """The chosen form (FP32 descriptors, the kernel contraction on int8
    slices with exact integer sums, the gradient contraction split-BF16x3,
    FP64 variance contraction) against the FP64 path on the same model:
    energy per atom and force (bar 1 meV/A), and sigma: |d sigma| / sigma_f
    and the relative error of atoms with sigma >= 0.01 sigma_f (bar 1e-3;
    closer to the sparse set sigma^2 is a cancellation of terms of size
    sigma_f^2 whose relative error is unbounded in any precision)."""
    from flare_amd.workloads import FCC_BASIS, lattice
    cell, pos = lattice(FCC_BASIS, 4.0, (6, 6, 6))
    rng = np.random.default_rng(8)
    nos = settings[0]
    species = rng.integers(0, nos, len(pos)).astype(np.int32)
    pos = pos + rng.normal(scale=0.08, size=pos.shape)
    cut = np.full((nos, nos), 6.0)
    calc = C.B2("chebyshev", "quadratic", [0.0, 6.0], [], settings, cut)
    s64 = gpu_structure(C, cell, species, pos, 6.0, calc, form=0)
    d = s64.descriptors[0]
    picks = [list(rng.choice(n, 150, replace=False)) for n in d.n_clusters_by_type]
    sgp = make_sgp(C, s64, picks, 2.0, 2.0, seed=4, calc=calc)
    sgp.predict_local_uncertainties(s64)
    s32 = gpu_structure(C, cell, species, pos, 6.0, calc, form=1)
    sgp.predict_local_uncertainties(s32)
    n = len(pos)
    dE = abs(s32.mean_efs[0] - s64.mean_efs[0]) / n
    dF = np.abs(s32.mean_efs[1:1 + 3 * n] - s64.mean_efs[1:1 + 3 * n]).max()
    v64 = s64.local_uncertainties[0]
    v32 = s32.local_uncertainties[0]
    sd64 = np.sqrt(np.abs(v64))
    sd32 = np.sqrt(np.abs(v32))
    sig_f = 2.0
    big = sd64 >= 0.01 * sig_f
    sig_abs = np.max(np.abs(sd32 - sd64)) / sig_f
    sig_rel = np.max(np.abs(sd32 - sd64)[big] / sd64[big]) if big.any() else 0.0
    say(f"chosen form: dE {dE * 1e3:.3e} meV/atom, dF {dF * 1e3:.3e} meV/A, "
        f"|dsigma|/sigma_f {sig_abs:.3e}, sigma rel (sigma >= 0.01 sigma_f, "
        f"{big.mean():.0%} of atoms) {sig_rel:.3e}")
    assert dF < 1e-3  # 1 meV/A
    assert dE < 1e-3
    assert sig_abs < 1e-3
    assert sig_rel < 1e-3


@pytest.mark.parametrize("shape", range(7))
def test_split_tile_shapes_bitwise(C, monkeypatch, shape):
    """Each tile shape of the split contraction, pinned for every contraction
    that admits it, gives the autotuned evaluation's results bit for bit
    (every output element sums its k terms in the same order)."""
    from flare_amd.workloads import FCC_BASIS, lattice
    cell, pos = lattice(FCC_BASIS, 4.0, (7, 7, 7))
    rng = np.random.default_rng(12)
    species = rng.integers(0, 2, len(pos)).astype(np.int32)
    pos = pos + rng.normal(scale=0.08, size=pos.shape)
    cut = np.full((2, 2), 5.0)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.0], [], [2, 8, 4], cut)
    s64 = gpu_structure(C, cell, species, pos, 5.0, calc, form=0)
    d = s64.descriptors[0]
    picks = [list(rng.choice(n, 300, replace=False)) for n in d.n_clusters_by_type]
    sgp = make_sgp(C, s64, picks, 2.0, 2.0, seed=4, calc=calc)
    s0 = gpu_structure(C, cell, species, pos, 5.0, calc, form=1)
    sgp.predict_local_uncertainties(s0)
    monkeypatch.setenv("FLARE_MI_TILE_SHAPE", str(shape))
    s1 = gpu_structure(C, cell, species, pos, 5.0, calc, form=1)
    sgp.predict_local_uncertainties(s1)
    assert np.array_equal(s1.mean_efs, s0.mean_efs)
    assert np.array_equal(s1.local_uncertainties[0], s0.local_uncertainties[0])


def test_determinism(C):
    cell, species, pos = random_cell(60, 21, nos=2, scale=9.0, skew=0.4)
    cut = np.full((2, 2), 5.0)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.0], [], [2, 6, 4], cut)
    s = gpu_structure(C, cell, species, pos, 5.0, calc)
    sgp = make_sgp(C, s, [[0, 1, 2], [0, 1, 2]], 1.0, 2.0, seed=3, calc=calc)
    outs = []
    for _ in range(3):
        t = gpu_structure(C, cell, species, pos, 5.0, calc)
        sgp.predict_local_uncertainties(t)
        outs.append((t.mean_efs.copy(), t.local_uncertainties[0].copy()))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])


def test_errors_are_raised(C):
    cut = np.full((1, 1), 4.0)
    calc = C.B2("chebyshev", "quadratic", [0.0, 4.0], [], [1, 3, 2], cut)
    with pytest.raises(ValueError):
        C.Structure(np.eye(3) * 5, [0, 0], np.array([[0, 0, 0], [np.nan, 0, 0]]), 4.0, [calc])
    with pytest.raises(ValueError):
        C.Structure(np.zeros((3, 3)), [0], np.zeros((1, 3)), 4.0, [calc])
    s = C.Structure(np.eye(3) * 6, [0, 0], np.array([[0, 0, 0], [1.5, 0.2, 0.1]]), 4.0, [calc])
    sgp = C.SparseGP([C.NormalizedDotProduct(1.0, 2.0)], 0.1, 0.1, 0.1)
    sgp.add_all_environments(s)
    with pytest.raises(RuntimeError):  # stale L_inv (sparse_gp.cpp:645-649)
        sgp.predict_mean(s)
