"""B3 on the GPU against the oracle (b3.cpp:80-443): descriptors in both
forms, and E/F/stress/local variances of a B3 model (FP64 drop-in bar), a
batch equal to its structures alone, and the chosen form's force error."""
import numpy as np
import pytest

from helpers import model_state, random_cell, rel_err, var_err

pytestmark = pytest.mark.gpu

NOS, NMAX, LMAX = 2, 3, 3
CUT = np.array([[4.6, 4.2], [4.2, 4.4]])
TOL = 1e-10


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def C():
    from flare_amd import _C_flare
    return _C_flare


@pytest.fixture(scope="module")
def system(C, oracle):
    cell, species, pos = random_cell(20, 41, nos=NOS, scale=7.0, skew=0.4)
    calc = C.B3("chebyshev", "quadratic", [0.0, 4.6], [], [NOS, NMAX, LMAX], CUT)
    ref = oracle.B3Values(oracle.Neighbors(cell, species, pos, 4.6, "literal"), NOS, NMAX, LMAX,
                          CUT)
    return dict(cell=cell, species=species, pos=pos, calc=calc, ref=ref)


@pytest.mark.parametrize("form", [0, 1])
def test_b3_descriptors(C, system, form):
    say("B3 descriptors, form", form)
    s = C.Structure(system["cell"], system["species"], system["pos"], 4.6, [system["calc"]],
                    form=form)
    d, ref = s.descriptors[0], system["ref"]
    assert d.n_descriptors == ref.n_descriptors == system["calc"].n_descriptors
    tol = TOL if form == 0 else 2e-5
    for t in range(NOS):
        assert rel_err(d.descriptors[t], ref.descriptors[t]) < tol
        assert rel_err(d.descriptor_norms[t], ref.descriptor_norms[t]) < tol
        assert list(d.atom_indices[t]) == list(ref.atom_indices[t])


def model(C, oracle, system, power):
    ref = system["ref"]
    picks = [list(range(min(4, ref.n_clusters_by_type[t]))) for t in range(NOS)]
    X, N, alpha, Linv = model_state(ref, picks, 1.3, power, seed=11)
    m = oracle.Model(1.3, power, X, N, alpha, L_inv=Linv)
    s = C.Structure(system["cell"], system["species"], system["pos"], 4.6, [system["calc"]])
    sgp = C.SparseGP([C.NormalizedDotProduct(1.3, power)], 0.01, 0.05, 0.005)
    sgp.add_specific_environments(s, [int(ref.atom_indices[t][k]) for t in range(NOS)
                                      for k in picks[t]])
    sgp.set_state(alpha, L_inv=Linv)
    return m, sgp


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_b3_predict_fp64(C, oracle, system, power):
    say("B3 predict, power", power)
    m, sgp = model(C, oracle, system, power)
    s = C.Structure(system["cell"], system["species"], system["pos"], 4.6, [system["calc"]])
    sgp.predict_local_uncertainties(s)
    mean, var = m.predict(system["ref"], "local")
    n = s.noa
    assert abs(s.mean_efs[0] - mean[0]) / abs(mean[0]) < TOL
    assert rel_err(s.mean_efs[1:1 + 3 * n], mean[1:1 + 3 * n]) < TOL
    assert rel_err(s.mean_efs[-6:], mean[-6:]) < TOL
    assert var_err(s.local_uncertainties[0], var, 1.3 ** 2) < TOL


def test_b3_batch_and_chosen_form(C, oracle, system):
    say("B3 batch and chosen form")
    from flare_amd.batch import Batch
    m, sgp = model(C, oracle, system, 2.0)
    structs = [(system["cell"], system["species"], system["pos"])]
    cell2, sp2, pos2 = random_cell(14, 43, nos=NOS, scale=6.5, skew=0.3)
    structs.append((cell2, sp2, pos2))
    for form in (0, 1):
        means, vars_ = Batch(structs, 4.6, system["calc"], form=form).predict(sgp)
        for k, st in enumerate(structs):
            s = C.Structure(*st, 4.6, [system["calc"]], form=form)
            sgp.predict_local_uncertainties(s)
            assert np.array_equal(means[k], s.mean_efs), (form, k)
            assert np.array_equal(vars_[k][0], s.local_uncertainties[0]), (form, k)
    s64 = C.Structure(*structs[0], 4.6, [system["calc"]])
    s32 = C.Structure(*structs[0], 4.6, [system["calc"]], form=1)
    sgp.predict_mean(s64)
    sgp.predict_mean(s32)
    n = s64.noa
    dF = np.abs(s32.mean_efs[1:1 + 3 * n] - s64.mean_efs[1:1 + 3 * n]).max()
    say(f"B3 chosen form: dF {dF * 1e3:.3e} meV/A")
    assert dF < 1e-3


def test_b3_limits(C, system):
    with pytest.raises(ValueError, match="l_max"):
        C.Structure(system["cell"], system["species"], system["pos"], 4.6,
                    [C.B3("chebyshev", "quadratic", [0.0, 4.6], [], [NOS, 2, 5], CUT)])
    s = C.Structure(system["cell"], system["species"], system["pos"], 4.6, [system["calc"]])
    with pytest.raises(ValueError, match="B2"):
        s.descriptors[0].descriptor_force_dervs


@pytest.mark.parametrize("form", [0, 1])
def test_b3_row_tiles(C, oracle, system, form):
    """Prediction over B3 row tiles (descriptors formed tile by tile, never
    held whole) equals the untiled prediction bit for bit, for tiles that
    split species and atoms unevenly."""
    say("B3 row tiles, form", form)
    from flare_amd import _lib
    m, sgp = model(C, oracle, system, 2.0)
    args = (system["cell"], system["species"], system["pos"], 4.6, [system["calc"]])
    whole = C.Structure(*args, form=form)
    sgp.predict_local_uncertainties(whole)
    for rows in (7, 16):
        ctx = _lib.Context(0)
        ctx.set_row_tile(rows)
        s = C.Structure(*args, form=form, context=ctx)
        sgp.predict_local_uncertainties(s)
        assert np.array_equal(s.mean_efs, whole.mean_efs), rows
        assert np.array_equal(s.local_uncertainties[0], whole.local_uncertainties[0]), rows
        with pytest.raises(RuntimeError, match="row tiles"):
            s.descriptors[0].descriptors


@pytest.mark.parametrize("nos,nmax,n_atoms", [(3, 6, 14), (4, 8, 9)])
def test_b3_lmax4_settings(C, oracle, nos, nmax, n_atoms):
    """l_max = 4 B3 against the oracle: [3, 6, 4] is the setting of the
    million-atom alloy (18 radial channels, pair-factorized kernels); [4, 8, 4]
    has 32 channels, 528 radial pairs, more than one workgroup takes, so both
    forms run the wave-per-atom kernels (one species has no atoms)."""
    say("B3 l_max 4 settings", nos, nmax, n_atoms)
    rc = 4.4
    cut = np.full((nos, nos), rc)
    cell, species, pos = random_cell(n_atoms, 7 + nos, nos=nos, scale=7.0, skew=0.3)
    calc = C.B3("chebyshev", "quadratic", [0.0, rc], [], [nos, nmax, 4], cut)
    ref = oracle.B3Values(oracle.Neighbors(cell, species, pos, rc, "literal"), nos, nmax, 4, cut)
    picks = [list(range(min(3, ref.n_clusters_by_type[t]))) for t in range(nos)]
    X, N, alpha, Linv = model_state(ref, picks, 1.1, 2.0, seed=5)
    m = oracle.Model(1.1, 2.0, X, N, alpha, L_inv=Linv)
    s0 = C.Structure(cell, species, pos, rc, [calc])
    sgp = C.SparseGP([C.NormalizedDotProduct(1.1, 2.0)], 0.01, 0.05, 0.005)
    sgp.add_specific_environments(s0, [int(ref.atom_indices[t][k]) for t in range(nos)
                                       for k in picks[t]])
    sgp.set_state(alpha, L_inv=Linv)
    for t in range(nos):
        assert rel_err(s0.descriptors[0].descriptors[t], ref.descriptors[t]) < TOL
    s = C.Structure(cell, species, pos, rc, [calc])
    sgp.predict_local_uncertainties(s)
    mean, var = m.predict(ref, "local")
    n = s.noa
    assert abs(s.mean_efs[0] - mean[0]) / abs(mean[0]) < TOL
    assert rel_err(s.mean_efs[1:1 + 3 * n], mean[1:1 + 3 * n]) < TOL
    assert rel_err(s.mean_efs[-6:], mean[-6:]) < TOL
    assert var_err(s.local_uncertainties[0], var, 1.1 ** 2) < TOL
    s32 = C.Structure(cell, species, pos, rc, [calc], form=1)
    sgp.predict_mean(s32)
    dF = np.abs(s32.mean_efs[1:1 + 3 * n] - s.mean_efs[1:1 + 3 * n]).max()
    say(f"  chosen form dF {dF * 1e3:.3e} meV/A")
    assert dF < 1e-3


@pytest.mark.parametrize("form", [0, 1])
def test_b2_and_b3_kernels_in_one_model(C, oracle, system, form):
    """Two descriptor calculators on one structure and a two-kernel model
    (kernel i acts on descriptor i, sparse_gp.cpp:643-757): the means add and
    the local variances are per kernel, equal bit for bit to the two
    single-kernel models, and (FP64) to the oracle's sum at the drop-in bar."""
    say("B2 + B3 model, form", form)
    from scipy.linalg import block_diag
    cell, species, pos = system["cell"], system["species"], system["pos"]
    b2 = C.B2("chebyshev", "quadratic", [0.0, 4.6], [], [NOS, 4, 3], CUT)
    b3 = system["calc"]
    ref3 = system["ref"]
    ref2 = oracle.B2Values(oracle.Neighbors(cell, species, pos, 4.6, "literal"), NOS, 4, 3, CUT)
    picks = [list(range(min(4, ref3.n_clusters_by_type[t]))) for t in range(NOS)]
    X2, N2, a2, L2 = model_state(ref2, picks, 1.7, 2.0, seed=21)
    X3, N3, a3, L3 = model_state(ref3, picks, 1.3, 2.0, seed=11)
    s0 = C.Structure(cell, species, pos, 4.6, [b2, b3])
    atoms = [int(ref3.atom_indices[t][k]) for t in range(NOS) for k in picks[t]]
    kern = [C.NormalizedDotProduct(1.7, 2.0), C.NormalizedDotProduct(1.3, 2.0)]
    both = C.SparseGP(kern, 0.01, 0.05, 0.005)
    both.add_specific_environments(s0, atoms)
    both.set_state(np.concatenate([a2, a3]), L_inv=block_diag(L2, L3))
    singles = []
    for k, (calc, a, L) in enumerate(((b2, a2, L2), (b3, a3, L3))):
        g = C.SparseGP([kern[k]], 0.01, 0.05, 0.005)
        g.add_specific_environments(C.Structure(cell, species, pos, 4.6, [calc]), atoms)
        g.set_state(a, L_inv=L)
        singles.append((calc, g))
    s = C.Structure(cell, species, pos, 4.6, [b2, b3], form=form)
    both.predict_local_uncertainties(s)
    parts = []
    for calc, g in singles:
        t = C.Structure(cell, species, pos, 4.6, [calc], form=form)
        g.predict_local_uncertainties(t)
        parts.append(t)
    assert np.array_equal(s.mean_efs, parts[0].mean_efs + parts[1].mean_efs)
    assert len(s.local_uncertainties) == 2
    for k in range(2):
        assert np.array_equal(s.local_uncertainties[k], parts[k].local_uncertainties[0])
    if form == 0:
        n = s.noa
        m2, v2 = oracle.Model(1.7, 2.0, X2, N2, a2, L_inv=L2).predict(ref2, "local")
        m3, v3 = oracle.Model(1.3, 2.0, X3, N3, a3, L_inv=L3).predict(ref3, "local")
        mean = m2 + m3
        assert abs(s.mean_efs[0] - mean[0]) / abs(mean[0]) < TOL
        assert rel_err(s.mean_efs[1:1 + 3 * n], mean[1:1 + 3 * n]) < TOL
        assert rel_err(s.mean_efs[-6:], mean[-6:]) < TOL
        assert var_err(s.local_uncertainties[0], v2, 1.7 ** 2) < TOL
        assert var_err(s.local_uncertainties[1], v3, 1.3 ** 2) < TOL
