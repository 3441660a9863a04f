"""Batched evaluation: structures of uneven sizes and cells in one device
object give, for each structure, the same result as evaluating it alone
(bit for bit, both forms), and agree with the oracle."""
import numpy as np
import pytest

from helpers import rel_err, var_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    from flare_amd import _C_flare as C
    from flare_amd.workloads import batch_structures
    structs = batch_structures(n_structures=14, seed=17, n_min=16, n_max=300, nos=3)
    cut = np.full((3, 3), 5.5)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 4, 3], cut)
    first = C.Structure(*structs[0], 5.5, [calc])
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    sgp.add_all_environments(first)
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(6).normal(scale=0.1, size=sgp.n_sparse))
    return C, structs, calc, sgp


@pytest.mark.parametrize("form", [0, 1])
def test_batch_equals_single(setup, form):
    from flare_amd.batch import Batch
    C, structs, calc, sgp = setup
    b = Batch(structs, 5.5, calc, form=form)
    means, vars_ = b.predict(sgp)
    for k, (cell, species, pos) in enumerate(structs):
        s = C.Structure(cell, species, pos, 5.5, [calc], form=form)
        sgp.predict_local_uncertainties(s)
        assert np.array_equal(means[k], s.mean_efs), k
        assert np.array_equal(vars_[k][0], s.local_uncertainties[0]), k


def test_batch_matches_oracle(setup, oracle):
    from flare_amd.batch import Batch
    C, structs, calc, sgp = setup
    b = Batch(structs[:5], 5.5, calc)
    means, vars_ = b.predict(sgp)
    cd = sgp.sparse_descriptors[0]
    m = oracle.Model(2.0, 2.0, cd.descriptors, cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv)
    for k, (cell, species, pos) in enumerate(structs[:5]):
        ref = oracle.B2Values(oracle.Neighbors(cell, species, pos, 5.5, "cells"), 3, 4, 3,
                              np.full((3, 3), 5.5), with_dervs=False)
        nb = oracle.Neighbors(cell, species, pos, 5.5, "cells")
        mean, var_atom, _ = m.predict_scalable(nb, 3, 4, 3, np.full((3, 3), 5.5))
        order = np.concatenate(ref.atom_indices)
        assert rel_err(means[k], mean) < 1e-10
        assert var_err(vars_[k][0], var_atom[order], 2.0 ** 2) < 1e-10


@pytest.mark.parametrize("form", [0, 1])
def test_two_kernel_batch(setup, form):
    """Kernel k of a two-kernel model acts on calculator k's descriptors in
    a batch as in a single structure (sparse_gp.cpp:651-660): summed means
    and each kernel's local variances equal the per-structure results."""
    from flare_amd.batch import Batch
    C, structs, calc, _ = setup
    calc2 = C.B2("chebyshev", "quadratic", [0.0, 4.5], [], [3, 3, 2], np.full((3, 3), 4.5))
    first = C.Structure(*structs[0], 5.5, [calc, calc2])
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0), C.NormalizedDotProduct(1.2, 1.0)],
                     0.01, 0.05, 0.005)
    sgp.add_all_environments(first)
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(9).normal(scale=0.1, size=sgp.n_sparse))
    b = Batch(structs[:6], 5.5, [calc, calc2], form=form)
    means, vars_ = b.predict(sgp)
    for k, (cell, species, pos) in enumerate(structs[:6]):
        s = C.Structure(cell, species, pos, 5.5, [calc, calc2], form=form)
        sgp.predict_local_uncertainties(s)
        assert np.array_equal(means[k], s.mean_efs), k
        assert len(vars_[k]) == 2
        for j in range(2):
            assert np.array_equal(vars_[k][j], s.local_uncertainties[j]), (k, j)
