"""Oracle parity at the benchmark configurations' own settings (the sizes
the oracle finishes in seconds; FP64 drop-in bar 1e-10 relative on E, F, S,
1e-10 sigma_f^2 on the local variances), and the chosen form's error budget
there (force 1 meV/A, |d sigma| / sigma_f and relative sigma error of atoms
with sigma >= 0.01 sigma_f at 1e-3):

  32k-atom alloy model  B2 [2,12,6] rc 6.0 (d = 2100), the 2,000 sparse
                        environments drawn from the 32,000-atom alloy (seed 3),
                        evaluated on a 4,000-atom sample of the alloy
  512-structure batch   B2 [3,8,4] rc 5.5 (d = 1500), 1,500 sparse from the
                        first structures (seed 6), real batch entries
                        including a slab and the 2,048-atom one, batched
  served mix            the same model through the server (FP64), a mix of
                        those structures
  1M-atom BCC model     B2 [3,10,6] rc 5.5 (d = 3255), 4,096 sparse (seed 9),
                        on a seeded 4,096-atom BCC alloy cell (seed 10)
"""
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


def check_chosen(C, sgp, calc, cell, species, pos, cutoff, sig_f, label):
    """Chosen form against the FP64 path on the same model."""
    s64 = C.Structure(cell, species, pos, cutoff, [calc], form=0)
    s32 = C.Structure(cell, species, pos, cutoff, [calc], form=1)
    sgp.predict_local_uncertainties(s64)
    sgp.predict_local_uncertainties(s32)
    n = len(pos)
    dF = np.abs(s32.mean_efs[1:1 + 3 * n] - s64.mean_efs[1:1 + 3 * n]).max()
    dE = abs(s32.mean_efs[0] - s64.mean_efs[0]) / n
    sd64 = np.sqrt(np.abs(s64.local_uncertainties[0]))
    sd32 = np.sqrt(np.abs(s32.local_uncertainties[0]))
    big = sd64 >= 0.01 * sig_f
    sig_abs = np.max(np.abs(sd32 - sd64)) / sig_f
    sig_rel = np.max(np.abs(sd32 - sd64)[big] / sd64[big]) if big.any() else 0.0
    say(f"{label} chosen form: dE {dE * 1e3:.2e} meV/atom, dF {dF * 1e3:.3f} meV/A, "
        f"|dsigma|/sigma_f {sig_abs:.2e}, rel (sigma >= 0.01 sigma_f, {big.mean():.0%}) "
        f"{sig_rel:.2e}")
    assert dF < 1e-3 and dE < 1e-3
    assert sig_abs < 1e-3 and sig_rel < 1e-3


def test_alloy_32k_model_on_4000_atoms(C, oracle):
    from flare_amd.workloads import fcc_alloy_32k, random_alloy
    say("32k-atom alloy model: building the 2,000-sparse model")
    cell, species, pos = fcc_alloy_32k()
    calc = C.B2("chebyshev", "quadratic", [0.0, 6.0], [], [2, 12, 6], np.full((2, 2), 6.0))
    big = C.Structure(cell, species, pos, 6.0, [calc])
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    atoms = sorted(np.random.default_rng(3).choice(len(pos), 2000, replace=False).tolist())
    sgp.add_specific_environments(big, atoms)
    del big
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(4).normal(scale=0.1, size=sgp.n_sparse))
    assert sgp.n_sparse == 2000
    c4, s4, p4 = random_alloy((10, 10, 10), a=4.0, nos=2, sigma=0.08, seed=12)
    s = C.Structure(c4, s4, p4, 6.0, [calc])
    sgp.predict_local_uncertainties(s)
    cd = sgp.sparse_descriptors[0]
    m = oracle.Model(2.0, 2.0, cd.descriptors, cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv,
                     fast=True)
    nb = oracle.Neighbors(c4, s4, p4, 6.0, "cells", fast=True)
    mean, var_atom, _ = m.predict_scalable(nb, 2, 12, 6, np.full((2, 2), 6.0))
    order = np.concatenate(s.descriptors[0].atom_indices)
    em, ev = rel_err(s.mean_efs, mean), var_err(s.local_uncertainties[0], var_atom[order], 4.0)
    say(f"4000 atoms vs oracle: mean {em:.2e}, variance {ev:.2e}")
    assert em < TOL64 and ev < TOL64
    check_chosen(C, sgp, calc, c4, s4, p4, 6.0, 2.0, "32k-atom alloy model")


@pytest.fixture(scope="module")
def batch_model(C):
    from flare_amd.workloads import batch_structures
    structs = batch_structures(512, seed=5)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 8, 4], np.full((3, 3), 5.5))
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    rng = np.random.default_rng(6)
    need = 1500
    for st in structs:
        if need <= 0:
            break
        s = C.Structure(*st, 5.5, [calc])
        take = min(need, len(st[1]))
        sgp.add_specific_environments(s, sorted(rng.choice(len(st[1]), take, replace=False)))
        need -= take
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(4).normal(scale=0.1, size=sgp.n_sparse))
    return structs, calc, sgp


def is_slab(s):
    """Atoms fill under three quarters of the third lattice vector."""
    frac = s[2] @ np.linalg.inv(s[0])
    return np.ptp(frac[:, 2]) < 0.75


def pick_entries(structs):
    """The largest structure, the smallest, a slab and two more."""
    n = [len(s[1]) for s in structs]
    slabs = [i for i, s in enumerate(structs) if is_slab(s)]
    picks = [int(np.argmax(n)), int(np.argmin(n)), 7, 100, slabs[0]]
    return sorted(set(picks))


def oracle_check(oracle, sgp, structs, idx, means, vars_, label):
    cd = sgp.sparse_descriptors[0]
    m = oracle.Model(2.0, 2.0, cd.descriptors, cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv,
                     fast=True)
    for j, i in enumerate(idx):
        cell, species, pos = structs[i]
        nb = oracle.Neighbors(cell, species, pos, 5.5, "cells", fast=True)
        mean, var_atom, _ = m.predict_scalable(nb, 3, 8, 4, np.full((3, 3), 5.5))
        order = np.argsort(species, kind="stable")
        em, ev = rel_err(means[j], mean), var_err(vars_[j], var_atom[order], 4.0)
        say(f"{label} structure {i} ({len(pos)} atoms): mean {em:.2e}, variance {ev:.2e}")
        assert em < TOL64 and ev < TOL64


def test_batch_512_entries(C, oracle, batch_model):
    from flare_amd.batch import Batch
    structs, calc, sgp = batch_model
    assert sgp.n_sparse == 1500
    idx = pick_entries(structs)
    assert any(is_slab(structs[i]) for i in idx)
    b = Batch([structs[i] for i in idx], 5.5, calc)
    means, vars_ = b.predict(sgp)
    oracle_check(oracle, sgp, structs, idx, means, [v[0] for v in vars_], "batched")
    i = idx[-1]
    check_chosen(C, sgp, calc, *structs[i], 5.5, 2.0, f"batch structure {i}")


def test_served_mix(C, oracle, batch_model):
    from flare_amd.serving import NativeServer
    structs, calc, sgp = batch_model
    idx = pick_entries(structs)
    srv = NativeServer(sgp, calc, 5.5, form=0, hbm_budget=8 << 30, n_workers=2)
    try:
        ids = [srv.submit(*structs[i]) for i in idx + list(range(20, 40))]
        res = [srv.poll(r, timeout_ms=-1) for r in ids]
    finally:
        srv.close()
    oracle_check(oracle, sgp, structs, idx, [r[0] for r in res[:len(idx)]],
                 [r[1][0] for r in res[:len(idx)]], "served")


def test_bcc_1m_model_on_4096_atoms(C, oracle):
    from flare_amd.workloads import BCC_BASIS, random_alloy
    say("1M-atom BCC model: 4,096 sparse environments")
    cut = np.full((3, 3), 5.5)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 10, 6], cut)
    cs, ss, ps = random_alloy((16, 16, 16), a=3.2, nos=3, sigma=0.1, seed=11, basis=BCC_BASIS)
    small = C.Structure(cs, ss, ps, 5.5, [calc])
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    picks = np.random.default_rng(9).choice(len(ss), 4096, replace=False)
    sgp.add_specific_environments(small, sorted(int(a) for a in picks))
    del small
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(4).normal(scale=0.1, size=sgp.n_sparse))
    cell, species, pos = random_alloy((16, 16, 8), a=3.2, nos=3, sigma=0.1, seed=10,
                                      basis=BCC_BASIS)
    assert len(pos) == 4096
    s = C.Structure(cell, species, pos, 5.5, [calc])
    sgp.predict_local_uncertainties(s)
    cd = sgp.sparse_descriptors[0]
    m = oracle.Model(2.0, 2.0, cd.descriptors, cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv,
                     fast=True)
    nb = oracle.Neighbors(cell, species, pos, 5.5, "cells", fast=True)
    mean, var_atom, _ = m.predict_scalable(nb, 3, 10, 6, cut)
    order = np.concatenate(s.descriptors[0].atom_indices)
    em, ev = rel_err(s.mean_efs, mean), var_err(s.local_uncertainties[0], var_atom[order], 4.0)
    say(f"4096 BCC atoms vs oracle: mean {em:.2e}, variance {ev:.2e}")
    assert em < TOL64 and ev < TOL64
    check_chosen(C, sgp, calc, cell, species, pos, 5.5, 2.0, "1M-atom BCC model")
