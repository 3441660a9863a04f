"""compute flare/std/atom files on the CPU: the L_inv, sparse-descriptor and
variance-map files written and read back (sparse_gp.cpp:1266-1466,
compute_flare_std_atom.cpp:329-686), and the variance map's algebra
(normalized_dot_product.cpp:769-830) against the closed form it must equal,
1 - sigma^2 |L^-1 k|^2 with L L^T = Kuu + jitter (power 1)."""
import types

import numpy as np
import pytest

from helpers import ndp_kuu

NOS, N, LMAX = 2, 3, 2
D = (NOS * N) * (NOS * N + 1) // 2 * (LMAX + 1)


def fake_sgp(seed=5, sigma=1.5, power=1.0, m=(7, 5)):
    """The attributes the writers read, from random sparse rows (one row
    of species 1 empty, as an isolated atom gives)."""
    from flare_amd import _C_flare as C
    rng = np.random.default_rng(seed)
    cd = C.ClusterDescriptor(NOS, D)
    cd.descriptors = [rng.normal(size=(k, D)) for k in m]
    cd.descriptors[1][2] = 0.0
    cd.descriptor_norms = [np.linalg.norm(x, axis=1) for x in cd.descriptors]
    kern = C.NormalizedDotProduct(sigma, power)
    Ms = sum(m)
    Kuu = np.zeros((Ms, Ms))
    o = 0
    for x, nr in zip(cd.descriptors, cd.descriptor_norms):
        k = len(x)
        Kuu[o:o + k, o:o + k] = ndp_kuu(x, nr, sigma, power)
        o += k
    jitter = 1e-8
    L = np.linalg.cholesky(Kuu + jitter * np.eye(Ms))
    sgp = types.SimpleNamespace(kernels=[kern], sparse_descriptors=[cd], n_kernels=1,
                                Kuu_jitter=jitter, L_inv=np.linalg.inv(L),
                                hyperparameters=np.array([sigma, 0.01, 0.05, 0.005]))
    calc = C.B2("chebyshev", "quadratic", [0.0, 4.5], [], [NOS, N, LMAX],
                np.array([[4.5, 4.0], [4.0, 4.25]]))
    return sgp, calc, Kuu


def test_L_inverse_and_sparse_files_roundtrip(tmp_path):
    from flare_amd.mapping import (read_L_inverse, read_sparse_descriptors, write_L_inverse,
                                   write_sparse_descriptors)
    sgp, calc, _ = fake_sgp(power=2.0)
    p1, p2 = str(tmp_path / "L_inv.flare"), str(tmp_path / "sparse_desc.flare")
    write_L_inverse(sgp, p1, "tester", [calc])
    write_sparse_descriptors(sgp, p2, "tester")
    info = read_L_inverse(p1)
    assert (info["power"], info["kernel_name"]) == (2, "NormalizedDotProduct")
    assert (info["n_species"], info["n_max"], info["l_max"], info["n_kernels"]) == (NOS, N, LMAX, 1)
    assert info["radial"] == "chebyshev" and info["cutoff_function"] == "quadratic"
    assert np.array_equal(info["cutoffs"], np.array([[4.5, 4.0], [4.0, 4.25]]))
    assert np.allclose(info["hyperparameters"], sgp.hyperparameters, rtol=1e-15, atol=0)
    assert info["n_clusters"] == 12
    assert np.allclose(info["L_inv"], np.tril(sgp.L_inv), rtol=1e-15, atol=1e-300)
    per_type = read_sparse_descriptors(p2, D)[0]
    cd = sgp.sparse_descriptors[0]
    for t in range(NOS):
        want = np.zeros_like(cd.descriptors[t])
        ok = cd.descriptor_norms[t] >= 1e-8
        want[ok] = cd.descriptors[t][ok] / cd.descriptor_norms[t][ok, None]
        assert np.allclose(per_type[t], want, rtol=1e-15, atol=1e-300)
    assert not per_type[1][2].any()  # the empty sparse row is written as zeros


def test_varmap_equals_L_inverse_form(tmp_path):
    from flare_amd.mapping import read_varmap_coefficients, write_varmap_coefficients
    sigma = 1.5
    sgp, calc, Kuu = fake_sgp(sigma=sigma, power=1.0)
    path = str(tmp_path / "varmap.flare")
    write_varmap_coefficients(sgp, path, "tester", 0, calc)
    info = read_varmap_coefficients(path)
    assert info["kernel_name"] == "NormalizedDotProduct" and info["n_descriptors"] == D
    beta = info["beta"]
    assert beta.shape == (NOS, D, D)
    cd = sgp.sparse_descriptors[0]
    rng = np.random.default_rng(9)
    o = 0
    for t in range(NOS):
        m = cd.n_clusters_by_type[t]
        bt = beta[t]
        assert not bt[:, m:].any()  # columns past M_s are zero
        Xh = np.zeros_like(cd.descriptors[t])
        ok = cd.descriptor_norms[t] >= 1e-8
        Xh[ok] = cd.descriptors[t][ok] / cd.descriptor_norms[t][ok, None]
        L = np.linalg.cholesky(Kuu[o:o + m, o:o + m] + sgp.Kuu_jitter * np.eye(m))
        o += m
        for _ in range(5):
            b = rng.normal(size=D)
            b /= np.linalg.norm(b)
            q = bt.T @ b
            k = sigma ** 2 * (Xh @ b)  # power 1 kernel column
            z = np.linalg.solve(L, k)
            var_map = 1.0 - q @ q
            var_ref = 1.0 - (z @ z) / sigma ** 2
            assert var_map == pytest.approx(var_ref, abs=1e-9)


def test_varmap_power_two_rejected():
    from flare_amd.mapping import varmap_coefficients
    sgp, _, _ = fake_sgp(power=2.0)
    with pytest.raises(ValueError, match="power 1"):
        varmap_coefficients(sgp)
