"""Result conventions of SGP_Calculator (calculator.py:57-122, 188-206) on
the host: energy plus single-atom energies, forces reshaped, FLARE stress
(xx, xy, xz, yy, yz, zz) to ASE Voigt (xx, yy, zz, yz, xz, xy) with the
sign flipped, DTC/SOR stds as signed square roots of the force variances,
local stds sorted from species order to atom order and divided by
|sigma_f|. A stub model stands in for the GPU predictions."""
import numpy as np
import pytest

from flare_amd.calculator import predict_results, signed_sqrt, sort_variances


class Values:
    def __init__(self, atom_indices):
        self.atom_indices = atom_indices
        self.n_types = len(atom_indices)


class Struc:
    def __init__(self):
        self.species = [1, 0, 1, 0, 0]
        self.noa = 5
        self.descriptors = [Values([np.array([1, 3, 4]), np.array([0, 2])])]


class Stub:
    """Writes known outputs the way SparseGP.predict_* do."""

    def __init__(self):
        self.called = []
        n = 5
        self.mean = np.arange(1 + 3 * n + 6, dtype=float) * 0.1 - 0.5

    def _m(self, s, kind):
        self.called.append(kind)
        s.mean_efs = self.mean.copy()

    def predict_SOR(self, s, kind="SOR"):
        self._m(s, kind)
        s.variance_efs = np.linspace(-1, 1, len(self.mean))

    def predict_DTC(self, s):
        self.predict_SOR(s, "DTC")

    def predict_local_uncertainties(self, s):
        self._m(s, "local")
        s.local_uncertainties = [np.array([4.0, 9.0, -1.0, 16.0, 0.25])]

    def predict_mean(self, s):
        self._m(s, "mean")


def test_energy_forces_stress():
    s, m = Struc(), Stub()
    r = predict_results(m, s, "local", {0: 10.0, 1: -3.0}, 2.0)
    assert r["energy"] == pytest.approx(m.mean[0] + 3 * 10.0 + 2 * -3.0)
    np.testing.assert_array_equal(r["forces"], m.mean[1:16].reshape(5, 3))
    xx, xy, xz, yy, yz, zz = m.mean[-6:]
    np.testing.assert_array_equal(r["stress"], -np.array([xx, yy, zz, yz, xz, xy]))


def test_local_stds_atom_order():
    s, m = Struc(), Stub()
    r = predict_results(m, s, "local", None, -2.0)
    # Species order: species 0 atoms 1, 3, 4 then species 1 atoms 0, 2.
    var_atom = np.zeros(5)
    var_atom[[1, 3, 4, 0, 2]] = [4.0, 9.0, -1.0, 16.0, 0.25]
    np.testing.assert_array_equal(sort_variances(s, [4.0, 9.0, -1.0, 16.0, 0.25]), var_atom)
    want = np.zeros((5, 3))
    want[:, 0] = np.array([4.0, 2.0, 0.5, 3.0, -1.0]) / 2.0
    np.testing.assert_allclose(r["stds"], want, rtol=1e-15)


@pytest.mark.parametrize("kind", ["SOR", "DTC"])
def test_dense_stds(kind):
    s, m = Struc(), Stub()
    r = predict_results(m, s, kind, None, 1.0)
    v = np.linspace(-1, 1, 22)[1:-6]
    np.testing.assert_allclose(r["stds"].reshape(-1), np.sign(v) * np.sqrt(np.abs(v)))
    assert m.called == [kind]


def test_signed_sqrt_zero():
    assert signed_sqrt([0.0])[0] == 0.0 and signed_sqrt([-4.0])[0] == -2.0
