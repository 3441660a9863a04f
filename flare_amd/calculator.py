"""Result conventions of FLARE's SGP_Calculator on top of the GPU path
(flare/bffs/sgp/calculator.py:28-122, 188-206), without the ASE dependency:
``predict_on_structure`` fills a results dict with the energy (plus single
atom energies), forces, the stress in ASE order and sign, and the
uncertainties of the chosen variance type.
"""
import numpy as np

from . import _C_flare
from ._lib import FORM_FP64


def sort_variances(structure, variances):
    """Species-sorted variances back to atom order (calculator.py:188-206)."""
    assert len(variances) == structure.noa
    out = np.zeros(len(variances))
    vals = structure.descriptors[0]
    v = 0
    for s in range(vals.n_types):
        for a in vals.atom_indices[s]:
            out[a] = variances[v]
            v += 1
    return out


def signed_sqrt(v):
    v = np.asarray(v, dtype=np.float64)
    return np.where(v > 0, np.sqrt(np.abs(v)), -np.sqrt(np.abs(v)))


class SGPCalculator:
    """E/F/S/stds from a SparseGP (calculator.py:16-122).

    ``species_map`` maps atomic numbers to coded species; ``hyps[0]`` is
    the signal sigma that local stds are divided by (calculator.py:120).
    """

    implemented_properties = ["energy", "forces", "stress", "stds"]

    def __init__(self, sparse_gp, descriptor_calculators, cutoff, species_map,
                 single_atom_energies=None, variance_type="local", form=FORM_FP64):
        self.sparse_gp = sparse_gp
        self.descriptor_calculators = descriptor_calculators
        self.cutoff = cutoff
        self.species_map = species_map
        self.single_atom_energies = single_atom_energies
        self.variance_type = variance_type
        self.form = form
        self.results = {}

    def calculate(self, cell, numbers, positions):
        coded = [self.species_map[z] for z in numbers]
        structure = _C_flare.Structure(cell, coded, positions, self.cutoff,
                                       self.descriptor_calculators, form=self.form)
        self.predict_on_structure(structure)
        return self.results

    def predict_on_structure(self, structure):
        if self.variance_type == "SOR":
            self.sparse_gp.predict_SOR(structure)
        elif self.variance_type == "DTC":
            self.sparse_gp.predict_DTC(structure)
        elif self.variance_type == "local":
            self.sparse_gp.predict_local_uncertainties(structure)
        else:
            self.sparse_gp.predict_mean(structure)
        mean = np.array(structure.mean_efs, copy=True)
        self.results["energy"] = float(mean[0])
        self.results["forces"] = mean[1:-6].reshape(-1, 3)
        if self.single_atom_energies is not None:
            for s in structure.species:
                self.results["energy"] += self.single_atom_energies[s]
        fs = mean[-6:]
        # FLARE stress (xx, xy, xz, yy, yz, zz) to ASE Voigt, sign flipped.
        self.results["stress"] = -np.array([fs[0], fs[3], fs[5], fs[4], fs[2], fs[1]])
        if self.variance_type in ("SOR", "DTC"):
            self.results["stds"] = signed_sqrt(structure.variance_efs[1:-6]).reshape(-1, 3)
        elif self.variance_type == "local":
            v = sort_variances(structure, structure.local_uncertainties[0])
            stds = np.zeros((len(v), 3))
            stds[:, 0] = signed_sqrt(v) / abs(self.sparse_gp.hyperparameters[0])
            self.results["stds"] = stds
        return self.results
