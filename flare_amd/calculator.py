"""FLARE's SGP_Calculator (flare/bffs/sgp/calculator.py:16-206) on the GPU
path, without the ASE dependency: ``calculate`` takes any object with
``numbers``, ``cell`` and ``positions`` (an ASE Atoms works) and fills
``results`` with the energy (plus single-atom energies), forces, the
stress in ASE order and sign, and the uncertainties of the model's variance
type; ``build_map`` writes the LAMMPS files (calculator.py:169-185).
"""
import json
from copy import deepcopy

import numpy as np

from . import _C_flare
from .sgp_wrapper import SGP_Wrapper, _plain


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
    """sqrt of positive variances, -sqrt|v| of negative ones (calculator.py:
    91-118 reports numerically negative variances that way)."""
    v = np.asarray(v, dtype=np.float64)
    return np.where(v > 0, np.sqrt(np.abs(v)), -np.sqrt(np.abs(v)))


def predict_results(sparse_gp, structure, variance_type, single_atom_energies, sigma_f):
    """calculator.py:57-122: run the prediction of the variance type on the
    structure and return the results dict. The reference predicts nothing
    for another variance type; here the mean alone is predicted then."""
    if variance_type == "SOR":
        sparse_gp.predict_SOR(structure)
    elif variance_type == "DTC":
        sparse_gp.predict_DTC(structure)
    elif variance_type == "local":
        sparse_gp.predict_local_uncertainties(structure)
    else:
        sparse_gp.predict_mean(structure)
    mean = np.array(structure.mean_efs, copy=True)
    results = {"energy": float(mean[0]), "forces": mean[1:-6].reshape(-1, 3)}
    if single_atom_energies is not None:
        for s in structure.species:
            results["energy"] += single_atom_energies[s]
    fs = mean[-6:]
    # FLARE stress (xx, xy, xz, yy, yz, zz) to ASE Voigt, sign flipped.
    results["stress"] = -np.array([fs[0], fs[3], fs[5], fs[4], fs[2], fs[1]])
    if variance_type in ("SOR", "DTC"):
        results["stds"] = signed_sqrt(structure.variance_efs[1:-6]).reshape(-1, 3)
    elif variance_type == "local":
        v = sort_variances(structure, structure.local_uncertainties[0])
        stds = np.zeros((len(v), 3))
        stds[:, 0] = signed_sqrt(v) / abs(sigma_f)
        results["stds"] = stds
    return results


class SGP_Calculator:
    """calculator.py:16-186 over an SGP_Wrapper."""

    implemented_properties = ["energy", "forces", "stress", "stds"]

    def __init__(self, sgp_model, use_mapping=False):
        self.gp_model = sgp_model
        self.results = {}
        self.use_mapping = use_mapping
        self.mgp_model = None

    def calculate(self, atoms=None, properties=None, system_changes=None):
        coded = [self.gp_model.species_map[int(z)] for z in atoms.numbers]
        structure = _C_flare.Structure(np.asarray(atoms.cell, dtype=np.float64), coded,
                                       atoms.positions, self.gp_model.cutoff,
                                       self.gp_model.descriptor_calculators)
        self.predict_on_structure(structure)
        return self.results

    def predict_on_structure(self, structure):
        self.results = predict_results(self.gp_model.sparse_gp, structure,
                                       self.gp_model.variance_type,
                                       self.gp_model.single_atom_energies,
                                       self.gp_model.hyps[0])
        return self.results

    def get_property(self, name, atoms=None):
        if atoms is not None:
            self.calculate(atoms)
        return self.results[name]

    def get_uncertainties(self, atoms):
        return self.get_property("stds", atoms)

    def calculation_required(self, atoms, quantities):
        return True

    def build_map(self, filename="lmp.flare", contributor="user", map_uncertainty=False):
        """calculator.py:169-185: the pair_style flare coefficient file and,
        for compute flare/std/atom, the variance map or the L_inv and
        sparse-descriptor files."""
        self.gp_model.sparse_gp.write_mapping_coefficients(filename, contributor, 0)
        if map_uncertainty:
            self.gp_model.write_varmap_coefficients(f"map_unc_{filename}", contributor, 0)
        else:
            self.gp_model.sparse_gp.write_L_inverse(f"L_inv_{filename}", contributor)
            self.gp_model.sparse_gp.write_sparse_descriptors(f"sparse_desc_{filename}",
                                                             contributor)

    def as_dict(self):
        out = {k: v for k, v in vars(self).items() if k not in ("gp_model", "mgp_model")}
        out["class"] = self.__class__.__name__
        out["gp_model"] = self.gp_model.as_dict()
        return out

    @staticmethod
    def from_dict(dct):
        sgp, _ = SGP_Wrapper.from_dict(dct["gp_model"])
        calc = SGP_Calculator(sgp, use_mapping=dct["use_mapping"])
        calc.results = dct.get("results", {})
        return calc

    def __deepcopy__(self, memo):
        d = self.as_dict()
        d["results"] = deepcopy(d["results"])
        return SGP_Calculator.from_dict(d)

    def write_model(self, name):
        if name[-5:] != ".json":
            name += ".json"
        with open(name, "w") as f:
            json.dump(_plain(self.as_dict()), f)

    @staticmethod
    def from_file(name):
        with open(name) as f:
            d = json.loads(f.readline())
        sgp, kernels = SGP_Wrapper.from_dict(d["gp_model"])
        return SGP_Calculator(sgp, use_mapping=d["use_mapping"]), kernels
