"""SGP_Wrapper: the Python model object of FLARE's sparse GP
(flare/bffs/sgp/sparse_gp.py:20-420) on the GPU path, with its model file
(as_dict / from_dict / write_model / from_file, sparse_gp.py:120-320).

The file stores the kernels, the B2 calculator, hyperparameters, species
map and the training structures as ASE Atoms dictionaries (numbers,
positions, cell, and the labels under "results"); loading replays the
training (update_db per structure with its sparse atoms, then
update_matrices_QR), as the reference does. Structures are duck-typed: any
object with ``numbers``, ``cell`` and ``positions`` (an ASE Atoms works),
or a _C_flare.Structure with coded species.
"""
import json
import warnings

import numpy as np

from . import _C_flare as C


class AtomsRecord:
    """The fields of an ASE Atoms dictionary the model file uses."""

    def __init__(self, numbers, positions, cell, energy=None, forces=None, stress=None,
                 info=None):
        self.numbers = np.asarray(numbers, dtype=np.int64)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.energy = np.zeros(0) if energy is None else np.asarray(energy, np.float64).reshape(-1)
        self.forces = None if forces is None else np.asarray(forces, np.float64).reshape(-1, 3)
        self.stress = None if stress is None else np.asarray(stress, np.float64).reshape(-1)
        self.info = dict(info or {})

    @staticmethod
    def _array(v):
        # ASE's JSON encoding of arrays ({"__ndarray__": [shape, dtype, flat]})
        # or plain lists (NumpyEncoder).
        if isinstance(v, dict) and "__ndarray__" in v:
            shape, dtype, flat = v["__ndarray__"]
            return np.array(flat, dtype=dtype).reshape(shape)
        if isinstance(v, dict) and "array" in v:
            return AtomsRecord._array(v["array"])
        return None if v is None else np.asarray(v)

    @classmethod
    def from_dict(cls, d):
        res = d.get("results") or {}
        e = res.get("energy")
        return cls(cls._array(d["numbers"]), cls._array(d["positions"]), cls._array(d["cell"]),
                   energy=None if e is None else np.atleast_1d(cls._array(e)),
                   forces=cls._array(res.get("forces")), stress=cls._array(res.get("stress")),
                   info={k: cls._array(v) if isinstance(v, (dict, list)) else v
                         for k, v in (d.get("info") or {}).items()})

    def as_dict(self):
        res = {"energy": self.energy.tolist()}
        if self.forces is not None:
            res["forces"] = self.forces.tolist()
        if self.stress is not None:
            res["stress"] = self.stress.tolist()
        return {"numbers": self.numbers.tolist(), "positions": self.positions.tolist(),
                "cell": self.cell.tolist(), "pbc": [True, True, True],
                "info": {k: np.asarray(v).tolist() for k, v in self.info.items()},
                "results": res}


class SGP_Wrapper:
    """sparse_gp.py:20-120. kernels: NormalizedDotProduct list;
    descriptor_calculators: [B2]; species_map: atomic number -> coded
    species."""

    def __init__(self, kernels, descriptor_calculators, cutoff, sigma_e, sigma_f, sigma_s,
                 species_map, variance_type="SOR", single_atom_energies=None,
                 energy_training=True, force_training=True, stress_training=True,
                 max_iterations=10, opt_method="BFGS", bounds=None):
        self.sparse_gp = C.SparseGP(kernels, sigma_e, sigma_f, sigma_s)
        self.descriptor_calculators = descriptor_calculators
        self.cutoff = cutoff
        self.hyps_mask = None
        self.species_map = species_map
        self.variance_type = variance_type
        self.single_atom_energies = single_atom_energies
        self.energy_training = energy_training
        self.force_training = force_training
        self.stress_training = stress_training
        self.max_iterations = max_iterations
        self.opt_method = opt_method
        self.bounds = bounds
        self.atom_indices = []
        self.rel_efs_noise = []
        self.hyp_labels = [f"Hyp{n}" for n in range(len(self.hyps))]
        self.sgp_var = None
        if isinstance(kernels[0], C.NormalizedDotProduct):
            if kernels[0].power == 1:
                self.sgp_var_flag = "self"
                self.sgp_var = self.sparse_gp
            else:
                self.sgp_var_flag = "new"
        else:
            warnings.warn("kernels[0] should be NormalizedDotProduct for variance mapping")
            self.sgp_var_flag = None

    training_data = property(lambda self: self.sparse_gp.training_structures)
    hyps = property(lambda self: self.sparse_gp.hyperparameters)
    hyps_and_labels = property(lambda self: (self.hyps, self.hyp_labels))
    likelihood = property(lambda self: self.sparse_gp.log_marginal_likelihood)
    likelihood_gradient = property(lambda self: self.sparse_gp.likelihood_gradient)
    force_noise = property(lambda self: self.sparse_gp.force_noise)

    def __len__(self):
        return len(self.training_data)

    def update_db(self, structure, forces, custom_range=(), energy=None, stress=None,
                  mode="specific", sgp=None, update_qr=True, atom_indices=(-1,),
                  rel_e_noise=1.0, rel_f_noise=1.0, rel_s_noise=1.0):
        """sparse_gp.py:322-418."""
        if isinstance(structure, C.Structure):
            coded = list(structure.species)
        else:
            coded = [self.species_map[int(z)] for z in structure.numbers]
        s = C.Structure(structure.cell, coded, structure.positions, self.cutoff,
                        self.descriptor_calculators)
        if energy is not None and self.energy_training:
            sae = 0.0
            if self.single_atom_energies is not None:
                sae = sum(self.single_atom_energies[c] for c in coded)
            s.energy = np.array([[energy - sae]])
        if forces is not None and self.force_training:
            s.forces = np.asarray(forces, dtype=np.float64).reshape(-1)
        if stress is not None and self.stress_training:
            st = np.asarray(stress, dtype=object)
            if st.shape != () or st.item() is not None:
                s.stresses = np.asarray(stress, dtype=np.float64).reshape(-1)
        if sgp is None:
            sgp = self.sparse_gp
            self.atom_indices.append(list(atom_indices))
        sgp.add_training_structure(s, list(atom_indices), rel_e_noise, rel_f_noise, rel_s_noise)
        self.rel_efs_noise.append([rel_e_noise, rel_f_noise, rel_s_noise])
        if mode == "all":
            if len(custom_range):
                raise Exception("Set mode='specific' for a user-defined custom_range")
            sgp.add_all_environments(s)
        elif mode == "uncertain":
            if len(custom_range) != 1:
                raise Exception("The custom_range should be set as [n_added] if mode='uncertain'")
            sgp.add_uncertain_environments(s, list(custom_range))
        elif mode == "specific":
            if len(custom_range) == 0:
                warnings.warn("The mode='specific' but no custom_range is given, will not add "
                              "sparse envs")
            else:
                sgp.add_specific_environments(s, list(custom_range))
        elif mode == "random":
            if len(custom_range) != 1:
                raise Exception("The custom_range should be set as [n_added] if mode='random'")
            sgp.add_random_environments(s, list(custom_range))
        else:
            raise NotImplementedError
        if update_qr:
            sgp.update_matrices_QR()

    def write_varmap_coefficients(self, filename, contributor, kernel_idx):
        """sparse_gp.py write_varmap_coefficients for a power-1 kernel (the
        model itself; other powers need a separately trained power-1 model,
        which this build does not make)."""
        if self.sgp_var_flag != "self":
            raise NotImplementedError("variance maps need a power-1 kernel")
        self.sparse_gp.write_varmap_coefficients(filename, contributor, kernel_idx)

    # Model file (sparse_gp.py:120-320) --------------------------------------
    def as_dict(self):
        out = {}
        for key, val in vars(self).items():
            if key not in ("sparse_gp", "sgp_var", "descriptor_calculators"):
                out[key] = val
        b2 = self.descriptor_calculators[0]
        out["descriptor_calculators"] = [{
            "type": "B2", "radial_basis": b2.radial_basis, "cutoff_function": b2.cutoff_function,
            "radial_hyps": b2.radial_hyps, "cutoff_hyps": b2.cutoff_hyps,
            "descriptor_settings": b2.descriptor_settings, "cutoffs": b2.cutoffs.tolist()}]
        out["hyps"], out["hyp_labels"] = list(self.hyps), self.hyp_labels
        out["Kuu_jitter"] = self.sparse_gp.Kuu_jitter
        out["kernels"] = [(k.kernel_name, k.sigma, k.power) for k in self.sparse_gp.kernels]
        inv = {v: k for k, v in self.species_map.items()}
        strucs = []
        for s, st in enumerate(self.training_data):
            sae = 0.0
            if self.single_atom_energies is not None:
                sae = sum(self.single_atom_energies[c] for c in st.species)
            e = np.asarray(st.energy, dtype=np.float64).reshape(-1)
            rec = AtomsRecord([inv[c] for c in st.species], st.positions, st.cell,
                              energy=e + sae if e.size else None,
                              forces=np.asarray(st.forces).reshape(st.noa, 3)
                              if np.size(st.forces) else None,
                              stress=np.asarray(st.stresses) if np.size(st.stresses) else None,
                              info={"rel_efs_noise": np.array(self.rel_efs_noise[s])})
            strucs.append(rec.as_dict())
        out["training_structures"] = strucs
        out["sparse_indice"] = self.sparse_gp.sparse_indices
        return out

    def write_model(self, name):
        if name[-5:] != ".json":
            name += ".json"
        with open(name, "w") as f:
            json.dump(_plain(self.as_dict()), f)

    @staticmethod
    def from_dict(in_dict):
        """sparse_gp.py:215-313: rebuild and retrain."""
        (kname, sigma, power), = in_dict["kernels"]
        if kname != "NormalizedDotProduct":
            raise ValueError(f"kernel {kname!r} is not supported (NormalizedDotProduct)")
        kernel = C.NormalizedDotProduct(float(sigma), int(power))
        (b2,) = in_dict["descriptor_calculators"]
        assert b2["type"] == "B2"
        calc = C.B2(b2["radial_basis"], b2["cutoff_function"], b2["radial_hyps"],
                    b2["cutoff_hyps"], b2["descriptor_settings"], b2["cutoffs"])
        sae = in_dict.get("single_atom_energies")
        sae = None if sae is None else {int(k): v for k, v in sae.items()}
        species_map = {int(k): v for k, v in in_dict["species_map"].items()}
        hyps = in_dict["hyps"]
        gp = SGP_Wrapper([kernel], [calc], in_dict["cutoff"], hyps[-3], hyps[-2], hyps[-1],
                         species_map, variance_type=in_dict["variance_type"],
                         single_atom_energies=sae,
                         energy_training=in_dict["energy_training"],
                         force_training=in_dict["force_training"],
                         stress_training=in_dict["stress_training"],
                         max_iterations=in_dict["max_iterations"],
                         opt_method=in_dict["opt_method"], bounds=in_dict["bounds"])
        gp.sparse_gp.Kuu_jitter = in_dict["Kuu_jitter"]
        for s, d in enumerate(in_dict["training_structures"]):
            rec = AtomsRecord.from_dict(d)
            custom_range = in_dict["sparse_indice"][0][s]
            atom_indices = in_dict["atom_indices"][s] if "atom_indices" in in_dict else [-1]
            energy = rec.energy[0] if rec.energy.size else None
            rel = rec.info.get("rel_efs_noise", [1, 1, 1])
            stress = rec.stress if in_dict["stress_training"] else None
            gp.update_db(rec, rec.forces, custom_range=custom_range, energy=energy,
                         stress=stress, mode="specific", sgp=None, update_qr=False,
                         atom_indices=atom_indices, rel_e_noise=rel[0], rel_f_noise=rel[1],
                         rel_s_noise=rel[2])
        gp.sparse_gp.update_matrices_QR()
        return gp, [kernel]

    @staticmethod
    def from_file(filename):
        with open(filename) as f:
            return SGP_Wrapper.from_dict(json.loads(f.readline()))


def _plain(obj):
    """numpy values to JSON types (what the reference's NumpyEncoder does)."""
    if isinstance(obj, dict):
        return {str(k): _plain(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_plain(v) for v in obj]
    if isinstance(obj, np.ndarray):
        return _plain(obj.tolist())
    if isinstance(obj, np.generic):
        return obj.item()
    return obj
