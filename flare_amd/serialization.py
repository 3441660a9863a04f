"""Model files: SparseGP::to_json / from_json (sparse_gp.cpp:1468-1479)
in the reference's JSON layout, which nlohmann writes from the member lists
of sparse_gp.h:102-108, structure.h:104-109, descriptor.h:50-56,87-89 and
normalized_dot_product.h:68-69, with Eigen matrices as lists of rows,
vectors as lists and empty Eigen objects as null (json.h:9-72).

A model read here is resident for prediction at once (alpha, L_inv, Sigma,
R_inv and the sparse set come from the file); its training structures are
rebuilt on the GPU from their cell, species, positions, cutoff and
descriptor calculators, so the kernel matrices can grow with more data.
"""
import json
import math

import numpy as np


def _vec(a):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    return None if a.size == 0 else [_num(x) for x in a]


def _ivec(a):
    a = np.asarray(a).reshape(-1)
    return None if a.size == 0 else [int(x) for x in a]


def _num(x):
    x = float(x)
    return x if math.isfinite(x) else None  # nlohmann writes NaN / inf as null


def _mat(a):
    a = np.asarray(a, dtype=np.float64)
    if a.size == 0:
        return None
    a = a.reshape(a.shape[0], -1)
    return [[_num(x) for x in row] for row in a]


def _read_vec(j, dtype=np.float64):
    if j is None:
        return np.zeros(0, dtype)
    return np.array([np.nan if x is None else x for x in j], dtype=dtype)


def _read_mat(j):
    if j is None or len(j) == 0:
        return np.zeros((0, 0))
    return np.array([[np.nan if x is None else x for x in row] for row in j], dtype=np.float64)


def kernel_to_json(kern):
    """NLOHMANN_DEFINE_TYPE_INTRUSIVE(NormalizedDotProduct, sigma, sig2,
    power, kernel_name, kernel_hyperparameters)."""
    return {"sigma": kern.sigma, "sig2": kern.sigma ** 2, "power": float(kern.power),
            "kernel_name": kern.kernel_name,
            "kernel_hyperparameters": _vec(kern.kernel_hyperparameters)}


def kernel_from_json(j):
    """kernel.cpp:103-131: kernels by name."""
    from . import _C_flare as C
    name = j["kernel_name"]
    if name != "NormalizedDotProduct":
        raise ValueError(f"kernel {name!r} is not supported (NormalizedDotProduct)")
    return C.NormalizedDotProduct(j["sigma"], j["power"])


def descriptor_to_json(calc):
    """b2.cpp:548-558."""
    return {"radial_basis": calc.radial_basis, "cutoff_function": calc.cutoff_function,
            "radial_hyps": list(calc.radial_hyps), "cutoff_hyps": list(calc.cutoff_hyps),
            "descriptor_settings": list(calc.descriptor_settings), "cutoffs": _mat(calc.cutoffs),
            "descriptor_name": calc.descriptor_name}


def descriptor_from_json(j):
    """descriptor.cpp:24-39, b2.cpp:560-569 (the reference passes radial_hyps
    as cutoff_hyps; the quadratic and cosine cutoffs ignore them)."""
    from . import _C_flare as C
    if j.get("descriptor_name") != "B2":
        raise ValueError("only B2 descriptors are read from model files "
                         "(descriptor.cpp:28-38 reads no other)")
    return C.B2(j["radial_basis"], j["cutoff_function"], j["radial_hyps"], j["radial_hyps"],
                j["descriptor_settings"], _read_mat(j["cutoffs"]))


def cluster_to_json(cd):
    """descriptor.h:87-89."""
    return {"descriptors": [_mat(x) for x in cd.descriptors],
            "descriptor_norms": [_vec(x) for x in cd.descriptor_norms],
            "cutoff_values": [_vec(x) for x in cd.cutoff_values],
            "n_clusters_by_type": list(cd.n_clusters_by_type),
            "cumulative_type_count": list(cd.cumulative_type_count),
            "n_descriptors": cd.n_descriptors, "n_types": cd.n_types,
            "n_clusters": cd.n_clusters}


def cluster_from_json(j):
    from . import _C_flare as C
    cd = C.ClusterDescriptor(int(j["n_types"]), int(j["n_descriptors"]))
    nd = cd.n_descriptors
    for s in range(cd.n_types):
        d = _read_mat(j["descriptors"][s])
        cd.descriptors[s] = d.reshape(-1, nd) if d.size else np.zeros((0, nd))
        cd.descriptor_norms[s] = _read_vec(j["descriptor_norms"][s])
        cd.cutoff_values[s] = _read_vec(j["cutoff_values"][s])
    return cd


def values_to_json(v):
    """DescriptorValues, descriptor.h:50-56."""
    return {"n_descriptors": v.n_descriptors, "n_types": v.n_types, "n_atoms": v.n_atoms,
            "volume": v.volume,
            "descriptors": [_mat(x) for x in v.descriptors],
            "descriptor_force_dervs": [_mat(x) for x in v.descriptor_force_dervs],
            "neighbor_coordinates": [_mat(x) for x in v.neighbor_coordinates],
            "descriptor_norms": [_vec(x) for x in v.descriptor_norms],
            "descriptor_force_dots": [_vec(x) for x in v.descriptor_force_dots],
            "cutoff_values": [_vec(x) for x in v.cutoff_values],
            "cutoff_dervs": [_vec(x) for x in v.cutoff_dervs],
            "neighbor_counts": [_ivec(x) for x in v.neighbor_counts],
            "cumulative_neighbor_counts": [_ivec(x) for x in v.cumulative_neighbor_counts],
            "atom_indices": [_ivec(x) for x in v.atom_indices],
            "neighbor_indices": [_ivec(x) for x in v.neighbor_indices],
            "n_clusters": v.n_clusters, "n_clusters_by_type": list(v.n_clusters_by_type),
            "cumulative_type_count": list(v.cumulative_type_count),
            "n_neighbors_by_type": list(v.n_neighbors_by_type)}


def structure_to_json(st):
    """structure.h:104-109 (neighbour list and descriptors from the
    device)."""
    cell = np.asarray(st.cell)
    ct = cell.T
    cd = cell @ cell.T
    rel = st.relative_positions
    return {"neighbor_count": _ivec(st.neighbor_count), "cutoff": st.cutoff,
            "cumulative_neighbor_count": _ivec(st.cumulative_neighbor_count),
            "structure_indices": _ivec(st.structure_indices),
            "neighbor_species": _ivec(st.neighbor_species),
            "cell": _mat(cell), "cell_transpose": _mat(ct),
            "cell_transpose_inverse": _mat(np.linalg.inv(ct)), "cell_dot": _mat(cd),
            "cell_dot_inverse": _mat(np.linalg.inv(cd)), "positions": _mat(st.positions),
            "wrapped_positions": _mat(st.wrapped_positions), "relative_positions": _mat(rel),
            "single_sweep_cutoff": st.single_sweep_cutoff, "volume": st.volume,
            "sweep": st.sweep, "n_neighbors": st.n_neighbors, "species": list(st.species),
            "noa": st.noa, "energy": _vec(st.energy), "forces": _vec(st.forces),
            "stresses": _vec(st.stresses), "mean_efs": _vec(st.mean_efs),
            "variance_efs": _vec(st.variance_efs),
            "mean_contributions": [_vec(x) for x in st.mean_contributions],
            "local_uncertainties": [_vec(x) for x in st.local_uncertainties],
            "descriptor_calculators": [descriptor_to_json(c) for c in st.descriptor_calculators],
            "descriptors": [values_to_json(v) for v in st.descriptors]}


def structure_from_json(j, context=None):
    """A GPU Structure rebuilt from the stored geometry and calculators,
    with the stored labels."""
    from . import _C_flare as C
    calcs = [descriptor_from_json(c) for c in (j.get("descriptor_calculators") or [])]
    cell = _read_mat(j["cell"])
    pos = _read_mat(j["positions"]).reshape(-1, 3)
    st = C.Structure(cell, j["species"], pos, j.get("cutoff"), calcs, context=context)
    st.energy = _read_vec(j.get("energy"))
    st.forces = _read_vec(j.get("forces"))
    st.stresses = _read_vec(j.get("stresses"))
    return st


def sparse_gp_to_json(sgp):
    """sparse_gp.h:102-108, member by member."""
    nan = float("nan")
    return {
        "hyperparameters": _vec(sgp.hyperparameters),
        "kernels": [kernel_to_json(k) for k in sgp.kernels],
        "Kuu_kernels": [_mat(x) for x in sgp.Kuu_kernels],
        "Kuf_kernels": [_mat(x) for x in sgp.Kuf_kernels],
        "Kuu": _mat(sgp.Kuu), "Kuf": _mat(sgp.Kuf), "n_kernels": sgp.n_kernels,
        "Kuu_jitter": sgp.Kuu_jitter, "Sigma": _mat(sgp.Sigma),
        "Kuu_inverse": _mat(sgp.Kuu_inverse), "R_inv": _mat(sgp.R_inv), "L_inv": _mat(sgp.L_inv),
        "alpha": _vec(sgp.alpha), "R_inv_diag": _vec(sgp.R_inv_diag),
        "L_diag": _vec(sgp.L_diag),
        "sparse_descriptors": [cluster_to_json(c) for c in sgp.sparse_descriptors],
        "training_structures": [structure_to_json(s) for s in sgp.training_structures],
        "sparse_indices": sgp.sparse_indices,
        "noise_vector": _vec(sgp.noise_vector), "y": _vec(sgp.y),
        "label_count": _vec(sgp.label_count),
        "n_energy_labels": sgp.n_energy_labels, "n_force_labels": sgp.n_force_labels,
        "n_stress_labels": sgp.n_stress_labels, "n_sparse": sgp.n_sparse,
        "n_labels": sgp.n_labels, "n_strucs": sgp.n_strucs,
        "energy_noise": sgp.energy_noise, "force_noise": sgp.force_noise,
        "stress_noise": sgp.stress_noise,
        "log_marginal_likelihood": _num(getattr(sgp, "log_marginal_likelihood", nan)),
        "data_fit": _num(sgp.data_fit), "complexity_penalty": _num(sgp.complexity_penalty),
        "trace_term": _num(getattr(sgp, "trace_term", 0.0)),
        "constant_term": _num(sgp.constant_term),
        "likelihood_gradient": _vec(sgp.likelihood_gradient),
    }


def sparse_gp_from_json(j, context=None):
    from . import _C_flare as C
    kernels = [kernel_from_json(k) for k in j["kernels"]]
    sgp = C.SparseGP(kernels, j["energy_noise"], j["force_noise"], j["stress_noise"])
    sgp.hyperparameters = _read_vec(j["hyperparameters"])
    sgp.Kuu_jitter = float(j["Kuu_jitter"])
    sgp.sparse_descriptors = [cluster_from_json(c) for c in j["sparse_descriptors"]]
    sgp.sparse_indices = [list(map(list, x)) for x in (j.get("sparse_indices") or [])]
    sgp.Kuu_kernels = [_read_mat(x) for x in j["Kuu_kernels"]]
    sgp.Kuf_kernels = [_read_mat(x) for x in j["Kuf_kernels"]]
    sgp.Kuu, sgp.Kuf = _read_mat(j["Kuu"]), _read_mat(j["Kuf"])
    for name in ("Sigma", "Kuu_inverse", "R_inv", "L_inv"):
        setattr(sgp, name, _read_mat(j[name]))
    for name in ("alpha", "R_inv_diag", "L_diag", "noise_vector", "y", "likelihood_gradient"):
        setattr(sgp, name, _read_vec(j.get(name)))
    sgp.label_count = _read_vec(j.get("label_count")).astype(np.int64)
    if sgp.label_count.size == 0:
        sgp.label_count = np.zeros(1, np.int64)
    for name in ("n_energy_labels", "n_force_labels", "n_stress_labels", "n_labels",
                 "n_strucs"):
        setattr(sgp, name, int(j[name]))
    for name in ("log_marginal_likelihood", "data_fit", "complexity_penalty", "trace_term",
                 "constant_term"):
        v = j.get(name)
        setattr(sgp, name, float("nan") if v is None else float(v))
    sgp.training_structures = [structure_from_json(s, context)
                               for s in j.get("training_structures") or []]
    _restore_label_kinds(sgp)
    return sgp


def _restore_label_kinds(sgp):
    """The per-label kind vectors (e/f/s_noise_one and their inverses) and
    the force atoms of each structure are not in the file
    (sparse_gp.h:102-108); they follow from the label layout of each
    structure (energy, forces, stress) and the noise vector."""
    n = sgp.n_labels
    e1, f1, s1 = np.zeros(n), np.zeros(n), np.zeros(n)
    sgp.training_atom_indices = []
    for j, st in enumerate(sgp.training_structures):
        a, b = int(sgp.label_count[j]), int(sgp.label_count[j + 1])
        ne, ns = np.size(st.energy), np.size(st.stresses)
        nf = b - a - ne - ns
        e1[a:a + ne] = 1
        f1[a + ne:a + ne + nf] = 1
        s1[b - ns:b] = 1
        # All atoms' forces, or an unknown subset (left as None: adding
        # sparse environments then needs the subset and raises).
        sgp.training_atom_indices.append(list(range(st.noa)) if nf == 3 * st.noa else
                                         (None if nf else []))
    nv = sgp.noise_vector if sgp.noise_vector.size == n else np.ones(n)
    sgp.e_noise_one = e1 * nv * sgp.energy_noise ** 2
    sgp.f_noise_one = f1 * nv * sgp.force_noise ** 2
    sgp.s_noise_one = s1 * nv * sgp.stress_noise ** 2
    with np.errstate(divide="ignore"):
        sgp.inv_e_noise_one = np.where(e1 > 0, 1 / np.where(e1 > 0, sgp.e_noise_one, 1), 0)
        sgp.inv_f_noise_one = np.where(f1 > 0, 1 / np.where(f1 > 0, sgp.f_noise_one, 1), 0)
        sgp.inv_s_noise_one = np.where(s1 > 0, 1 / np.where(s1 > 0, sgp.s_noise_one, 1), 0)


def dump(obj, file_name):
    with open(file_name, "w") as f:
        json.dump(obj, f)


def load(file_name):
    with open(file_name) as f:
        return json.load(f)
