"""Drop-in for the prediction path of FLARE's ``_C_flare`` module
(src/ace_binding.cpp:28-263): ``Structure``, ``DescriptorValues``,
``ClusterDescriptor``, ``B2``, ``NormalizedDotProduct`` and ``SparseGP`` with
the same names, argument meaning and error behaviour, backed by the HIP
library libflare_mi through its C ABI.

Neighbour lists and B2 descriptors are built on the GPU when a
``Structure`` is constructed with a cutoff and descriptor calculators
(structure.cpp:27-43) and stay resident there; ``SparseGP.predict_*``
mutate the passed ``Structure`` (sparse_gp.cpp:643-757). Host copies of
the descriptor containers are made only when their fields are read.

Out of this module's scope (not on the prediction path): the dense
descriptor gradients ``descriptor_force_dervs`` / ``descriptor_force_dots``,
which the GPU path never forms, and the sparse-GP training updates that
need them (``add_training_structure``, likelihoods).
"""
import numpy as np

from . import _lib
from ._lib import CUTOFF_COSINE, CUTOFF_QUADRATIC, FORM_CHOSEN, FORM_FP64

_CUTOFFS = {"quadratic": CUTOFF_QUADRATIC, "cosine": CUTOFF_COSINE}
_EMPTY = 1e-8


class Descriptor:
    """Base of the descriptor calculators (descriptor.h:12-25)."""

    descriptor_name = ""

    def compute_struc(self, structure):
        raise NotImplementedError


class B2(Descriptor):
    """B2 power spectrum (descriptors/b2.h:12-53; b2.cpp:13-50).

    B2(radial_basis, cutoff_function, radial_hyps, cutoff_hyps,
       descriptor_settings[, cutoffs])
    """

    descriptor_name = "B2"

    def __init__(self, radial_basis, cutoff_function, radial_hyps, cutoff_hyps,
                 descriptor_settings, cutoffs=None):
        self._radial_basis = str(radial_basis)
        self._cutoff_function = str(cutoff_function)
        self._radial_hyps = [float(x) for x in radial_hyps]
        self._cutoff_hyps = [float(x) for x in cutoff_hyps]
        self._descriptor_settings = [int(x) for x in descriptor_settings]
        nos = self._descriptor_settings[0]
        if cutoffs is None:
            cutoffs = np.full((nos, nos), self._radial_hyps[1])
        self._cutoffs = np.array(cutoffs, dtype=np.float64).reshape(nos, nos)

    radial_basis = property(lambda self: self._radial_basis)
    cutoff_function = property(lambda self: self._cutoff_function)
    radial_hyps = property(lambda self: list(self._radial_hyps))
    cutoff_hyps = property(lambda self: list(self._cutoff_hyps))
    descriptor_settings = property(lambda self: list(self._descriptor_settings))
    cutoffs = property(lambda self: self._cutoffs.copy())

    @property
    def n_descriptors(self):
        nos, N, lmax = self._descriptor_settings
        nr = nos * N
        return nr * (nr + 1) // 2 * (lmax + 1)

    def _settings(self):
        # The reference silently leaves an unknown basis or cutoff function
        # unset and fails when it is called (radial.cpp:381-408,
        # cutoffs.cpp:87-103); here the error is raised up front.
        if self._radial_basis != "chebyshev":
            raise ValueError(f"radial basis {self._radial_basis!r} is not supported (chebyshev)")
        if self._cutoff_function not in _CUTOFFS:
            raise ValueError(f"cutoff function {self._cutoff_function!r} is not supported "
                             "(quadratic, cosine)")
        nos, N, lmax = self._descriptor_settings
        cut = np.ascontiguousarray(self._cutoffs.reshape(-1))
        s = _lib.B2Settings(nos, N, lmax, _CUTOFFS[self._cutoff_function],
                            float(self._radial_hyps[0]),
                            cut.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)))
        return s, cut

    def compute_struc(self, structure):
        """Descriptors of a structure whose neighbour list is built."""
        return structure._compute_b2(self)


class B3(B2):
    """B3 bispectrum (descriptors/b3.h; b3.cpp:13-50), same arguments as B2;
    l_max <= 4 (the reference's Wigner 3j table, wigner3j.cpp:1983-1985).
    Unlike the reference's constructor with a cutoff matrix (b3.cpp:37-56,
    which leaves the Wigner coefficients unset), both constructors work."""

    descriptor_name = "B3"
    _N_LS = [1, 5, 15, 34, 65]  # triangle-valid (l1, l2, l3), b3.cpp:200-211

    @property
    def n_descriptors(self):
        nos, N, lmax = self._descriptor_settings
        nr = nos * N
        return nr * (nr + 1) * (nr + 2) // 6 * self._N_LS[lmax]

    def _settings(self):
        if self._descriptor_settings[2] > 4:
            raise ValueError("B3 supports l_max <= 4")
        return super()._settings()


class DescriptorValues:
    """Per-structure descriptor container (descriptor.h:31-57), read lazily
    from the device. Fields are per species lists, in species-sorted order
    with atom order kept within a species (b2.cpp:104-176)."""

    def __init__(self, structure=None, calc=None, handle=None):
        self._structure = structure
        self._calc = calc
        self._cache = {}
        # The device structure holding this calculator's descriptors.
        self._handle = handle if handle is not None else (
            structure._h if structure is not None else None)
        if structure is not None:
            L = _lib.load()
            nos = calc.descriptor_settings[0]
            nd = _lib.C.c_int()
            ncl = np.zeros(nos, np.int32)
            nnb = np.zeros(nos, np.int32)
            _lib.check(L.fm_structure_b2_sizes(self._handle, _lib.C.byref(nd), _lib.ptr(ncl),
                                               _lib.ptr(nnb)))
            self.n_descriptors = int(nd.value)
            self.n_types = nos
            self.n_atoms = structure.noa
            self.volume = structure.volume
            self.n_clusters_by_type = [int(x) for x in ncl]
            self.n_neighbors_by_type = [int(x) for x in nnb]
            self.cumulative_type_count = [0]
            for x in self.n_clusters_by_type:
                self.cumulative_type_count.append(self.cumulative_type_count[-1] + x)
            self.n_clusters = int(ncl.sum())

    def _load(self):
        if self._cache or self._structure is None:
            return
        L = _lib.load()
        nd = self.n_descriptors
        out = {k: [] for k in ("descriptors", "descriptor_norms", "neighbor_counts",
                               "cumulative_neighbor_counts", "atom_indices",
                               "neighbor_indices", "neighbor_coordinates")}
        for s in range(self.n_types):
            ns, nb = self.n_clusters_by_type[s], self.n_neighbors_by_type[s]
            d = np.zeros((ns, nd))
            nrm = np.zeros(ns)
            nc = np.zeros(ns, np.int32)
            cnc = np.zeros(ns, np.int32)
            ai = np.zeros(ns, np.int32)
            ni = np.zeros(nb, np.int32)
            co = np.zeros((nb, 3))
            _lib.check(L.fm_structure_get_b2(self._handle, s, _lib.ptr(d), _lib.ptr(nrm),
                                             _lib.ptr(nc), _lib.ptr(cnc), _lib.ptr(ai),
                                             _lib.ptr(ni), _lib.ptr(co)))
            for k, v in zip(out, (d, nrm, nc, cnc, ai, ni, co)):
                out[k].append(v)
        out["cutoff_values"] = [np.ones(n) for n in self.n_clusters_by_type]
        self._cache = out

    def __getattr__(self, name):
        if name in ("descriptors", "descriptor_norms", "neighbor_counts",
                    "cumulative_neighbor_counts", "atom_indices", "neighbor_indices",
                    "neighbor_coordinates", "cutoff_values"):
            self._load()
            return self._cache[name]
        if name in ("descriptor_force_dervs", "descriptor_force_dots"):
            self._load_dervs()
            return self._cache[name]
        if name == "cutoff_dervs":  # zero for B2 (b2.cpp:138)
            return [np.zeros(3 * n) for n in self.n_neighbors_by_type]
        raise AttributeError(name)

    def _load_dervs(self):
        """Dense dB/dr (b2.cpp:181-242), built on the device on first use.
        Prediction never needs them (forces come from the adjoint); they
        serve the DTC/SOR kernels and callers that read them."""
        if "descriptor_force_dervs" in self._cache or self._structure is None:
            return
        self._load()
        L = _lib.load()
        dervs, dots = [], []
        for s in range(self.n_types):
            nb = self.n_neighbors_by_type[s]
            d = np.zeros((3 * nb, self.n_descriptors))
            f = np.zeros(3 * nb)
            _lib.check(L.fm_structure_get_b2_dervs(self._handle, s, _lib.ptr(d),
                                                   _lib.ptr(f)))
            dervs.append(d)
            dots.append(f)
        self._cache["descriptor_force_dervs"] = dervs
        self._cache["descriptor_force_dots"] = dots


class ClusterDescriptor:
    """Sparse-set container (descriptor.h:61-90): raw descriptor rows and
    norms per species, species-major (the row order of alpha and L_inv)."""

    def __init__(self, n_types=0, n_descriptors=0):
        self.n_types = n_types
        self.n_descriptors = n_descriptors
        self.descriptors = [np.zeros((0, n_descriptors)) for _ in range(n_types)]
        self.descriptor_norms = [np.zeros(0) for _ in range(n_types)]
        self.cutoff_values = [np.zeros(0) for _ in range(n_types)]

    @property
    def n_clusters_by_type(self):
        return [len(x) for x in self.descriptor_norms]

    @property
    def cumulative_type_count(self):
        c = [0]
        for x in self.n_clusters_by_type[:-1]:
            c.append(c[-1] + x)
        return c

    @property
    def n_clusters(self):
        return int(sum(self.n_clusters_by_type))

    def add_clusters_by_type(self, values, clusters):
        """descriptor.cpp:104-140."""
        if self.n_types == 0:
            self.__init__(values.n_types, values.n_descriptors)
        for s in range(self.n_types):
            idx = np.asarray(clusters[s], dtype=np.int64)
            if idx.size == 0:
                continue
            self.descriptors[s] = np.concatenate([self.descriptors[s], values.descriptors[s][idx]])
            self.descriptor_norms[s] = np.concatenate([self.descriptor_norms[s],
                                                       values.descriptor_norms[s][idx]])
            self.cutoff_values[s] = np.concatenate([self.cutoff_values[s],
                                                    values.cutoff_values[s][idx]])

    def add_all_clusters(self, values):
        """descriptor.cpp:142-176."""
        self.add_clusters_by_type(values, [range(n) for n in values.n_clusters_by_type])


class Kernel:
    kernel_name = ""


class NormalizedDotProduct(Kernel):
    """k = sigma^2 (x . y / |x||y|)^power (kernels/normalized_dot_product.cpp:15-26)."""

    kernel_name = "NormalizedDotProduct"

    def __init__(self, sigma, power):
        self._sigma = float(sigma)
        self.power = float(power)

    sigma = property(lambda self: self._sigma)

    @property
    def kernel_hyperparameters(self):
        return np.array([self._sigma])

    def _set_hyperparameters(self, hyps):
        """normalized_dot_product.cpp:625-629."""
        self._sigma = float(np.asarray(hyps).reshape(-1)[0])

    def envs_envs(self, envs1, envs2, hyps=None):
        """normalized_dot_product.cpp:28-89 on host FP64 (a sparse-set
        sized product, used to build Kuu)."""
        sig = float(self._sigma if hyps is None else np.asarray(hyps)[0])
        K = np.zeros((envs1.n_clusters, envs2.n_clusters))
        c1, c2 = envs1.cumulative_type_count, envs2.cumulative_type_count
        for s in range(envs1.n_types):
            a, b = envs1.descriptors[s], envs2.descriptors[s]
            if a.shape[0] == 0 or b.shape[0] == 0:
                continue
            na, nb = envs1.descriptor_norms[s], envs2.descriptor_norms[s]
            dot = a @ b.T
            with np.errstate(divide="ignore", invalid="ignore"):
                v = sig * sig * (dot / np.outer(na, nb)) ** self.power
            v[na < _EMPTY, :] = 0.0
            v[:, nb < _EMPTY] = 0.0
            K[c1[s]:c1[s] + a.shape[0], c2[s]:c2[s] + b.shape[0]] = v
        return K

    def envs_struc(self, envs, struc, hyps=None):
        """normalized_dot_product.cpp:117-214 on the GPU: the
        n_sparse x (1 + 3N + 6) kernel matrix between a sparse set and a
        structure's descriptors (DescriptorValues of a Structure)."""
        sig = float(self._sigma if hyps is None else np.asarray(hyps)[0])
        st = struc._structure
        m = _DeviceModel.from_sparse_set(st._ctx, sig, self.power, envs)
        K = np.zeros((envs.n_clusters, 1 + 3 * st.noa + 6))
        _lib.check(_lib.load().fm_envs_struc(m.h, struc._handle, _lib.ptr(K)))
        return K

    def self_kernel_struc(self, struc, hyps=None):
        """normalized_dot_product.cpp:450-586 on the GPU."""
        sig = float(self._sigma if hyps is None else np.asarray(hyps)[0])
        st = struc._structure
        empty = ClusterDescriptor(struc.n_types, struc.n_descriptors)
        m = _DeviceModel.from_sparse_set(st._ctx, sig, self.power, empty)
        out = np.zeros(1 + 3 * st.noa + 6)
        _lib.check(_lib.load().fm_self_kernel_struc(m.h, struc._handle, _lib.ptr(out)))
        return out


class Structure:
    """Periodic structure (structure.h:9-114). With a cutoff and descriptor
    calculators the neighbour list and descriptors are computed on the GPU
    at construction (structure.cpp:27-43).

    ``form`` selects the arithmetic of the descriptors and of later
    predictions: ``FORM_FP64`` (default, the drop-in) or ``FORM_CHOSEN``.
    """

    def __init__(self, cell, species, positions, cutoff=None, descriptor_calculators=None,
                 form=FORM_FP64, context=None):
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)
        self.species = [int(s) for s in species]
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.noa = len(self.species)
        if self.positions.shape[0] != self.noa:
            raise ValueError("positions and species disagree in length")
        self.cell_transpose = self.cell.T.copy()
        self.volume = abs(float(np.linalg.det(self.cell)))
        self.form = form
        self.energy = np.zeros(0)
        self.forces = np.zeros(0)
        self.stresses = np.zeros(0)
        self.mean_efs = np.zeros(0)
        self.variance_efs = np.zeros(0)
        self.local_uncertainties = []
        self.mean_contributions = []
        self.descriptors = []
        self.descriptor_calculators = list(descriptor_calculators or [])
        self._h = None
        self._hs = []  # device structure of each descriptor calculator (_hs[0] is _h)
        self._slot = 0
        self._ctx = context or _lib.default_context()
        self._nbr_cache = None
        self.cutoff = cutoff
        # The three-argument constructor (structure.cpp:7-25) builds no
        # neighbour list; the device structure is still created so that
        # wrapped_positions comes from the same code.
        self._build(cutoff if cutoff is not None else 1e-3)
        if cutoff is not None:
            self.compute_descriptors()

    def _create_device(self, cutoff):
        L = _lib.load()
        h = _lib.C.c_void_p()
        cell = _lib.f64(self.cell.reshape(-1))
        sp = _lib.i32(self.species)
        pos = _lib.f64(self.positions.reshape(-1))
        _lib.check(L.fm_structure_create(self._ctx.h, _lib.ptr(cell), self.noa, _lib.ptr(sp),
                                         _lib.ptr(pos), float(cutoff), _lib.C.byref(h)))
        return h

    def _build(self, cutoff):
        L = _lib.load()
        self._build_cutoff = cutoff
        h = self._create_device(cutoff)
        self._h = h
        self._hs = [h]
        sweep = _lib.C.c_int()
        nnb = _lib.C.c_int64()
        vol = _lib.C.c_double()
        ssc = _lib.C.c_double()
        _lib.check(L.fm_structure_info(h, _lib.C.byref(sweep), _lib.C.byref(nnb),
                                       _lib.C.byref(vol), _lib.C.byref(ssc)))
        self.sweep = int(sweep.value)
        self.n_neighbors = int(nnb.value)
        self.volume = float(vol.value)
        self.single_sweep_cutoff = float(ssc.value)

    def __del__(self):
        try:
            for h in self._hs:
                _lib.load().fm_structure_destroy(h)
        except Exception:
            pass

    def _neighbors(self):
        if self._nbr_cache is None:
            L = _lib.load()
            n, nnb = self.noa, self.n_neighbors
            w = np.zeros((n, 3))
            cnt = np.zeros(n, np.int32)
            cum = np.zeros(n + 1, np.int32)
            si = np.zeros(nnb, np.int32)
            ns = np.zeros(nnb, np.int32)
            rel = np.zeros((nnb, 4))
            _lib.check(L.fm_structure_get_neighbors(self._h, _lib.ptr(w), _lib.ptr(cnt),
                                                    _lib.ptr(cum), _lib.ptr(si), _lib.ptr(ns),
                                                    _lib.ptr(rel)))
            self._nbr_cache = dict(wrapped_positions=w, neighbor_count=cnt,
                                   cumulative_neighbor_count=cum, structure_indices=si,
                                   neighbor_species=ns, relative_positions=rel)
        return self._nbr_cache

    wrapped_positions = property(lambda self: self._neighbors()["wrapped_positions"])
    neighbor_count = property(lambda self: self._neighbors()["neighbor_count"])
    cumulative_neighbor_count = property(
        lambda self: self._neighbors()["cumulative_neighbor_count"])
    structure_indices = property(lambda self: self._neighbors()["structure_indices"])
    neighbor_species = property(lambda self: self._neighbors()["neighbor_species"])
    relative_positions = property(lambda self: self._neighbors()["relative_positions"])

    def wrap_positions(self):
        return self.wrapped_positions.copy()

    @staticmethod
    def to_json(file_name, struc):
        """structure.cpp to_json: the reference's JSON layout."""
        from . import serialization
        serialization.dump(serialization.structure_to_json(struc), file_name)

    @staticmethod
    def from_json(file_name):
        from . import serialization
        return serialization.structure_from_json(serialization.load(file_name))

    def compute_descriptors(self):
        """structure.cpp:45-50. Each calculator's descriptors live in a device
        structure of their own (the first is the one built at construction;
        the others repeat its neighbour list)."""
        out = []
        for k, calc in enumerate(self.descriptor_calculators):
            self._slot = k
            out.append(calc.compute_struc(self))
        self._slot = 0
        self.descriptors = out

    def _compute_b2(self, calc):
        while len(self._hs) <= self._slot:
            self._hs.append(self._create_device(self._build_cutoff))
        h = self._hs[self._slot]
        settings, keep = calc._settings()
        fn = (_lib.load().fm_structure_compute_b3 if isinstance(calc, B3)
              else _lib.load().fm_structure_compute_b2)
        _lib.check(fn(h, _lib.C.byref(settings), int(self.form)))
        del keep
        return DescriptorValues(self, calc, h)


class SparseGP:
    """Sparse GP (bffs/sparse_gp.h, sparse_gp.cpp): the sparse set, the
    training labels and the kernel matrices that train it, the trained state
    (alpha, L_inv, Kuu_inverse, R_inv, Sigma) and the predictions.

    Kernel matrices between sparse environments and training structures
    (Kuf columns, normalized_dot_product.cpp:117-214) are built on the GPU;
    the sparse-set-sized algebra (Kuu, the Cholesky and Householder QR of
    update_matrices_QR, likelihoods) is FP64 on the host, as the reference
    keeps it FP64. Predictions run on the GPU in the Structure's form.
    """

    def __init__(self, kernels=None, energy_noise=1.0, force_noise=1.0, stress_noise=1.0):
        # sparse_gp.cpp:14-56.
        self.kernels = list(kernels or [])
        self.n_kernels = len(self.kernels)
        self.energy_noise = float(energy_noise)
        self.force_noise = float(force_noise)
        self.stress_noise = float(stress_noise)
        self.Kuu_jitter = 1e-8  # sparse_gp.cpp:19
        h = [np.asarray(k.kernel_hyperparameters, dtype=np.float64) for k in self.kernels]
        self.hyperparameters = np.concatenate(
            h + [np.array([self.energy_noise, self.force_noise, self.stress_noise])])
        self.sparse_descriptors = [ClusterDescriptor() for _ in self.kernels]
        self.sparse_indices = [[] for _ in self.kernels]
        self.Kuu_kernels = [np.zeros((0, 0)) for _ in self.kernels]
        self.Kuf_kernels = [np.zeros((0, 0)) for _ in self.kernels]
        self.Kuu = np.zeros((0, 0))
        self.Kuf = np.zeros((0, 0))
        self.training_structures = []
        self.training_atom_indices = []
        self.n_energy_labels = self.n_force_labels = self.n_stress_labels = 0
        self.n_labels = self.n_strucs = 0
        self.label_count = np.zeros(1, np.int64)
        self.y = np.zeros(0)
        self.noise_vector = np.zeros(0)
        self.e_noise_one = np.zeros(0)
        self.f_noise_one = np.zeros(0)
        self.s_noise_one = np.zeros(0)
        self.inv_e_noise_one = np.zeros(0)
        self.inv_f_noise_one = np.zeros(0)
        self.inv_s_noise_one = np.zeros(0)
        self.alpha = np.zeros(0)
        self.L_inv = np.zeros((0, 0))
        self.L_diag = np.zeros(0)
        self.Kuu_inverse = np.zeros((0, 0))
        self.Sigma = np.zeros((0, 0))
        self.R_inv = np.zeros((0, 0))
        self.R_inv_diag = np.zeros(0)
        self.log_marginal_likelihood = self.data_fit = 0.0
        self.complexity_penalty = self.constant_term = 0.0
        self.likelihood_gradient = np.zeros(0)
        self.Kuf_e_noise_Kfu, self.Kuf_f_noise_Kfu, self.Kuf_s_noise_Kfu = [], [], []
        self._models = {}

    @property
    def n_sparse(self):
        return int(sum(c.n_clusters for c in self.sparse_descriptors))

    def _invalidate(self):
        self._models = {}

    # Sparse set (sparse_gp.cpp:58-341) ---------------------------------
    def _initialize_sparse_descriptors(self, structure):
        """sparse_gp.cpp:58-71."""
        for k, cd in enumerate(self.sparse_descriptors):
            if cd.n_types == 0 and k < len(structure.descriptors):
                v = structure.descriptors[k]
                cd.__init__(v.n_types, v.n_descriptors)

    def _add_clusters(self, structure, by_type):
        """Kuu and Kuf rows of the new clusters (update_Kuu / update_Kuf /
        stack, sparse_gp.cpp:343-472), then the clusters themselves. by_type[k]
        lists, per species, cluster rows of structure.descriptors[k]."""
        new = []
        for k in range(self.n_kernels):
            cl = ClusterDescriptor()
            cl.add_clusters_by_type(structure.descriptors[k], by_type[k])
            new.append(cl)
        self._update_Kuu(new)
        self._update_Kuf(new)
        self._stack()
        for k in range(self.n_kernels):
            self.sparse_descriptors[k].add_clusters_by_type(structure.descriptors[k],
                                                            by_type[k])
        self._invalidate()

    @staticmethod
    def _type_major(old, new):
        """Row order of (old rows; new rows) once the new clusters join the
        sparse set: per species, the old rows then the new ones
        (sparse_gp.cpp:343-402)."""
        order = []
        M0 = old.n_clusters
        co, cn = old.cumulative_type_count, new.cumulative_type_count
        for s in range(new.n_types):
            no = old.n_clusters_by_type[s] if old.n_types else 0
            if no:
                order.extend(range(co[s], co[s] + no))
            order.extend(range(M0 + cn[s], M0 + cn[s] + new.n_clusters_by_type[s]))
        return np.asarray(order, dtype=np.int64)

    def _update_Kuu(self, new):
        """sparse_gp.cpp:343-402: the previous block, its kernel with the new
        clusters and the new clusters' own block, laid out species-major."""
        for k, kern in enumerate(self.kernels):
            old = self.sparse_descriptors[k]
            self_block = kern.envs_envs(new[k], new[k])
            if old.n_clusters:
                prev = kern.envs_envs(old, new[k])
                full = np.block([[self.Kuu_kernels[k], prev], [prev.T, self_block]])
            else:
                full = self_block
            order = self._type_major(old, new[k])
            self.Kuu_kernels[k] = full[np.ix_(order, order)]

    def _label_columns(self, j, K):
        """Columns of structure j's labels in an envs_struc matrix K
        (energy, forces of the training atoms, stress; sparse_gp.cpp:423-456)."""
        st = self.training_structures[j]
        if self.training_atom_indices[j] is None:
            raise RuntimeError(f"the force atoms of training structure {j} are not known "
                               "(a model file does not store them)")
        cols = []
        if np.size(st.energy):
            cols.append(K[:, :1])
        if np.size(st.forces):
            for a in self.training_atom_indices[j]:
                cols.append(K[:, 1 + 3 * a:4 + 3 * a])
        if np.size(st.stresses):
            cols.append(K[:, 1 + 3 * st.noa:])
        return np.concatenate(cols, axis=1) if cols else K[:, :0]

    def _update_Kuf(self, new):
        """sparse_gp.cpp:404-472: kernels between the new clusters and every
        training structure, on the GPU."""
        for k, kern in enumerate(self.kernels):
            old = self.sparse_descriptors[k]
            rows = np.zeros((new[k].n_clusters, self.n_labels))
            if new[k].n_clusters:
                for j, st in enumerate(self.training_structures):
                    K = kern.envs_struc(new[k], st.descriptors[k])
                    a, b = self.label_count[j], self.label_count[j + 1]
                    rows[:, a:b] = self._label_columns(j, K)
            prev = self.Kuf_kernels[k] if old.n_clusters else np.zeros((0, self.n_labels))
            full = np.concatenate([prev, rows], axis=0)
            self.Kuf_kernels[k] = full[self._type_major(old, new[k])]

    def _stack(self):
        """stack_Kuu / stack_Kuf (sparse_gp.cpp:585-605)."""
        M = sum(b.shape[0] for b in self.Kuu_kernels)
        self.Kuu = np.zeros((M, M))
        self.Kuf = np.zeros((M, self.n_labels))
        o = 0
        for ku, kf in zip(self.Kuu_kernels, self.Kuf_kernels):
            m = ku.shape[0]
            self.Kuu[o:o + m, o:o + m] = ku
            if m:
                self.Kuf[o:o + m, :] = kf
            o += m

    def _atom_indices_of_clusters(self, vals, clusters):
        ai = np.concatenate([np.asarray(a) for a in vals.atom_indices]) if vals.n_clusters else []
        return [int(ai[c]) for c in clusters]

    @staticmethod
    def _clusters_by_type(vals, clusters):
        """ClusterDescriptor::add_clusters (descriptor.cpp:77-100): cluster
        indices (species-sorted order) to per-species rows, order kept."""
        cum = vals.cumulative_type_count
        out = [[] for _ in range(vals.n_types)]
        for c in clusters:
            for s in range(vals.n_types):
                if cum[s] <= c < cum[s + 1]:
                    out[s].append(int(c) - cum[s])
                    break
        return out

    def add_specific_environments(self, structure, atoms):
        """sparse_gp.cpp:140-186: per species, the clusters whose atom is in
        ``atoms``, in the structure's species-sorted order (an atom listed
        twice is added twice)."""
        self._initialize_sparse_descriptors(structure)
        atoms = [int(a) for a in atoms]
        by_kernel = []
        for k in range(self.n_kernels):
            self.sparse_indices[k].append(list(atoms))
            vals = structure.descriptors[k]
            by_type = []
            for s in range(vals.n_types):
                rows = []
                for r, a in enumerate(np.asarray(vals.atom_indices[s])):
                    rows.extend([r] * atoms.count(int(a)))
                by_type.append(rows)
            by_kernel.append(by_type)
        self._add_clusters(structure, by_kernel)

    def add_all_environments(self, structure):
        """sparse_gp.cpp:309-341."""
        self._initialize_sparse_descriptors(structure)
        by_kernel = [[list(range(n)) for n in structure.descriptors[k].n_clusters_by_type]
                     for k in range(self.n_kernels)]
        self._add_clusters(structure, by_kernel)
        for k in range(self.n_kernels):
            self.sparse_indices[k].append(list(range(structure.noa)))

    def add_random_environments(self, structure, n_added, seed=None):
        """sparse_gp.cpp:259-307. The reference draws with std::random_shuffle
        (unseeded); here numpy's generator with an optional seed draws the
        same kind of sample: n_added[k] clusters without replacement."""
        self._initialize_sparse_descriptors(structure)
        rng = np.random.default_rng(seed)
        by_kernel, picks = [], []
        for k in range(self.n_kernels):
            vals = structure.descriptors[k]
            n = min(int(n_added[k]), vals.n_clusters)
            pick = [int(c) for c in rng.permutation(vals.n_clusters)[:n]]
            picks.append(pick)
            by_kernel.append(self._clusters_by_type(vals, pick))
        self._add_clusters(structure, by_kernel)
        for k in range(self.n_kernels):
            self.sparse_indices[k].append(
                self._atom_indices_of_clusters(structure.descriptors[k], picks[k]))

    def add_uncertain_environments(self, structure, n_added):
        """sparse_gp.cpp:188-257: the n_added[k] clusters of largest local
        variance (stable sort, decreasing)."""
        if len(n_added) != self.n_kernels:
            raise RuntimeError("n_added must have the same size as the number of kernels")
        self._initialize_sparse_descriptors(structure)
        var = self.compute_cluster_uncertainties(structure)
        by_kernel, picks = [], []
        for k in range(self.n_kernels):
            order = sorted(range(len(var[k])), key=lambda i: -var[k][i])
            pick = order[:min(int(n_added[k]), len(order))]
            picks.append(pick)
            by_kernel.append(self._clusters_by_type(structure.descriptors[k], pick))
        self._add_clusters(structure, by_kernel)
        for k in range(self.n_kernels):
            self.sparse_indices[k].append(
                self._atom_indices_of_clusters(structure.descriptors[k], picks[k]))

    # Training labels (sparse_gp.cpp:474-583) ------------------------------
    def add_training_structure(self, structure, atom_indices=(-1,), rel_e_noise=1.0,
                               rel_f_noise=1.0, rel_s_noise=1.0):
        """Adds the structure's energy, forces (of ``atom_indices``, all atoms
        for [-1]) and stress labels; their Kuf columns come from envs_struc
        on the GPU (FP64 descriptors)."""
        self._initialize_sparse_descriptors(structure)
        if structure.form != FORM_FP64:
            raise ValueError("training structures need FP64 descriptors")
        n_atoms = structure.noa
        energy = np.asarray(structure.energy, dtype=np.float64).reshape(-1)
        forces = np.asarray(structure.forces, dtype=np.float64).reshape(-1)
        stresses = np.asarray(structure.stresses, dtype=np.float64).reshape(-1)
        atom_indices = [int(a) for a in atom_indices]
        if forces.size == 0:
            atoms = []
        elif atom_indices[0] == -1:
            atoms = list(range(n_atoms))
        else:
            atoms = atom_indices
        if forces.size and forces.size != 3 * n_atoms:
            raise ValueError("forces must hold 3 values per atom")
        self.training_atom_indices.append(atoms)
        n_e, n_f, n_s = energy.size, 3 * len(atoms), stresses.size
        n_new = n_e + n_f + n_s
        y = np.concatenate([energy] + [forces[3 * a:3 * a + 3] for a in atoms] + [stresses])
        e1 = np.r_[np.full(n_e, 1 / rel_e_noise ** 2), np.zeros(n_f + n_s)]
        f1 = np.r_[np.zeros(n_e), np.full(n_f, 1 / rel_f_noise ** 2), np.zeros(n_s)]
        s1 = np.r_[np.zeros(n_e + n_f), np.full(n_s, 1 / rel_s_noise ** 2)]
        ie = np.r_[np.full(n_e, rel_e_noise ** 2), np.zeros(n_f + n_s)]
        iff = np.r_[np.zeros(n_e), np.full(n_f, rel_f_noise ** 2), np.zeros(n_s)]
        is_ = np.r_[np.zeros(n_e + n_f), np.full(n_s, rel_s_noise ** 2)]
        noise = (e1 / self.energy_noise ** 2 + f1 / self.force_noise ** 2
                 + s1 / self.stress_noise ** 2)
        self.y = np.r_[self.y, y]
        self.noise_vector = np.r_[self.noise_vector, noise]
        self.e_noise_one = np.r_[self.e_noise_one, e1]
        self.f_noise_one = np.r_[self.f_noise_one, f1]
        self.s_noise_one = np.r_[self.s_noise_one, s1]
        self.inv_e_noise_one = np.r_[self.inv_e_noise_one, ie]
        self.inv_f_noise_one = np.r_[self.inv_f_noise_one, iff]
        self.inv_s_noise_one = np.r_[self.inv_s_noise_one, is_]
        self.label_count = np.r_[self.label_count, self.n_labels + n_new]
        self.training_structures.append(structure)
        self.n_strucs += 1
        j = self.n_strucs - 1
        for k, kern in enumerate(self.kernels):
            cd = self.sparse_descriptors[k]
            cols = np.zeros((cd.n_clusters, n_new))
            if cd.n_clusters:
                cols = self._label_columns(j, kern.envs_struc(cd, structure.descriptors[k]))
            prev = self.Kuf_kernels[k] if self.Kuf_kernels[k].size else np.zeros(
                (cd.n_clusters, self.n_labels))
            self.Kuf_kernels[k] = np.concatenate([prev, cols], axis=1)
        self.n_energy_labels += n_e
        self.n_force_labels += n_f
        self.n_stress_labels += n_s
        self.n_labels += n_new
        self._stack()
        self._invalidate()

    # Trained state (sparse_gp.cpp:607-641) --------------------------------
    def update_matrices_QR(self):
        """L = chol(Kuu + jitter I), L_inv, Kuu^-1 = L_inv^T L_inv; the
        Householder QR of A = [Lambda^1/2 Kuf^T; L^T]; R_inv = R^-1,
        alpha = R_inv (Q^T b)[:M] with b = [Lambda^1/2 y; 0], Sigma =
        R_inv R_inv^T. FP64 on the host (LAPACK geqrf: the Householder
        convention of Eigen's HouseholderQR)."""
        from scipy.linalg import solve_triangular
        M = self.Kuu.shape[0]
        eye = np.eye(M)
        L = np.linalg.cholesky(self.Kuu + self.Kuu_jitter * eye)
        self.L_inv = solve_triangular(L, eye, lower=True)
        self.L_diag = np.diag(self.L_inv).copy()
        self.Kuu_inverse = self.L_inv.T @ self.L_inv
        ns = np.sqrt(self.noise_vector)
        A = np.vstack([ns[:, None] * self.Kuf.T, L.T])
        b = np.r_[ns * self.y, np.zeros(M)]
        Q, R = np.linalg.qr(A, mode="reduced")
        self.R_inv = solve_triangular(R, eye, lower=False)
        self.R_inv_diag = np.diag(self.R_inv).copy()
        self.alpha = self.R_inv @ (Q.T @ b)
        self.Sigma = self.R_inv @ self.R_inv.T
        self._invalidate()

    def update_L_inverse(self):
        """L_inv = chol(Kuu + jitter I)^-1 and Kuu^-1 = L_inv^T L_inv
        (sparse_gp.cpp:610-618) alone, for a model whose alpha is set
        directly (no training labels)."""
        from scipy.linalg import solve_triangular
        M = self.Kuu.shape[0]
        L = np.linalg.cholesky(self.Kuu + self.Kuu_jitter * np.eye(M))
        self.L_inv = solve_triangular(L, np.eye(M), lower=True)
        self.L_diag = np.diag(self.L_inv).copy()
        self.Kuu_inverse = self.L_inv.T @ self.L_inv
        self._invalidate()

    def set_alpha(self, alpha):
        alpha = np.asarray(alpha, dtype=np.float64).reshape(-1)
        if alpha.size != self.n_sparse:
            raise ValueError("alpha must have one entry per sparse environment")
        self.alpha = alpha.copy()
        self._invalidate()

    def set_state(self, alpha, L_inv=None, Sigma=None, Kuu_inverse=None, R_inv=None):
        """Sets a trained state directly (alpha, L_inv, and R_inv or Sigma;
        Sigma = R_inv R_inv^T, a Sigma given alone is factored by Cholesky).
        update_matrices_QR produces the same state from training labels."""
        self.set_alpha(alpha)
        if L_inv is not None:
            self.L_inv = np.array(L_inv, dtype=np.float64)
            self.L_diag = np.diag(self.L_inv).copy()
            self.Kuu_inverse = (self.L_inv.T @ self.L_inv if Kuu_inverse is None
                                else np.array(Kuu_inverse, dtype=np.float64))
        if R_inv is not None:
            self.R_inv = np.array(R_inv, dtype=np.float64)
            self.Sigma = self.R_inv @ self.R_inv.T if Sigma is None else np.array(
                Sigma, dtype=np.float64)
        elif Sigma is not None:
            self.Sigma = np.array(Sigma, dtype=np.float64)
            self.R_inv = np.linalg.cholesky(self.Sigma)
        self._invalidate()

    def set_hyperparameters(self, hyps):
        """sparse_gp.cpp:1166-1198: kernel hyperparameters rescale Kuu and Kuf
        (both are sigma^2 times a sigma-free matrix for these kernels), the
        noises reset the noise vector, then update_matrices_QR."""
        hyps = np.asarray(hyps, dtype=np.float64).reshape(-1)
        h = 0
        for k, kern in enumerate(self.kernels):
            nh = len(kern.kernel_hyperparameters)
            new = hyps[h:h + nh]
            scale = new[0] ** 2 / kern.sigma ** 2
            self.Kuu_kernels[k] = self.Kuu_kernels[k] * scale
            self.Kuf_kernels[k] = self.Kuf_kernels[k] * scale
            kern._set_hyperparameters(new)
            h += nh
        self._stack()
        self.hyperparameters = hyps.copy()
        self.energy_noise, self.force_noise, self.stress_noise = (float(x) for x in hyps[h:h + 3])
        self.noise_vector = (self.e_noise_one / self.energy_noise ** 2
                             + self.f_noise_one / self.force_noise ** 2
                             + self.s_noise_one / self.stress_noise ** 2)
        self.update_matrices_QR()

    # Likelihoods (sparse_gp.cpp:759-1164) ----------------------------------
    def compute_likelihood_stable(self):
        """sparse_gp.cpp:759-787 (the reference also prints two lines of
        intermediate terms; this does not)."""
        r = self.y - self.Kuf.T @ self.alpha
        self.data_fit = -0.5 * float(self.y @ (self.noise_vector * r))
        self.constant_term = -0.5 * self.n_labels * np.log(2 * np.pi)
        noise_det = float(np.sum(np.log(self.noise_vector)))
        kuu_inv_det = -2 * float(np.sum(np.log(np.abs(self.L_diag))))
        sigma_inv_det = 2 * float(np.sum(np.log(np.abs(self.R_inv_diag))))
        self.complexity_penalty = 0.5 * (noise_det + kuu_inv_det + sigma_inv_det)
        self.log_marginal_likelihood = self.complexity_penalty + self.data_fit + self.constant_term

    def precompute_KnK(self):
        """sparse_gp.cpp:874-897: Kuf_i N_x Kuf_j^T / (sigma_i sigma_j)^2 per
        noise kind x."""
        self.Kuf_e_noise_Kfu, self.Kuf_f_noise_Kfu, self.Kuf_s_noise_Kfu = [], [], []
        for ki, kf_i in zip(self.kernels, self.Kuf_kernels):
            for kj, kf_j in zip(self.kernels, self.Kuf_kernels):
                s4 = ki.sigma ** 2 * kj.sigma ** 2
                self.Kuf_e_noise_Kfu.append((kf_i * self.e_noise_one) @ kf_j.T / s4)
                self.Kuf_f_noise_Kfu.append((kf_i * self.f_noise_one) @ kf_j.T / s4)
                self.Kuf_s_noise_Kfu.append((kf_i * self.s_noise_one) @ kf_j.T / s4)

    def _sizes(self):
        return [b.shape[0] for b in self.Kuu_kernels]

    def _KnK(self, precomputed):
        """compute_KnK (sparse_gp.cpp:899-935)."""
        if not precomputed:
            return tuple((self.Kuf * one) @ self.Kuf.T
                         for one in (self.e_noise_one, self.f_noise_one, self.s_noise_one))
        M = self.n_sparse
        out = [np.zeros((M, M)) for _ in range(3)]
        sz = self._sizes()
        off = np.r_[0, np.cumsum(sz)]
        c = 0
        for i, ki in enumerate(self.kernels):
            for j, kj in enumerate(self.kernels):
                s4 = ki.sigma ** 2 * kj.sigma ** 2
                for o, pre in zip(out, (self.Kuf_e_noise_Kfu, self.Kuf_f_noise_Kfu,
                                        self.Kuf_s_noise_Kfu)):
                    o[off[i]:off[i + 1], off[j]:off[j + 1]] += pre[c] * s4
                c += 1
        return tuple(out)

    def _dKnK(self, i):
        """compute_dKnK (sparse_gp.cpp:937-1012)."""
        M = self.n_sparse
        out = np.zeros((M, M))
        off = np.r_[0, np.cumsum(self._sizes())]
        si = self.kernels[i].sigma
        for j, kj in enumerate(self.kernels):
            c = i * self.n_kernels + j
            s3 = 2 * si * kj.sigma ** 2
            blk = (self.Kuf_e_noise_Kfu[c] * s3 / self.energy_noise ** 2
                   + self.Kuf_f_noise_Kfu[c] * s3 / self.force_noise ** 2
                   + self.Kuf_s_noise_Kfu[c] * s3 / self.stress_noise ** 2)
            out[off[i]:off[i + 1], off[j]:off[j + 1]] += blk
        return out

    def compute_likelihood_gradient_stable(self, precomputed_KnK=False):
        """sparse_gp.cpp:789-872: the likelihood of the current state and
        its gradient with respect to every hyperparameter."""
        r = self.y - self.Kuf.T @ self.alpha
        self.compute_likelihood_stable()
        M = self.n_sparse
        nh = len(self.hyperparameters)
        grad = np.zeros(nh)
        off = np.r_[0, np.cumsum(self._sizes())]
        h = 0
        for i, kern in enumerate(self.kernels):
            n_hyps = len(kern.kernel_hyperparameters)
            cur = self.hyperparameters[h:h + n_hyps]
            ku = self.Kuu_kernels[i]
            ku0 = ku * (cur[0] ** 2 / kern.sigma ** 2)
            ku1 = ku * (2 * cur[0] / kern.sigma ** 2)
            dKuu = np.zeros((M, M))
            dKuu[off[i]:off[i + 1], off[i]:off[i + 1]] = ku1
            if precomputed_KnK:
                dKnK = self._dKnK(i)
            else:
                dKuf = np.zeros((M, self.n_labels))
                dKuf[off[i]:off[i + 1]] = self.Kuf_kernels[i] * (2 * cur[0] / kern.sigma ** 2)
                dKnK = (dKuf * self.noise_vector) @ self.Kuf.T
            Pi = dKnK + dKnK.T + dKuu
            comp = 0.5 * np.trace(np.linalg.inv(ku0) @ ku1) - 0.5 * np.trace(Pi @ self.Sigma)
            if precomputed_KnK:
                dKuf = np.zeros((M, self.n_labels))
                dKuf[off[i]:off[i + 1]] = self.Kuf[off[i]:off[i + 1]]
                dK_alpha = (2.0 / cur[0]) * dKuf.T @ self.alpha
            else:
                dK_alpha = dKuf.T @ self.alpha
            fit = float(dK_alpha @ (self.noise_vector * r)) - 0.5 * float(
                self.alpha @ dKuu @ self.alpha)
            grad[h] = comp + fit
            h += n_hyps
        KnK_e, KnK_f, KnK_s = self._KnK(precomputed_KnK)
        for q, (n_lab, sig, KnK, one) in enumerate((
                (self.n_energy_labels, self.energy_noise, KnK_e, self.e_noise_one),
                (self.n_force_labels, self.force_noise, KnK_f, self.f_noise_one),
                (self.n_stress_labels, self.stress_noise, KnK_s, self.s_noise_one))):
            comp = -n_lab / sig + np.trace(KnK @ self.Sigma) / sig ** 3
            fit = float(r @ (one * r)) / sig ** 3
            grad[h + q] = comp + fit
        self.likelihood_gradient = grad
        return self.log_marginal_likelihood

    def compute_likelihood(self):
        """sparse_gp.cpp:1014-1045: the direct O(n_labels^3) likelihood
        through a column-pivoted QR of Qff + Lambda^-1."""
        if self.n_labels == 0:
            return
        from scipy.linalg import qr, solve_triangular
        Q = self.Kuf.T @ self.Kuu_inverse @ self.Kuf + np.diag(1 / self.noise_vector)
        q, R, piv = qr(Q, pivoting=True)
        z = solve_triangular(R, q.T @ self.y)
        x = np.empty_like(z)
        x[piv] = z
        self.complexity_penalty = -0.5 * float(np.sum(np.log(np.abs(np.diag(R)))))
        self.data_fit = -0.5 * float(self.y @ x)
        self.constant_term = -0.5 * self.n_labels * np.log(2 * np.pi)
        self.log_marginal_likelihood = self.complexity_penalty + self.data_fit + self.constant_term

    def compute_likelihood_gradient(self, hyperparameters):
        """sparse_gp.cpp:1047-1164: likelihood and gradient at the given
        hyperparameters through the dense n_labels x n_labels Qff."""
        from scipy.linalg import lu_factor, lu_solve
        hyps = np.asarray(hyperparameters, dtype=np.float64)
        M, n = self.n_sparse, self.n_labels
        Kuu = np.zeros((M, M))
        Kuf = np.zeros((M, n))
        dKuu, dKuf = [], []
        off = np.r_[0, np.cumsum(self._sizes())]
        h = 0
        for i, kern in enumerate(self.kernels):
            nh = len(kern.kernel_hyperparameters)
            c = hyps[h:h + nh]
            sl = slice(off[i], off[i + 1])
            Kuu[sl, sl] = self.Kuu_kernels[i] * (c[0] ** 2 / kern.sigma ** 2)
            Kuf[sl] = self.Kuf_kernels[i] * (c[0] ** 2 / kern.sigma ** 2)
            du, df = np.zeros((M, M)), np.zeros((M, n))
            du[sl, sl] = self.Kuu_kernels[i] * (2 * c[0] / kern.sigma ** 2)
            df[sl] = self.Kuf_kernels[i] * (2 * c[0] / kern.sigma ** 2)
            dKuu.append(du)
            dKuf.append(df)
            h += nh
        Kinv = np.linalg.inv(Kuu + self.Kuu_jitter * np.eye(M))
        se, sf, ss = hyps[h:h + 3]
        noise = se ** 2 * self.inv_e_noise_one + sf ** 2 * self.inv_f_noise_one \
            + ss ** 2 * self.inv_s_noise_one
        Qff = Kuf.T @ Kinv @ Kuf + np.diag(noise)
        grads = [dF.T @ Kinv @ Kuf - Kuf.T @ Kinv @ dU @ Kinv @ Kuf + Kuf.T @ Kinv @ dF
                 for dU, dF in zip(dKuu, dKuf)]
        grads += [np.diag(2 * se * self.inv_e_noise_one), np.diag(2 * sf * self.inv_f_noise_one),
                  np.diag(2 * ss * self.inv_s_noise_one)]
        lu = lu_factor(Qff)
        Qinv = lu_solve(lu, np.eye(n))
        self.complexity_penalty = -0.5 * float(np.sum(np.log(np.abs(np.diag(lu[0])))))
        Qy = Qinv @ self.y
        self.data_fit = -0.5 * float(self.y @ Qy)
        self.constant_term = -n * np.log(2 * np.pi) / 2
        self.log_marginal_likelihood = self.complexity_penalty + self.data_fit + self.constant_term
        g = np.zeros(len(hyps))
        for q, dQ in enumerate(grads):
            QdQ = Qinv @ dQ
            g[q] = (-np.trace(QdQ) + float(self.y @ QdQ @ Qy)) / 2
        self.likelihood_gradient = g
        return self.log_marginal_likelihood

    # Device models ---------------------------------------------------------
    def _model(self, k, ctx):
        # Each cached model holds its context, so an id cannot be reused by
        # another context while its entry exists.
        key = (k, id(ctx))
        m = self._models.get(key)
        if m is None or m.ctx is not ctx:
            m = _DeviceModel(self, k, ctx)
            self._models[key] = m
        return m

    # Predictions -----------------------------------------------------------
    def _check_state(self):
        if self.L_inv.shape[0] != self.n_sparse:
            raise RuntimeError("L_inv must be up to date to make predictions. Please call "
                               "update_matrices_QR and try again.")

    def _run(self, structure, flags):
        """Kernel k acts on descriptor k of the structure (sparse_gp.cpp:
        643-665, 731-757); means add up, local variances are per kernel."""
        self._check_state()
        if len(structure.descriptors) < self.n_kernels:
            raise ValueError("the structure has fewer descriptor calculators than the model "
                             "has kernels")
        n = structure.noa
        mean = np.zeros(1 + 3 * n + 6)
        local = []
        for k in range(self.n_kernels):
            m = self._model(k, structure._ctx)
            out = np.zeros(1 + 3 * n + 6)
            var = np.zeros(n) if flags & _lib.PREDICT_LOCAL else None
            _lib.check(_lib.load().fm_predict(m.h, structure.descriptors[k]._handle, flags,
                                              int(structure.form), _lib.ptr(out),
                                              _lib.ptr(var), None))
            mean += out
            if var is not None:
                local.append(var)
        return mean, local

    def predict_mean(self, structure):
        """sparse_gp.cpp:643-665."""
        structure.mean_efs, _ = self._run(structure, _lib.PREDICT_MEAN)

    def predict_local_uncertainties(self, structure):
        """sparse_gp.cpp:731-757."""
        mean, local = self._run(structure, _lib.PREDICT_MEAN | _lib.PREDICT_LOCAL)
        structure.mean_efs = mean
        structure.local_uncertainties = local

    def compute_cluster_uncertainties(self, structure):
        """sparse_gp.cpp:93-139: per kernel, the local variances of the
        structure's clusters in species-sorted order."""
        if self.L_inv.shape[0] != self.Kuu.shape[0] or self.L_inv.shape[0] != self.n_sparse:
            raise RuntimeError("L_inv must be up to date to evaluate uncertainties. Please "
                               "call update_matrices_QR and try again.")
        _, local = self._run(structure, _lib.PREDICT_MEAN | _lib.PREDICT_LOCAL)
        return local

    def _dense(self, structure, kind):
        """DTC / SOR: mean from every kernel; variances K_self - diag(K^T
        Kuu^-1 K) + diag(K^T Sigma K) (DTC) or diag(K^T R_inv R_inv^T K)
        (SOR) with K stacked over kernels (sparse_gp.cpp:667-729). A single
        kernel runs whole on the GPU; with several, each kernel's matrix K_k
        and self kernel come from the GPU and the stacked products are FP64
        on the host."""
        self._check_state()
        if self.R_inv.shape != (self.n_sparse, self.n_sparse):
            raise RuntimeError("R_inv (or Sigma) must be set for DTC and SOR predictions")
        n = structure.noa
        if self.n_kernels == 1:
            mean = np.zeros(1 + 3 * n + 6)
            var = np.zeros(1 + 3 * n + 6)
            m = self._model(0, structure._ctx)
            _lib.check(_lib.load().fm_predict_dense(m.h, structure.descriptors[0]._handle,
                                                    kind, _lib.ptr(mean), _lib.ptr(var)))
        else:
            Ks, kself = [], np.zeros(1 + 3 * n + 6)
            for k, kern in enumerate(self.kernels):
                vals = structure.descriptors[k]
                Ks.append(kern.envs_struc(self.sparse_descriptors[k], vals))
                kself += kern.self_kernel_struc(vals)
            K = np.concatenate(Ks, axis=0)
            mean = K.T @ self.alpha
            if kind == _lib.DENSE_SOR:
                V = K.T @ self.R_inv
                var = np.sum(V * V, axis=1)
            else:
                Q1 = self.L_inv @ K
                Q2 = self.R_inv.T @ K
                var = kself - np.sum(Q1 * Q1, axis=0) + np.sum(Q2 * Q2, axis=0)
        structure.mean_efs = mean
        structure.variance_efs = var

    def predict_SOR(self, structure):
        """sparse_gp.cpp:667-692: variance diag(K^T R_inv R_inv^T K)."""
        self._dense(structure, _lib.DENSE_SOR)

    def predict_DTC(self, structure):
        """sparse_gp.cpp:694-729: variance K_self - diag(K^T Kuu^-1 K) +
        diag(K^T Sigma K)."""
        self._dense(structure, _lib.DENSE_DTC)

    # Files (sparse_gp.cpp:1200-1466) ---------------------------------------
    def _calculator(self, kernel_index, calc):
        """The descriptor calculator the files describe: the first training
        structure's (sparse_gp.cpp:1225-1227), or the one given."""
        if calc is not None:
            return calc
        if self.training_structures:
            return self.training_structures[0].descriptor_calculators[kernel_index]
        raise ValueError("no training structure holds the descriptor calculator; pass calc=")

    def write_mapping_coefficients(self, file_name, contributor, kernel_index, calc=None):
        """sparse_gp.cpp:1200-1264 (pair_style flare coefficient file)."""
        from . import mapping
        mapping.write_mapping_coefficients(self, file_name, contributor, kernel_index,
                                           self._calculator(kernel_index, calc))

    def write_varmap_coefficients(self, file_name, contributor, kernel_index, calc=None):
        """sparse_gp.cpp:1266-1339 (power-1 variance map)."""
        from . import mapping
        mapping.write_varmap_coefficients(self, file_name, contributor, kernel_index,
                                          self._calculator(kernel_index, calc))

    def write_L_inverse(self, file_name, contributor, calcs=None):
        """sparse_gp.cpp:1341-1402."""
        from . import mapping
        if calcs is None:
            calcs = [self._calculator(k, None) for k in range(self.n_kernels)]
        mapping.write_L_inverse(self, file_name, contributor, calcs)

    def write_sparse_descriptors(self, file_name, contributor):
        """sparse_gp.cpp:1404-1466."""
        from . import mapping
        mapping.write_sparse_descriptors(self, file_name, contributor)

    # Model files (sparse_gp.cpp:1468-1479) ---------------------------------
    @staticmethod
    def to_json(file_name, sgp):
        """Writes the model in the reference's JSON layout."""
        from . import serialization
        serialization.dump(serialization.sparse_gp_to_json(sgp), file_name)

    @staticmethod
    def from_json(file_name, context=None):
        """Reads a model written by to_json (here or by the reference)."""
        from . import serialization
        return serialization.sparse_gp_from_json(serialization.load(file_name), context)


class _DeviceModel:
    """One kernel's resident model on a context."""

    def __init__(self, sgp, k, ctx):
        d, self._keep = self.describe(sgp, k)
        h = _lib.C.c_void_p()
        _lib.check(_lib.load().fm_model_create(ctx.h, _lib.C.byref(d), _lib.C.byref(h)))
        self.h = h
        self.ctx = ctx

    @staticmethod
    def describe(sgp, k):
        """fm_model_desc of kernel k (and the arrays it points into)."""
        kern = sgp.kernels[k]
        if not isinstance(kern, NormalizedDotProduct):
            raise ValueError("only NormalizedDotProduct kernels are supported on the GPU path")
        cd = sgp.sparse_descriptors[k]
        off = sum(c.n_clusters for c in sgp.sparse_descriptors[:k])
        M = cd.n_clusters
        Linv = Rinv = None
        if sgp.L_inv.shape[0] == sgp.n_sparse and M:
            Linv = sgp.L_inv[off:off + M, off:off + M]
        if sgp.n_kernels == 1 and sgp.R_inv.shape == (M, M) and M:
            Rinv = sgp.R_inv
        return _DeviceModel._desc(kern.sigma, kern.power, cd, sgp.alpha[off:off + M], Linv, Rinv)

    @staticmethod
    def _desc(sigma, power, cd, alpha, Linv=None, Rinv=None):
        M, nd = cd.n_clusters, cd.n_descriptors
        desc = _lib.f64(np.concatenate(cd.descriptors, axis=0) if M else np.zeros((0, nd)))
        norms = _lib.f64(np.concatenate(cd.descriptor_norms) if M else np.zeros(0))
        alpha = _lib.f64(alpha)
        Linv = _lib.f64(Linv) if Linv is not None else None
        Rinv = _lib.f64(Rinv) if Rinv is not None else None
        nbt = _lib.i32(cd.n_clusters_by_type)
        keep = (desc, norms, alpha, Linv, Rinv, nbt)
        P = _lib.C.POINTER(_lib.C.c_double)
        d = _lib.ModelDesc(float(sigma), float(power), cd.n_types, nd,
                           nbt.ctypes.data_as(_lib.C.POINTER(_lib.C.c_int32)),
                           desc.ctypes.data_as(P), norms.ctypes.data_as(P),
                           alpha.ctypes.data_as(P),
                           Linv.ctypes.data_as(P) if Linv is not None else None,
                           Rinv.ctypes.data_as(P) if Rinv is not None else None)
        return d, keep

    def _create(self, ctx, sigma, power, cd, alpha, Linv=None, Rinv=None):
        d, self._keep = self._desc(sigma, power, cd, alpha, Linv, Rinv)
        h = _lib.C.c_void_p()
        _lib.check(_lib.load().fm_model_create(ctx.h, _lib.C.byref(d), _lib.C.byref(h)))
        self.h = h
        self.ctx = ctx  # the context outlives the models made on it

    @classmethod
    def from_sparse_set(cls, ctx, sigma, power, cd):
        self = cls.__new__(cls)
        self._create(ctx, sigma, power, cd, np.zeros(cd.n_clusters))
        return self

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None:
                _lib.load().fm_model_destroy(self.h)
        except Exception:
            pass
