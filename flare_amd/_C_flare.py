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
    """Sparse GP prediction state and predictions (bffs/sparse_gp.h).

    The state (sparse set, alpha, L_inv) is set by ``add_*_environments`` +
    ``update_L_inverse`` (Kuu + jitter Cholesky in FP64 on the host, as
    sparse_gp.cpp:607-618) and ``set_alpha``, or all at once by
    ``set_state``. Predictions run on the GPU in the Structure's form.
    """

    def __init__(self, kernels=None, energy_noise=1.0, force_noise=1.0, stress_noise=1.0):
        self.kernels = list(kernels or [])
        self.n_kernels = len(self.kernels)
        self.energy_noise = float(energy_noise)
        self.force_noise = float(force_noise)
        self.stress_noise = float(stress_noise)
        self.Kuu_jitter = 1e-8  # sparse_gp.cpp:19
        self.sparse_descriptors = [ClusterDescriptor() for _ in self.kernels]
        self.alpha = np.zeros(0)
        self.L_inv = np.zeros((0, 0))
        self.Kuu = np.zeros((0, 0))
        self.Kuu_inverse = np.zeros((0, 0))
        self.Sigma = np.zeros((0, 0))
        self.R_inv = np.zeros((0, 0))
        self._models = {}

    @property
    def hyperparameters(self):
        h = [k.kernel_hyperparameters for k in self.kernels]
        return np.concatenate(h + [np.array([self.energy_noise, self.force_noise,
                                             self.stress_noise])])

    @property
    def n_sparse(self):
        return int(sum(c.n_clusters for c in self.sparse_descriptors))

    def _invalidate(self):
        self._models = {}

    # Sparse set (sparse_gp.cpp:141-341) ---------------------------------
    def add_specific_environments(self, structure, atoms):
        """Adds the given atoms of each descriptor (atom indices in structure
        order) to the sparse set."""
        atoms = [int(a) for a in atoms]
        for k, vals in enumerate(structure.descriptors):
            by_type = [[] for _ in range(vals.n_types)]
            pos = {}
            for s in range(vals.n_types):
                for r, a in enumerate(vals.atom_indices[s]):
                    pos[int(a)] = (s, r)
            for a in atoms:
                s, r = pos[a]
                by_type[s].append(r)
            self.sparse_descriptors[k].add_clusters_by_type(vals, by_type)
        self._invalidate()

    def add_all_environments(self, structure):
        for k, vals in enumerate(structure.descriptors):
            self.sparse_descriptors[k].add_all_clusters(vals)
        self._invalidate()

    def add_random_environments(self, structure, n_added, seed=None):
        rng = np.random.default_rng(seed)
        for k, vals in enumerate(structure.descriptors):
            by_type = [[] for _ in range(vals.n_types)]
            n = structure.noa
            pick = rng.choice(n, size=min(int(n_added[k]), n), replace=False)
            idx = {}
            for s in range(vals.n_types):
                for r, a in enumerate(vals.atom_indices[s]):
                    idx[int(a)] = (s, r)
            for a in sorted(int(x) for x in pick):
                s, r = idx[a]
                by_type[s].append(r)
            self.sparse_descriptors[k].add_clusters_by_type(vals, by_type)
        self._invalidate()

    # State -----------------------------------------------------------------
    def update_Kuu(self):
        blocks = [k.envs_envs(c, c) for k, c in zip(self.kernels, self.sparse_descriptors)]
        M = sum(b.shape[0] for b in blocks)
        self.Kuu = np.zeros((M, M))
        o = 0
        for b in blocks:
            self.Kuu[o:o + b.shape[0], o:o + b.shape[0]] = b
            o += b.shape[0]
        return self.Kuu

    def update_L_inverse(self):
        """L_inv = chol(Kuu + jitter I)^-1 and Kuu^-1 = L_inv^T L_inv
        (sparse_gp.cpp:610-618), FP64 on the host."""
        self.update_Kuu()
        M = self.Kuu.shape[0]
        L = np.linalg.cholesky(self.Kuu + self.Kuu_jitter * np.eye(M))
        from scipy.linalg import solve_triangular
        self.L_inv = solve_triangular(L, np.eye(M), lower=True)
        self.Kuu_inverse = self.L_inv.T @ self.L_inv
        self._invalidate()

    def set_alpha(self, alpha):
        alpha = np.asarray(alpha, dtype=np.float64).reshape(-1)
        if alpha.size != self.n_sparse:
            raise ValueError("alpha must have one entry per sparse environment")
        self.alpha = alpha.copy()
        self._invalidate()

    def set_state(self, alpha, L_inv=None, Sigma=None, Kuu_inverse=None, R_inv=None):
        """Sets the trained state. R_inv (the inverse of the QR factor R,
        sparse_gp.cpp:635-640) defines Sigma = R_inv R_inv^T; a Sigma given
        alone is factored as Sigma = R_inv R_inv^T by Cholesky."""
        self.set_alpha(alpha)
        if L_inv is not None:
            self.L_inv = np.array(L_inv, dtype=np.float64)
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

    def _model(self, k, ctx):
        key = (k, id(ctx))
        m = self._models.get(key)
        if m is None:
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

    def _dense(self, structure, kind):
        self._check_state()
        if self.n_kernels != 1:
            raise ValueError("DTC and SOR predictions take a model with one kernel")
        if self.R_inv.shape != (self.n_sparse, self.n_sparse):
            raise RuntimeError("R_inv (or Sigma) must be set for DTC and SOR predictions")
        n = structure.noa
        mean = np.zeros(1 + 3 * n + 6)
        var = np.zeros(1 + 3 * n + 6)
        m = self._model(0, structure._ctx)
        _lib.check(_lib.load().fm_predict_dense(m.h, structure._h, kind, _lib.ptr(mean),
                                                _lib.ptr(var)))
        structure.mean_efs = mean
        structure.variance_efs = var

    def predict_SOR(self, structure):
        """sparse_gp.cpp:667-692: variance diag(K^T R_inv R_inv^T K)."""
        self._dense(structure, _lib.DENSE_SOR)

    def predict_DTC(self, structure):
        """sparse_gp.cpp:694-729: variance K_self - diag(K^T Kuu^-1 K) +
        diag(K^T Sigma K)."""
        self._dense(structure, _lib.DENSE_DTC)


class _DeviceModel:
    """One kernel's resident model on a context."""

    def __init__(self, sgp, k, ctx):
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
        self._create(ctx, kern.sigma, kern.power, cd, sgp.alpha[off:off + M], Linv, Rinv)

    def _create(self, ctx, sigma, power, cd, alpha, Linv=None, Rinv=None):
        M, nd = cd.n_clusters, cd.n_descriptors
        desc = _lib.f64(np.concatenate(cd.descriptors, axis=0) if M else np.zeros((0, nd)))
        norms = _lib.f64(np.concatenate(cd.descriptor_norms) if M else np.zeros(0))
        alpha = _lib.f64(alpha)
        Linv = _lib.f64(Linv) if Linv is not None else None
        Rinv = _lib.f64(Rinv) if Rinv is not None else None
        self._keep = (desc, norms, alpha, Linv, Rinv)
        nbt = _lib.i32(cd.n_clusters_by_type)
        P = _lib.C.POINTER(_lib.C.c_double)
        d = _lib.ModelDesc(float(sigma), float(power), cd.n_types, nd,
                           nbt.ctypes.data_as(_lib.C.POINTER(_lib.C.c_int32)),
                           desc.ctypes.data_as(P), norms.ctypes.data_as(P),
                           alpha.ctypes.data_as(P),
                           Linv.ctypes.data_as(P) if Linv is not None else None,
                           Rinv.ctypes.data_as(P) if Rinv is not None else None)
        h = _lib.C.c_void_p()
        _lib.check(_lib.load().fm_model_create(ctx.h, _lib.C.byref(d), _lib.C.byref(h)))
        self.h = h

    @classmethod
    def from_sparse_set(cls, ctx, sigma, power, cd):
        self = cls.__new__(cls)
        self._create(ctx, sigma, power, cd, np.zeros(cd.n_clusters))
        return self

    def __del__(self):
        try:
            _lib.load().fm_model_destroy(self.h)
        except Exception:
            pass
