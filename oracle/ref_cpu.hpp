// ref_cpu: FP64 CPU restatement of FLARE's sparse-GP force-evaluation path.
//
// TEST INFRASTRUCTURE ONLY. This is the oracle that the HIP product path is
// checked against; nothing in the product (flare_amd/, include/) links, loads
// or calls it. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use it.
//
// Each function cites the reference file:line whose algorithm it restates
// (paths relative to the mir-group/flare source tree). The checks that pin
// each part live in tests/test_oracle_*.py.
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace ref {

// ---------------------------------------------------------------------------
// Geometry (src/flare_pp/structure.cpp)
// ---------------------------------------------------------------------------
struct Geometry {
  int noa = 0;
  double cell[9];            // rows are lattice vectors
  double cell_t_inv[9];      // (cell^T)^-1
  double cell_dot[9];        // cell * cell^T
  double cell_dot_inv[9];    // (cell * cell^T)^-1
  double volume = 0, single_sweep_cutoff = 0, cutoff = 0;
  int sweep = 0;
  std::vector<int> species;
  std::vector<double> positions, wrapped;  // noa x 3
};

struct NeighborList {
  std::vector<int> count;       // noa
  std::vector<int> cumulative;  // noa + 1
  std::vector<int> structure_indices, neighbor_species;  // n_nb
  std::vector<double> rel;      // n_nb x 4 (r, x, y, z)
  std::vector<int> image;       // n_nb x 3 (s1, s2, s3); not stored by the reference
};

void inverse3(const double *m, double *out);   // cofactor inverse
double det3(const double *m);
void make_geometry(Geometry &g, const double *cell, int noa, const int *species,
                   const double *positions, double cutoff);
// structure.cpp:52-118 (O(N^2 * images))
void neighbors_literal(const Geometry &g, NeighborList &nl);
// Same output as neighbors_literal through a cell list (O(N)).
void neighbors_cells(const Geometry &g, NeighborList &nl);

// ---------------------------------------------------------------------------
// Bond basis (radial.cpp, cutoffs.cpp, y_grad.cpp)
// ---------------------------------------------------------------------------
enum CutoffKind { CUT_QUADRATIC = 0, CUT_COSINE = 1 };
void chebyshev(double *vals, double *derivs, double r, int N, double r1, double r2);
void cutoff_fn(int kind, double r, double rc, double &fc, double &dfc);
// g[N], gx/gy/gz[N]; radial.cpp:344-379
void radial_basis(double *g, double *gx, double *gy, double *gz, double x,
                  double y, double z, double r, double rc, int N, double r1,
                  int cutoff_kind);
// Real spherical harmonics (no Condon-Shortley phase), index l*l + (m + l).
void real_sph_harm(double *Y, double *Yx, double *Yy, double *Yz, double x,
                   double y, double z, int lmax);

// ---------------------------------------------------------------------------
// B2 descriptor (descriptors/b2.cpp) and DescriptorValues (descriptor.h)
// ---------------------------------------------------------------------------
struct B2Settings {
  int nos = 1, N = 1, lmax = 0;
  int cutoff_kind = CUT_QUADRATIC;
  double r1 = 0.0;
  std::vector<double> cutoffs;  // nos x nos, [central][neighbour]
  int n_descriptors() const {
    int nr = nos * N;
    return nr * (nr + 1) / 2 * (lmax + 1);
  }
  int n_bond() const { return nos * N * (lmax + 1) * (lmax + 1); }
};

struct DescriptorValues {
  int n_descriptors = 0, n_types = 0, n_atoms = 0, n_clusters = 0;
  double volume = 0;
  std::vector<std::vector<double>> descriptors, descriptor_force_dervs,
      neighbor_coordinates, descriptor_norms, descriptor_force_dots,
      cutoff_values;
  std::vector<std::vector<int>> neighbor_counts, cumulative_neighbor_counts,
      atom_indices, neighbor_indices;
  std::vector<int> n_clusters_by_type, cumulative_type_count,
      n_neighbors_by_type;
  bool has_force_dervs = false;
};

// b2.cpp:76-179 + 181-242 + 244-399. When with_dervs is false the dense
// derivative blocks are left empty (scalable mode).
void compute_b2_struc(const Geometry &g, const NeighborList &nl,
                      const B2Settings &s, bool with_dervs,
                      DescriptorValues &out);
// Per-atom single-bond coefficients c (n_bond) for atom i.
void single_bond_atom(const Geometry &g, const NeighborList &nl,
                      const B2Settings &s, int i, double *c);
void b2_from_c(const B2Settings &s, const double *c, double *B);
// Species-sorted repack shared by B2 and B3 (b2.cpp:104-176).
void pack_descriptor_values(const Geometry &g, const NeighborList &nl, const B2Settings &s,
                            int nd, const std::vector<std::vector<int>> &bonds,
                            const std::vector<std::vector<double>> &B,
                            const std::vector<std::vector<double>> &dB,
                            const std::vector<std::vector<double>> &fdots,
                            const std::vector<double> &norms, bool with_dervs,
                            DescriptorValues &out);

// ---------------------------------------------------------------------------
// B3 descriptor (descriptors/b3.cpp); lmax <= 4 as the reference's Wigner
// table (wigner3j.cpp:1983-1985). Settings as B2.
// ---------------------------------------------------------------------------
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
// Complex Y_l^m with the Condon-Shortley phase, index l*l + (m + l)
// (y_grad.cpp:24-665), and gradients.
void complex_sph_harm(std::complex<double> *Y, std::complex<double> *Yx,
                      std::complex<double> *Yy, std::complex<double> *Yz, double x, double y,
                      double z, int lmax);
int b3_n_ls(int lmax);
void compute_b3_struc(const Geometry &g, const NeighborList &nl, const B2Settings &s,
                      bool with_dervs, DescriptorValues &out);

// ---------------------------------------------------------------------------
// Sparse set and model state (descriptor.h:61-90, sparse_gp.h)
// ---------------------------------------------------------------------------
struct SparseSet {
  int n_types = 0, n_descriptors = 0, n_clusters = 0;
  std::vector<std::vector<double>> descriptors;       // raw rows per species
  std::vector<std::vector<double>> descriptor_norms;  // per species
  std::vector<int> n_clusters_by_type, cumulative_type_count;
};

struct NDPKernel {
  double sigma = 1.0, power = 2.0;
};

struct Model {
  NDPKernel kernel;
  SparseSet sparse;
  std::vector<double> alpha;        // M
  std::vector<double> L_inv;        // M x M (row-major)
  std::vector<double> Kuu_inverse;  // M x M
  std::vector<double> Sigma;        // M x M
  std::vector<double> R_inv;        // M x M
};

// normalized_dot_product.cpp:28-89
void envs_envs(const SparseSet &a, const SparseSet &b, const NDPKernel &k,
               std::vector<double> &K);
// normalized_dot_product.cpp:117-214 (needs dense derivatives)
void envs_struc(const SparseSet &envs, const DescriptorValues &d,
                const NDPKernel &k, std::vector<double> &K);
// normalized_dot_product.cpp:450-586
void self_kernel_struc(const DescriptorValues &d, const NDPKernel &k,
                       std::vector<double> &v);
SparseSet cluster_from_values(const DescriptorValues &d);
SparseSet sparse_from_values(const DescriptorValues &d,
                             const std::vector<std::vector<int>> &by_type);

// sparse_gp.cpp:643-665
void predict_mean(const Model &m, const DescriptorValues &d,
                  std::vector<double> &mean_efs);
// sparse_gp.cpp:93-139 (species-sorted order)
void local_uncertainties(const Model &m, const DescriptorValues &d,
                         std::vector<double> &var);
// sparse_gp.cpp:694-729
void predict_dtc(const Model &m, const DescriptorValues &d,
                 std::vector<double> &mean_efs, std::vector<double> &var_efs);
// sparse_gp.cpp:667-692
void predict_sor(const Model &m, const DescriptorValues &d,
                 std::vector<double> &mean_efs, std::vector<double> &var_efs);

// Scalable mode: cell-list neighbours, per-atom adjoint (the arithmetic of
// lammps_descriptor.cpp:250-315 / pair_flare.cpp:141-174, with K.alpha in
// place of the mapped beta). Outputs mean_efs (1 + 3N + 6) and local
// variances in atom order. Atom tiles go through a blocked GEMM.
void predict_scalable(const Model &m, const Geometry &g, const NeighborList &nl,
                      const B2Settings &s, std::vector<double> &mean_efs,
                      std::vector<double> &local_var_atom_order,
                      std::vector<double> *local_energy = nullptr);

// Mapping coefficients (normalized_dot_product.cpp:653-767) and the
// pair_style flare evaluation (lammps_descriptor.cpp:209-315,
// pair_flare.cpp:68-184).
void mapping_coefficients(const Model &m, std::vector<double> &beta,
                          int &beta_size);
std::string coefficient_file_text(const Model &m, const B2Settings &s,
                                  const std::string &date,
                                  const std::string &contributor);
struct MappedPotential {
  int power = 2, nos = 1, N = 1, lmax = 0, beta_size = 0;
  int cutoff_kind = CUT_QUADRATIC;
  bool normalized = true;
  std::string radial, kernel_name;
  std::vector<double> cutoffs;                  // nos x nos
  std::vector<std::vector<double>> beta_matrix; // per species (d x d) or d
};
bool parse_coefficient_file(const std::string &text, MappedPotential &p,
                            std::string &err);
// pair_flare.cpp:68-184 on a periodic structure: per-atom energy, forces,
// virial (xx, xy, xz, yy, yz, zz summed over bonds of -f (x) r).
void pair_flare_eval(const MappedPotential &p, const Geometry &g,
                     const NeighborList &nl, std::vector<double> &atom_energy,
                     std::vector<double> &forces, double virial[6]);

// Text writers for compute flare/std/atom (sparse_gp.cpp:1341-1466).
std::string L_inverse_file_text(const Model &m, const B2Settings &s,
                                const std::string &date,
                                const std::string &contributor);
std::string sparse_desc_file_text(const Model &m, const std::string &date,
                                  const std::string &contributor);

// Blocked row-major GEMM C = A * B^T (A: m x k, B: n x k), OpenMP.
void gemm_abt(int m, int n, int k, const double *A, const double *B, double *C);

}  // namespace ref
