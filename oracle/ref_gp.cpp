// Oracle: NormalizedDotProduct kernel, sparse-GP predictions, mapping
// coefficients, coefficient files and the pair_style flare evaluation
// (test infrastructure only).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "ref_cpu.hpp"

namespace ref {

static const double kEmpty = 1e-8;  // normalized_dot_product.cpp:45,131

// ---------------------------------------------------------------------------
// Blocked GEMM, C = A * B^T (row-major), used by the scalable mode: B is
// packed k-major per (j, k) block and a 4 x 8 register block of C is
// accumulated with 4-wide vectors.
// ---------------------------------------------------------------------------
typedef double v4df __attribute__((vector_size(32)));

void gemm_abt(int m, int n, int k, const double *A, const double *B, double *C) {
  const int MB = 64, NB = 128, KB = 256;
  int nib = (m + MB - 1) / MB, njb = (n + NB - 1) / NB;
#pragma omp parallel
  {
    std::vector<double> Bp((size_t)KB * NB);
#pragma omp for collapse(2) schedule(dynamic)
    for (int ib = 0; ib < nib; ib++)
      for (int jb = 0; jb < njb; jb++) {
        int i0 = ib * MB, j0 = jb * NB;
        int i1 = std::min(m, i0 + MB), j1 = std::min(n, j0 + NB);
        int nb = j1 - j0, nbp = (nb + 7) / 8 * 8;
        for (int i = i0; i < i1; i++)
          for (int j = j0; j < j1; j++) C[(size_t)i * n + j] = 0;
        for (int k0 = 0; k0 < k; k0 += KB) {
          int kb = std::min(KB, k - k0);
          for (int kk = 0; kk < kb; kk++)
            for (int jj = 0; jj < nbp; jj++)
              Bp[(size_t)kk * nbp + jj] = jj < nb ? B[(size_t)(j0 + jj) * k + k0 + kk] : 0.0;
          int i = i0;
          for (; i + 4 <= i1; i += 4) {
            for (int jj = 0; jj < nbp; jj += 8) {
              v4df acc[4][2];
              for (int r = 0; r < 4; r++) acc[r][0] = acc[r][1] = v4df{0, 0, 0, 0};
              const double *a0 = A + (size_t)i * k + k0;
              for (int kk = 0; kk < kb; kk++) {
                v4df b0, b1;
                __builtin_memcpy(&b0, &Bp[(size_t)kk * nbp + jj], 32);
                __builtin_memcpy(&b1, &Bp[(size_t)kk * nbp + jj + 4], 32);
                for (int r = 0; r < 4; r++) {
                  double av = a0[(size_t)r * k + kk];
                  v4df a = {av, av, av, av};
                  acc[r][0] += a * b0;
                  acc[r][1] += a * b1;
                }
              }
              for (int r = 0; r < 4; r++)
                for (int q = 0; q < 8; q++) {
                  int j = j0 + jj + q;
                  if (j < j1) C[(size_t)(i + r) * n + j] += acc[r][q >> 2][q & 3];
                }
            }
          }
          for (; i < i1; i++)
            for (int jj = 0; jj < nb; jj++) {
              double acc = 0;
              for (int kk = 0; kk < kb; kk++)
                acc += A[(size_t)i * k + k0 + kk] * Bp[(size_t)kk * nbp + jj];
              C[(size_t)i * n + j0 + jj] += acc;
            }
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Sparse sets (descriptor.cpp:104-176)
// ---------------------------------------------------------------------------
SparseSet cluster_from_values(const DescriptorValues &d) {
  SparseSet s;
  s.n_types = d.n_types;
  s.n_descriptors = d.n_descriptors;
  s.descriptors = d.descriptors;
  s.descriptor_norms = d.descriptor_norms;
  s.n_clusters_by_type = d.n_clusters_by_type;
  s.cumulative_type_count.assign(d.n_types, 0);
  for (int t = 1; t < d.n_types; t++)
    s.cumulative_type_count[t] = s.cumulative_type_count[t - 1] + s.n_clusters_by_type[t - 1];
  s.n_clusters = d.n_clusters;
  return s;
}

SparseSet sparse_from_values(const DescriptorValues &d,
                             const std::vector<std::vector<int>> &by_type) {
  SparseSet s;
  s.n_types = d.n_types;
  s.n_descriptors = d.n_descriptors;
  int nd = d.n_descriptors;
  s.descriptors.resize(d.n_types);
  s.descriptor_norms.resize(d.n_types);
  s.n_clusters_by_type.assign(d.n_types, 0);
  s.cumulative_type_count.assign(d.n_types, 0);
  for (int t = 0; t < d.n_types; t++) {
    for (int idx : by_type[t]) {
      s.descriptors[t].insert(s.descriptors[t].end(), d.descriptors[t].begin() + (size_t)idx * nd,
                              d.descriptors[t].begin() + (size_t)(idx + 1) * nd);
      s.descriptor_norms[t].push_back(d.descriptor_norms[t][idx]);
    }
    s.n_clusters_by_type[t] = (int)by_type[t].size();
    s.n_clusters += s.n_clusters_by_type[t];
    if (t > 0) s.cumulative_type_count[t] = s.cumulative_type_count[t - 1] + s.n_clusters_by_type[t - 1];
  }
  return s;
}

// ---------------------------------------------------------------------------
// NormalizedDotProduct
// ---------------------------------------------------------------------------
static inline double dotn(const double *a, const double *b, int n) {
  double acc = 0;
  for (int i = 0; i < n; i++) acc += a[i] * b[i];
  return acc;
}

// normalized_dot_product.cpp:28-89
void envs_envs(const SparseSet &a, const SparseSet &b, const NDPKernel &k,
               std::vector<double> &K) {
  double sig_sq = k.sigma * k.sigma;
  int nd = a.n_descriptors;
  K.assign((size_t)a.n_clusters * b.n_clusters, 0.0);
  for (int s = 0; s < a.n_types; s++) {
    int n1 = a.n_clusters_by_type[s], n2 = b.n_clusters_by_type[s];
    int c1 = a.cumulative_type_count[s], c2 = b.cumulative_type_count[s];
    for (int i = 0; i < n1; i++) {
      double ni = a.descriptor_norms[s][i];
      if (ni < kEmpty) continue;
      for (int j = 0; j < n2; j++) {
        double nj = b.descriptor_norms[s][j];
        if (nj < kEmpty) continue;
        double dot = dotn(&a.descriptors[s][(size_t)i * nd], &b.descriptors[s][(size_t)j * nd], nd);
        double nrm = dot / (ni * nj);
        K[(size_t)(c1 + i) * b.n_clusters + c2 + j] += sig_sq * std::pow(nrm, k.power);
      }
    }
  }
}

// normalized_dot_product.cpp:117-214
void envs_struc(const SparseSet &envs, const DescriptorValues &d,
                const NDPKernel &k, std::vector<double> &K) {
  if (!d.has_force_dervs) throw std::runtime_error("envs_struc needs dense derivatives");
  double sig_sq = k.sigma * k.sigma;
  int nd = d.n_descriptors, na = d.n_atoms;
  int n_out = 1 + 3 * na + 6;
  double vol_inv = 1 / d.volume;
  K.assign((size_t)envs.n_clusters * n_out, 0.0);
  for (int s = 0; s < envs.n_types; s++) {
    int n_sparse = envs.n_clusters_by_type[s], n_struc = d.n_clusters_by_type[s];
    int c_sparse = envs.cumulative_type_count[s];
    int nnb = d.n_neighbors_by_type[s];
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n_sparse; i++) {
      double norm_i = envs.descriptor_norms[s][i];
      if (norm_i < kEmpty) continue;
      const double *x = &envs.descriptors[s][(size_t)i * nd];
      double *row = &K[(size_t)(c_sparse + i) * n_out];
      // force_dot(i, :) = x . dB/dr rows
      std::vector<double> force_dot(3 * (size_t)nnb);
      for (int q = 0; q < 3 * nnb; q++)
        force_dot[q] = dotn(x, &d.descriptor_force_dervs[s][(size_t)q * nd], nd);
      for (int j = 0; j < n_struc; j++) {
        double norm_j = d.descriptor_norms[s][j];
        double norm_ij = norm_i * norm_j;
        double norm_ij3 = norm_ij * norm_j * norm_j;
        if (norm_j < kEmpty) continue;
        double dot = dotn(x, &d.descriptors[s][(size_t)j * nd], nd);
        double norm_dot = dot / norm_ij;
        double dval = k.power * std::pow(norm_dot, k.power - 1);
        row[0] += sig_sq * std::pow(norm_dot, k.power);
        int n_neigh = d.neighbor_counts[s][j];
        int c_neigh = d.cumulative_neighbor_counts[s][j];
        int atom_index = d.atom_indices[s][j];
        for (int kk = 0; kk < n_neigh; kk++) {
          int nbr = d.neighbor_indices[s][c_neigh + kk];
          int stress_counter = 0;
          for (int comp = 0; comp < 3; comp++) {
            int ind = c_neigh + kk;
            int fi = 3 * ind + comp;
            double f1 = force_dot[fi] / norm_ij;
            double f2 = dot * d.descriptor_force_dots[s][fi] / norm_ij3;
            double fk = sig_sq * dval * (f1 - f2);
            row[1 + 3 * nbr + comp] -= fk;
            row[1 + 3 * atom_index + comp] += fk;
            for (int comp2 = comp; comp2 < 3; comp2++) {
              double coord = d.neighbor_coordinates[s][3 * (size_t)ind + comp2];
              row[1 + 3 * na + stress_counter] -= fk * coord * vol_inv;
              stress_counter++;
            }
          }
        }
      }
    }
  }
}

// normalized_dot_product.cpp:450-586, literally (including its stress term,
// which pairs component m of the force kernel with coordinates n >= m).
void self_kernel_struc(const DescriptorValues &d, const NDPKernel &k,
                       std::vector<double> &v) {
  if (!d.has_force_dervs) throw std::runtime_error("self_kernel_struc needs dense derivatives");
  double sig_sq = k.sigma * k.sigma;
  int nd = d.n_descriptors, na = d.n_atoms;
  int n_out = 1 + 3 * na + 6;
  double vol_inv = 1 / d.volume, vol_inv_sq = vol_inv * vol_inv;
  v.assign(n_out, 0.0);
  for (int s = 0; s < d.n_types; s++) {
    int n_struc = d.n_clusters_by_type[s];
    int nnb = d.n_neighbors_by_type[s];
    const std::vector<double> &D = d.descriptors[s], &dD = d.descriptor_force_dervs[s];
    const std::vector<double> &fdots = d.descriptor_force_dots[s];
    // force_dot(f, j) = dB_f . B_j
    std::vector<double> force_dot((size_t)3 * nnb * n_struc);
#pragma omp parallel for
    for (int f = 0; f < 3 * nnb; f++)
      for (int j = 0; j < n_struc; j++)
        force_dot[(size_t)f * n_struc + j] = dotn(&dD[(size_t)f * nd], &D[(size_t)j * nd], nd);
    std::vector<double> par((size_t)n_struc * n_out, 0.0);
    bool bad = false;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n_struc; i++) {
      double norm_i = d.descriptor_norms[s][i];
      if (norm_i < kEmpty) continue;
      double norm_i2 = norm_i * norm_i, norm_i3 = norm_i2 * norm_i;
      double *prow = &par[(size_t)i * n_out];
      for (int j = i; j < n_struc; j++) {
        double norm_j = d.descriptor_norms[s][j];
        if (norm_j < kEmpty) continue;
        double mult = (i == j) ? 1 : 2;
        double norm_j2 = norm_j * norm_j, norm_j3 = norm_j2 * norm_j;
        double norm_ij = norm_i * norm_j;
        double dot = dotn(&D[(size_t)i * nd], &D[(size_t)j * nd], nd);
        double norm_dot = dot / norm_ij;
        double c1 = (k.power - 1) * k.power * std::pow(norm_dot, k.power - 2);
        if (std::fabs(norm_dot) < kEmpty && k.power < 2) bad = true;
        double c2 = k.power * std::pow(norm_dot, k.power - 1);
        prow[0] += sig_sq * mult * std::pow(norm_dot, k.power);
        int nn1 = d.neighbor_counts[s][i], cn1 = d.cumulative_neighbor_counts[s][i];
        int ci1 = d.atom_indices[s][i];
        int nn2 = d.neighbor_counts[s][j], cn2 = d.cumulative_neighbor_counts[s][j];
        int ci2 = d.atom_indices[s][j];
        for (int a = 0; a < nn1; a++) {
          int ind1 = cn1 + a;
          int ni1 = d.neighbor_indices[s][ind1];
          for (int b = 0; b < nn2; b++) {
            int ind2 = cn2 + b;
            int ni2 = d.neighbor_indices[s][ind2];
            int sc = 0;
            for (int m = 0; m < 3; m++) {
              int f1 = 3 * ind1 + m, f2 = 3 * ind2 + m;
              double v1 = force_dot[(size_t)f1 * n_struc + j] / norm_ij - norm_dot * fdots[f1] / norm_i2;
              double v2 = force_dot[(size_t)f2 * n_struc + i] / norm_ij - norm_dot * fdots[f2] / norm_j2;
              double v3 = dotn(&dD[(size_t)f1 * nd], &dD[(size_t)f2 * nd], nd) / norm_ij;
              double v4 = fdots[f1] * force_dot[(size_t)f2 * n_struc + i] / (norm_i3 * norm_j);
              double v5 = fdots[f2] * force_dot[(size_t)f1 * n_struc + j] / (norm_i * norm_j3);
              double v6 = fdots[f1] * fdots[f2] * norm_dot / (norm_i2 * norm_j2);
              double kv = sig_sq * mult * (c1 * v1 * v2 + c2 * (v3 - v4 - v5 + v6));
              if (ci1 == ci2) prow[1 + ci1 * 3 + m] += kv;
              if (ci1 == ni2) prow[1 + ci1 * 3 + m] -= kv;
              if (ni1 == ci2) prow[1 + ni1 * 3 + m] -= kv;
              if (ni1 == ni2) prow[1 + ni1 * 3 + m] += kv;
              for (int nn = m; nn < 3; nn++) {
                double co1 = d.neighbor_coordinates[s][3 * (size_t)ind1 + nn];
                double co2 = d.neighbor_coordinates[s][3 * (size_t)ind2 + nn];
                prow[1 + 3 * na + sc] += kv * co1 * co2 * vol_inv_sq;
                sc++;
              }
            }
          }
        }
      }
    }
    if (bad)
      throw std::invalid_argument("Dot product of descriptors is 0, and the negative power function in force-force kernel diverges.");
    for (int i = 0; i < n_struc; i++)
      for (int q = 0; q < n_out; q++) v[q] += par[(size_t)i * n_out + q];
  }
}

// ---------------------------------------------------------------------------
// Sparse GP predictions
// ---------------------------------------------------------------------------
// sparse_gp.cpp:643-665: mean_efs = K^T alpha.
void predict_mean(const Model &m, const DescriptorValues &d,
                  std::vector<double> &mean_efs) {
  std::vector<double> K;
  envs_struc(m.sparse, d, m.kernel, K);
  int M = m.sparse.n_clusters, n_out = 1 + 3 * d.n_atoms + 6;
  mean_efs.assign(n_out, 0.0);
  for (int j = 0; j < M; j++)
    for (int c = 0; c < n_out; c++) mean_efs[c] += K[(size_t)j * n_out + c] * m.alpha[j];
}

// sparse_gp.cpp:93-139: K_self - diag(Q1^T Q1), Q1 = L_inv * K(cluster, sparse)^T.
void local_uncertainties(const Model &m, const DescriptorValues &d,
                         std::vector<double> &var) {
  SparseSet cl = cluster_from_values(d);
  int M = m.sparse.n_clusters, n = cl.n_clusters, nd = d.n_descriptors;
  double sig_sq = m.kernel.sigma * m.kernel.sigma;
  std::vector<double> Kcs;
  envs_envs(cl, m.sparse, m.kernel, Kcs);
  var.assign(n, 0.0);
#pragma omp parallel for
  for (int c = 0; c < n; c++) {
    // K_self(c): the diagonal of envs_envs(cluster, cluster).
    int s = 0;
    while (c >= cl.cumulative_type_count[s] + cl.n_clusters_by_type[s]) s++;
    int ci = c - cl.cumulative_type_count[s];
    double nrm = cl.descriptor_norms[s][ci];
    double kself = 0;
    if (nrm >= kEmpty) {
      const double *b = &cl.descriptors[s][(size_t)ci * nd];
      kself = sig_sq * std::pow(dotn(b, b, nd) / (nrm * nrm), m.kernel.power);
    }
    double q2 = 0;
    for (int r = 0; r < M; r++) {
      double acc = 0;
      for (int j = 0; j < M; j++) acc += m.L_inv[(size_t)r * M + j] * Kcs[(size_t)c * M + j];
      q2 += acc * acc;
    }
    var[c] = kself - q2;
  }
}

static void diag_quad(const std::vector<double> &K, int M, int n_out,
                      const std::vector<double> &A, std::vector<double> &out) {
  out.assign(n_out, 0.0);
#pragma omp parallel for
  for (int c = 0; c < n_out; c++) {
    double acc = 0;
    for (int i = 0; i < M; i++) {
      double t = 0;
      for (int j = 0; j < M; j++) t += A[(size_t)i * M + j] * K[(size_t)j * n_out + c];
      acc += K[(size_t)i * n_out + c] * t;
    }
    out[c] = acc;
  }
}

// sparse_gp.cpp:694-729
void predict_dtc(const Model &m, const DescriptorValues &d,
                 std::vector<double> &mean_efs, std::vector<double> &var_efs) {
  std::vector<double> K;
  envs_struc(m.sparse, d, m.kernel, K);
  int M = m.sparse.n_clusters, n_out = 1 + 3 * d.n_atoms + 6;
  mean_efs.assign(n_out, 0.0);
  for (int j = 0; j < M; j++)
    for (int c = 0; c < n_out; c++) mean_efs[c] += K[(size_t)j * n_out + c] * m.alpha[j];
  std::vector<double> kself, q, v;
  self_kernel_struc(d, m.kernel, kself);
  diag_quad(K, M, n_out, m.Kuu_inverse, q);
  diag_quad(K, M, n_out, m.Sigma, v);
  var_efs.resize(n_out);
  for (int c = 0; c < n_out; c++) var_efs[c] = kself[c] - q[c] + v[c];
}

// sparse_gp.cpp:667-692
void predict_sor(const Model &m, const DescriptorValues &d,
                 std::vector<double> &mean_efs, std::vector<double> &var_efs) {
  std::vector<double> K;
  envs_struc(m.sparse, d, m.kernel, K);
  int M = m.sparse.n_clusters, n_out = 1 + 3 * d.n_atoms + 6;
  mean_efs.assign(n_out, 0.0);
  for (int j = 0; j < M; j++)
    for (int c = 0; c < n_out; c++) mean_efs[c] += K[(size_t)j * n_out + c] * m.alpha[j];
  var_efs.assign(n_out, 0.0);
#pragma omp parallel for
  for (int c = 0; c < n_out; c++) {
    double acc = 0;
    for (int j = 0; j < M; j++) {
      double t = 0;
      for (int i = 0; i < M; i++) t += K[(size_t)i * n_out + c] * m.R_inv[(size_t)i * M + j];
      acc += t * t;
    }
    var_efs[c] = acc;
  }
}

// ---------------------------------------------------------------------------
// Scalable mode: per-atom adjoint with K.alpha.
// ---------------------------------------------------------------------------
// u[n, lm] = dE/dc[n, lm] given w = dE/dB (lammps_descriptor.cpp:286-314).
static void adjoint_u(const B2Settings &s, const double *w, const double *c, double *u) {
  int nr = s.nos * s.N, H = (s.lmax + 1) * (s.lmax + 1);
  for (int q = 0; q < nr * H; q++) u[q] = 0;
  int k = 0;
  for (int n1 = 0; n1 < nr; n1++)
    for (int n2 = n1; n2 < nr; n2++)
      for (int l = 0; l <= s.lmax; l++) {
        double wk = w[k++];
        for (int m = 0; m < 2 * l + 1; m++) {
          int a1 = n1 * H + l * l + m, a2 = n2 * H + l * l + m;
          if (n1 == n2) {
            u[a1] += 2 * wk * c[a1];
          } else {
            u[a1] += wk * c[a2];
            u[a2] += wk * c[a1];
          }
        }
      }
}

// Per-bond force f = dE_i/d r_ij from u, recomputing the bond basis.
static void bond_forces(const Geometry &g, const NeighborList &nl,
                        const B2Settings &s, int i, const double *u,
                        std::vector<int> &bonds, std::vector<double> &f) {
  int N = s.N, H = (s.lmax + 1) * (s.lmax + 1), si = g.species[i];
  bonds.clear();
  for (int q = nl.cumulative[i]; q < nl.cumulative[i + 1]; q++) {
    double rc = s.cutoffs[si * s.nos + nl.neighbor_species[q]];
    if (nl.rel[4 * (size_t)q] <= rc) bonds.push_back(q);
  }
  f.assign(3 * bonds.size(), 0.0);
  std::vector<double> gv(N), gx(N), gy(N), gz(N), h(H), hx(H), hy(H), hz(H);
  for (size_t b = 0; b < bonds.size(); b++) {
    int q = bonds[b];
    int sj = nl.neighbor_species[q];
    double rc = s.cutoffs[si * s.nos + sj];
    const double *rel = &nl.rel[4 * (size_t)q];
    radial_basis(gv.data(), gx.data(), gy.data(), gz.data(), rel[1], rel[2], rel[3], rel[0], rc, N,
                 s.r1, s.cutoff_kind);
    real_sph_harm(h.data(), hx.data(), hy.data(), hz.data(), rel[1], rel[2], rel[3], s.lmax);
    double fx = 0, fy = 0, fz = 0;
    for (int n = 0; n < N; n++)
      for (int lm = 0; lm < H; lm++) {
        double uu = u[(sj * N + n) * H + lm];
        fx += (gx[n] * h[lm] + gv[n] * hx[lm]) * uu;
        fy += (gy[n] * h[lm] + gv[n] * hy[lm]) * uu;
        fz += (gz[n] * h[lm] + gv[n] * hz[lm]) * uu;
      }
    f[3 * b] = fx;
    f[3 * b + 1] = fy;
    f[3 * b + 2] = fz;
  }
}

void predict_scalable(const Model &m, const Geometry &g, const NeighborList &nl,
                      const B2Settings &s, std::vector<double> &mean_efs,
                      std::vector<double> &local_var,
                      std::vector<double> *local_energy) {
  int noa = g.noa, nd = s.n_descriptors(), nbd = s.n_bond();
  double sig_sq = m.kernel.sigma * m.kernel.sigma, p = m.kernel.power;
  int M = m.sparse.n_clusters;
  std::vector<double> C((size_t)noa * nbd), Bh((size_t)noa * nd), norms(noa);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < noa; i++) {
    double *c = &C[(size_t)i * nbd];
    single_bond_atom(g, nl, s, i, c);
    double *b = &Bh[(size_t)i * nd];
    b2_from_c(s, c, b);
    double nn = 0;
    for (int q = 0; q < nd; q++) nn += b[q] * b[q];
    norms[i] = std::sqrt(nn);
    double inv = norms[i] < kEmpty ? 0.0 : 1.0 / norms[i];
    for (int q = 0; q < nd; q++) b[q] *= inv;
  }
  std::vector<double> E(noa, 0.0), dEdB((size_t)noa * nd, 0.0);
  local_var.assign(noa, 0.0);
  for (int t = 0; t < m.sparse.n_types; t++) {
    int Ms = m.sparse.n_clusters_by_type[t], cs = m.sparse.cumulative_type_count[t];
    std::vector<int> atoms;
    for (int i = 0; i < noa; i++)
      if (g.species[i] == t) atoms.push_back(i);
    if (atoms.empty()) continue;
    // Normalised sparse rows (empty rows zero) and their transpose.
    std::vector<double> X((size_t)Ms * nd, 0.0), XT((size_t)nd * Ms, 0.0), a(Ms, 0.0);
    for (int j = 0; j < Ms; j++) {
      double nrm = m.sparse.descriptor_norms[t][j];
      if (nrm < kEmpty) continue;
      for (int q = 0; q < nd; q++) {
        double v = m.sparse.descriptors[t][(size_t)j * nd + q] / nrm;
        X[(size_t)j * nd + q] = v;
        XT[(size_t)q * Ms + j] = v;
      }
      a[j] = m.alpha[cs + j];
    }
    // Rows of L_inv that touch this species block.
    int r_end = M;
    while (r_end > cs + Ms) {
      bool zero = true;
      for (int j = 0; j < Ms && zero; j++) zero = m.L_inv[(size_t)(r_end - 1) * M + cs + j] == 0.0;
      if (!zero) break;
      r_end--;
    }
    const int T = 256;
    for (size_t a0 = 0; a0 < atoms.size(); a0 += T) {
      int nt = (int)std::min((size_t)T, atoms.size() - a0);
      std::vector<double> Bt((size_t)nt * nd), S((size_t)nt * Ms), W((size_t)nt * Ms),
          G((size_t)nt * nd), Kt((size_t)nt * Ms);
      for (int r = 0; r < nt; r++)
        std::memcpy(&Bt[(size_t)r * nd], &Bh[(size_t)atoms[a0 + r] * nd], nd * sizeof(double));
      gemm_abt(nt, Ms, nd, Bt.data(), X.data(), S.data());
      for (int r = 0; r < nt; r++) {
        int i = atoms[a0 + r];
        double e = 0;
        bool empty = norms[i] < kEmpty;
        for (int j = 0; j < Ms; j++) {
          double sv = S[(size_t)r * Ms + j];
          bool skip = empty || m.sparse.descriptor_norms[t][j] < kEmpty;
          double kv = skip ? 0.0 : sig_sq * std::pow(sv, p);
          double wv = skip ? 0.0 : sig_sq * a[j] * p * std::pow(sv, p - 1);
          e += kv * a[j];
          W[(size_t)r * Ms + j] = wv;
          Kt[(size_t)r * Ms + j] = kv;
        }
        E[i] = e;
      }
      gemm_abt(nt, nd, Ms, W.data(), XT.data(), G.data());
      for (int r = 0; r < nt; r++) {
        int i = atoms[a0 + r];
        if (norms[i] < kEmpty) continue;
        const double *b = &Bh[(size_t)i * nd];
        double *gg = &G[(size_t)r * nd];
        double gb = dotn(gg, b, nd);
        for (int q = 0; q < nd; q++) dEdB[(size_t)i * nd + q] = (gg[q] - gb * b[q]) / norms[i];
      }
      // Local variance: sigma^2 (b.b)^p - |L_inv k|^2 over the rows that the
      // species block touches.
      std::vector<double> Lt((size_t)(r_end - cs) * Ms);
      for (int r = cs; r < r_end; r++)
        for (int j = 0; j < Ms; j++) Lt[(size_t)(r - cs) * Ms + j] = m.L_inv[(size_t)r * M + cs + j];
      std::vector<double> Q((size_t)nt * (r_end - cs));
      gemm_abt(nt, r_end - cs, Ms, Kt.data(), Lt.data(), Q.data());
      for (int r = 0; r < nt; r++) {
        int i = atoms[a0 + r];
        if (norms[i] < kEmpty) {
          local_var[i] = 0;
          continue;
        }
        const double *b = &Bh[(size_t)i * nd];
        // envs_envs diagonal: (B.B)/(|B||B|) on the raw descriptor; B-hat is
        // stored, so b.b plays that role.
        double kself = sig_sq * std::pow(dotn(b, b, nd), p);
        double q2 = 0;
        for (int c = 0; c < r_end - cs; c++) q2 += Q[(size_t)r * (r_end - cs) + c] * Q[(size_t)r * (r_end - cs) + c];
        local_var[i] = kself - q2;
      }
    }
  }
  // Adjoint to bond forces, then a deterministic gather.
  std::vector<std::vector<double>> fb(noa);
  std::vector<std::vector<int>> bb(noa);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < noa; i++) {
    std::vector<double> u(nbd);
    adjoint_u(s, &dEdB[(size_t)i * nd], &C[(size_t)i * nbd], u.data());
    bond_forces(g, nl, s, i, u.data(), bb[i], fb[i]);
  }
  int n_out = 1 + 3 * noa + 6;
  mean_efs.assign(n_out, 0.0);
  double etot = 0;
  for (int i = 0; i < noa; i++) etot += E[i];
  mean_efs[0] = etot;
  double vol_inv = 1 / g.volume;
  double st[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < noa; i++) {
    for (size_t b = 0; b < bb[i].size(); b++) {
      int q = bb[i][b];
      int j = nl.structure_indices[q];
      const double *rel = &nl.rel[4 * (size_t)q];
      int sc = 0;
      for (int comp = 0; comp < 3; comp++) {
        double fv = fb[i][3 * b + comp];
        mean_efs[1 + 3 * j + comp] -= fv;
        mean_efs[1 + 3 * i + comp] += fv;
        for (int c2 = comp; c2 < 3; c2++) st[sc++] -= fv * rel[1 + c2] * vol_inv;
      }
    }
  }
  for (int q = 0; q < 6; q++) mean_efs[1 + 3 * noa + q] = st[q];
  if (local_energy) *local_energy = E;
}

// ---------------------------------------------------------------------------
// Mapping coefficients and the coefficient file
// ---------------------------------------------------------------------------
// normalized_dot_product.cpp:653-767
void mapping_coefficients(const Model &m, std::vector<double> &beta, int &beta_size) {
  int nd = m.sparse.n_descriptors, nos = m.sparse.n_types;
  double sig2 = m.kernel.sigma * m.kernel.sigma;
  int p = (int)m.kernel.power;
  if (p != 1 && p != 2) throw std::runtime_error("mapping needs power 1 or 2");
  beta_size = p == 1 ? nd : nd * (nd + 1) / 2;
  beta.assign((size_t)nos * beta_size, 0.0);
  for (int t = 0; t < nos; t++) {
    int nt = m.sparse.n_clusters_by_type[t], ct = m.sparse.cumulative_type_count[t];
    double *row = &beta[(size_t)t * beta_size];
    for (int j = 0; j < nt; j++) {
      double pn = m.sparse.descriptor_norms[t][j];
      if (pn < kEmpty) continue;
      const double *x = &m.sparse.descriptors[t][(size_t)j * nd];
      double av = m.alpha[ct + j];
      int bc = 0;
      for (int k = 0; k < nd; k++) {
        double pk = x[k] / pn;
        if (p == 1) {
          row[bc++] += sig2 * pk * av;
        } else {
          for (int l = k; l < nd; l++) {
            double pl = x[l] / pn;
            double bv = sig2 * pk * pl * av;
            row[bc++] += (k != l) ? 2 * bv : bv;
          }
        }
      }
    }
  }
}

static std::string fmt_sci(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.16e", v);
  return buf;
}

static void descriptor_header(std::ostringstream &o, const B2Settings &s, int coeff_size) {
  // b2.cpp:52-74
  o << "chebyshev\n";
  o << s.nos << " " << s.N << " " << s.lmax << " " << coeff_size << "\n";
  o << (s.cutoff_kind == CUT_QUADRATIC ? "quadratic" : "cosine") << "\n";
  for (int i = 0; i < s.nos * s.nos; i++) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.2f ", s.cutoffs[i]);
    o << buf;
  }
  o << "\n";
}

// sparse_gp.cpp:1200-1264
std::string coefficient_file_text(const Model &m, const B2Settings &s,
                                  const std::string &date,
                                  const std::string &contributor) {
  std::vector<double> beta;
  int bs;
  mapping_coefficients(m, beta, bs);
  std::ostringstream o;
  o << "DATE: " << date << " CONTRIBUTOR: " << contributor << "\n";
  char pw[32];
  std::snprintf(pw, sizeof(pw), "%.0f", m.kernel.power);
  o << pw << " NormalizedDotProduct\n";
  descriptor_header(o, s, bs);
  int count = 0;
  for (int t = 0; t < s.nos; t++) {
    if (count != 0) o << "\n";
    for (int j = 0; j < bs; j++) {
      double v = beta[(size_t)t * bs + j];
      o << (v > 0 ? "  " : " ") << fmt_sci(v);
      count++;
      if (count == 5) {
        count = 0;
        o << "\n";
      }
    }
  }
  return o.str();
}

// pair_flare.cpp:262-419 (read_file) + grab (:427-439).
bool parse_coefficient_file(const std::string &text, MappedPotential &p,
                            std::string &err) {
  std::istringstream in(text);
  std::string line;
  if (!std::getline(in, line)) { err = "missing header"; return false; }
  if (!std::getline(in, line)) { err = "missing kernel line"; return false; }
  {
    std::istringstream ls(line);
    ls >> p.power >> p.kernel_name;
  }
  if (!std::getline(in, line)) { err = "missing radial line"; return false; }
  {
    std::istringstream ls(line);
    ls >> p.radial;
  }
  if (!std::getline(in, line)) { err = "missing sizes"; return false; }
  {
    std::istringstream ls(line);
    ls >> p.nos >> p.N >> p.lmax >> p.beta_size;
  }
  if (!std::getline(in, line)) { err = "missing cutoff fn"; return false; }
  std::string cut;
  {
    std::istringstream ls(line);
    ls >> cut;
  }
  std::vector<double> vals;
  double v;
  size_t want_cut = (size_t)p.nos * p.nos;
  std::string rest((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  std::istringstream rs(rest);
  while (vals.size() < want_cut && rs >> v) vals.push_back(v);
  if (vals.size() < want_cut) { err = "short cutoff block"; return false; }
  p.cutoffs = vals;
  int nr = p.N * p.nos;
  int nd = nr * (nr + 1) / 2 * (p.lmax + 1);
  int check;
  if (p.power == 1) check = nd;
  else if (p.power == 2) check = nd * (nd + 1) / 2;
  else { err = "Power should be 1 or 2."; return false; }
  if (check != p.beta_size) { err = "Beta size doesn't match the number of descriptors."; return false; }
  if (p.radial != "chebyshev") { err = "Please use chebyshev radial basis function."; return false; }
  if (cut == "quadratic") p.cutoff_kind = CUT_QUADRATIC;
  else if (cut == "cosine") p.cutoff_kind = CUT_COSINE;
  else { err = "Please use quadratic or cosine cutoff function."; return false; }
  if (p.kernel_name == "NormalizedDotProduct") p.normalized = true;
  else if (p.kernel_name == "DotProduct") p.normalized = false;
  else { err = "Kernel string not recognized, expected <power> <kernel string>"; return false; }
  std::vector<double> beta;
  while (beta.size() < (size_t)p.beta_size * p.nos && rs >> v) beta.push_back(v);
  if (beta.size() < (size_t)p.beta_size * p.nos) { err = "short beta block"; return false; }
  p.beta_matrix.assign(p.nos, {});
  size_t bc = 0;
  for (int t = 0; t < p.nos; t++) {
    if (p.power == 1) {
      p.beta_matrix[t].assign(beta.begin() + bc, beta.begin() + bc + nd);
      bc += nd;
    } else {
      std::vector<double> &bm = p.beta_matrix[t];
      bm.assign((size_t)nd * nd, 0.0);
      for (int i = 0; i < nd; i++)
        for (int j = i; j < nd; j++) {
          if (i == j) bm[(size_t)i * nd + j] = beta[bc];
          else {
            double h = beta[bc] / 2;
            bm[(size_t)i * nd + j] = h;
            bm[(size_t)j * nd + i] = h;
          }
          bc++;
        }
    }
  }
  return true;
}

// pair_flare.cpp:68-184 with lammps_descriptor.cpp:209-315, on the periodic
// neighbour list (bond filter rsq < rc^2 as pair_flare.cpp:115).
void pair_flare_eval(const MappedPotential &p, const Geometry &g,
                     const NeighborList &nl, std::vector<double> &atom_energy,
                     std::vector<double> &forces, double virial[6]) {
  int noa = g.noa, N = p.N, H = (p.lmax + 1) * (p.lmax + 1), nr = p.nos * N;
  int nbd = nr * H, nd = nr * (nr + 1) / 2 * (p.lmax + 1);
  atom_energy.assign(noa, 0.0);
  forces.assign(3 * (size_t)noa, 0.0);
  for (int q = 0; q < 6; q++) virial[q] = 0;
  B2Settings s;
  s.nos = p.nos;
  s.N = N;
  s.lmax = p.lmax;
  s.cutoff_kind = p.cutoff_kind;
  s.cutoffs = p.cutoffs;
  std::vector<std::vector<double>> fb(noa);
  std::vector<std::vector<int>> bb(noa);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < noa; i++) {
    int si = g.species[i];
    std::vector<int> bonds;
    for (int q = nl.cumulative[i]; q < nl.cumulative[i + 1]; q++) {
      double rc = p.cutoffs[si * p.nos + nl.neighbor_species[q]];
      const double *rel = &nl.rel[4 * (size_t)q];
      double rsq = rel[1] * rel[1] + rel[2] * rel[2] + rel[3] * rel[3];
      if (rsq < rc * rc) bonds.push_back(q);
    }
    std::vector<double> c(nbd, 0.0), gv(N), gx(N), gy(N), gz(N), h(H), hx(H), hy(H), hz(H);
    for (int q : bonds) {
      int sj = nl.neighbor_species[q];
      double rc = p.cutoffs[si * p.nos + sj];
      const double *rel = &nl.rel[4 * (size_t)q];
      double r = std::sqrt(rel[1] * rel[1] + rel[2] * rel[2] + rel[3] * rel[3]);
      radial_basis(gv.data(), gx.data(), gy.data(), gz.data(), rel[1], rel[2], rel[3], r, rc, N, 0.0,
                   p.cutoff_kind);
      real_sph_harm(h.data(), hx.data(), hy.data(), hz.data(), rel[1], rel[2], rel[3], p.lmax);
      for (int n = 0; n < N; n++)
        for (int lm = 0; lm < H; lm++) c[(sj * N + n) * H + lm] += gv[n] * h[lm];
    }
    std::vector<double> B(nd);
    b2_from_c(s, c.data(), B.data());
    double n2 = dotn(B.data(), B.data(), nd);
    if (n2 < kEmpty) continue;  // pair_flare.cpp:134-136
    std::vector<double> w(nd);
    double e;
    const std::vector<double> &bm = p.beta_matrix[si];
    if (p.normalized) {
      if (p.power == 1) {
        double bn = std::sqrt(n2);
        e = dotn(B.data(), bm.data(), nd) / bn;
        for (int k = 0; k < nd; k++) w[k] = bm[k] / bn - e * B[k] / n2;
      } else {
        std::vector<double> bp(nd);
        for (int k = 0; k < nd; k++) bp[k] = dotn(&bm[(size_t)k * nd], B.data(), nd);
        e = dotn(B.data(), bp.data(), nd) / n2;
        for (int k = 0; k < nd; k++) w[k] = 2 * (bp[k] - e * B[k]) / n2;
      }
    } else {
      if (p.power == 1) {
        for (int k = 0; k < nd; k++) w[k] = bm[k];
        e = dotn(B.data(), w.data(), nd);
      } else {
        std::vector<double> bp(nd);
        for (int k = 0; k < nd; k++) bp[k] = dotn(&bm[(size_t)k * nd], B.data(), nd);
        e = dotn(B.data(), bp.data(), nd);
        for (int k = 0; k < nd; k++) w[k] = 2 * bp[k];
      }
    }
    atom_energy[i] = e;
    std::vector<double> u(nbd);
    adjoint_u(s, w.data(), c.data(), u.data());
    bb[i] = bonds;
    fb[i].assign(3 * bonds.size(), 0.0);
    for (size_t b = 0; b < bonds.size(); b++) {
      int q = bonds[b];
      int sj = nl.neighbor_species[q];
      double rc = p.cutoffs[si * p.nos + sj];
      const double *rel = &nl.rel[4 * (size_t)q];
      double r = std::sqrt(rel[1] * rel[1] + rel[2] * rel[2] + rel[3] * rel[3]);
      radial_basis(gv.data(), gx.data(), gy.data(), gz.data(), rel[1], rel[2], rel[3], r, rc, N, 0.0,
                   p.cutoff_kind);
      real_sph_harm(h.data(), hx.data(), hy.data(), hz.data(), rel[1], rel[2], rel[3], p.lmax);
      double f3[3] = {0, 0, 0};
      for (int n = 0; n < N; n++)
        for (int lm = 0; lm < H; lm++) {
          double uu = u[(sj * N + n) * H + lm];
          f3[0] += (gx[n] * h[lm] + gv[n] * hx[lm]) * uu;
          f3[1] += (gy[n] * h[lm] + gv[n] * hy[lm]) * uu;
          f3[2] += (gz[n] * h[lm] + gv[n] * hz[lm]) * uu;
        }
      for (int k = 0; k < 3; k++) fb[i][3 * b + k] = f3[k];
    }
  }
  for (int i = 0; i < noa; i++)
    for (size_t b = 0; b < bb[i].size(); b++) {
      int q = bb[i][b];
      int j = nl.structure_indices[q];
      const double *rel = &nl.rel[4 * (size_t)q];
      int sc = 0;
      for (int a = 0; a < 3; a++) {
        double fv = fb[i][3 * b + a];
        forces[3 * i + a] += fv;
        forces[3 * (size_t)j + a] -= fv;
        for (int c2 = a; c2 < 3; c2++) virial[sc++] -= fv * rel[1 + c2];
      }
    }
}

// sparse_gp.cpp:1341-1402
std::string L_inverse_file_text(const Model &m, const B2Settings &s,
                                const std::string &date,
                                const std::string &contributor) {
  std::ostringstream o;
  o << "L_inverse_block file DATE: " << date << " CONTRIBUTOR: " << contributor << "\n";
  char pw[32];
  std::snprintf(pw, sizeof(pw), "%.0f", m.kernel.power);
  o << pw << " NormalizedDotProduct\n";
  o << 4 << "\n";
  double hyps[1] = {m.kernel.sigma};
  o << fmt_sci(hyps[0]) << " ";
  // The three noise hyperparameters are not part of the prediction state;
  // they are written as zero by this restatement.
  for (int q = 0; q < 3; q++) o << fmt_sci(0.0) << " ";
  o << "\n";
  descriptor_header(o, s, 1);
  int M = m.sparse.n_clusters;
  o << M << "\n";
  int count = 1;
  for (int j = 0; j < M; j++)
    for (int k = 0; k <= j; k++) {
      o << fmt_sci(m.L_inv[(size_t)j * M + k]) << " ";
      if (count % 5 == 0) o << "\n";
      count++;
    }
  if (count % 5 != 0) o << "\n";
  return o.str();
}

// sparse_gp.cpp:1404-1466
std::string sparse_desc_file_text(const Model &m, const std::string &date,
                                  const std::string &contributor) {
  std::ostringstream o;
  o << "sparse_descriptors file DATE: " << date << " CONTRIBUTOR: " << contributor << "\n";
  o << 1 << "\n";
  o << 0 << " " << m.sparse.n_clusters << " " << m.sparse.n_types << "\n";
  int nd = m.sparse.n_descriptors;
  int count = 1;
  for (int t = 0; t < m.sparse.n_types; t++) {
    o << m.sparse.n_clusters_by_type[t] << "\n";
    for (int j = 0; j < m.sparse.n_clusters_by_type[t]; j++) {
      double nrm = m.sparse.descriptor_norms[t][j];
      for (int k = 0; k < nd; k++) {
        if (nrm < kEmpty) o << fmt_sci(0.0) << " ";
        else o << fmt_sci(m.sparse.descriptors[t][(size_t)j * nd + k] / nrm) << " ";
        if (count % 5 == 0) o << "\n";
        count++;
      }
    }
    if (count % 5 != 1) o << "\n";
  }
  return o.str();
}

}  // namespace ref
