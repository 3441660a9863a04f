// Oracle: plain C surface of ref_cpu for ctypes (test infrastructure only).
#include <complex>
#include <cstring>
#include <string>

#include "ref_cpu.hpp"

using namespace ref;

namespace {
struct GeomHandle {
  Geometry g;
  NeighborList nl;
};
struct DescHandle {
  DescriptorValues d;
};
B2Settings make_settings(int nos, int N, int lmax, int cutoff_kind, double r1,
                         const double *cutoffs) {
  B2Settings s;
  s.nos = nos;
  s.N = N;
  s.lmax = lmax;
  s.cutoff_kind = cutoff_kind;
  s.r1 = r1;
  s.cutoffs.assign(cutoffs, cutoffs + nos * nos);
  return s;
}
thread_local std::string g_err;
}  // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

void *ref_geom_new(const double *cell, int noa, const int *species,
                   const double *positions, double cutoff, int mode) {
  auto *h = new GeomHandle();
  make_geometry(h->g, cell, noa, species, positions, cutoff);
  if (mode == 0) neighbors_literal(h->g, h->nl);
  else neighbors_cells(h->g, h->nl);
  return h;
}
void ref_geom_free(void *p) { delete (GeomHandle *)p; }

void ref_geom_info(void *p, int *ints, double *dbls) {
  auto *h = (GeomHandle *)p;
  ints[0] = h->g.noa;
  ints[1] = h->g.sweep;
  ints[2] = h->nl.cumulative.empty() ? 0 : h->nl.cumulative.back();
  dbls[0] = h->g.volume;
  dbls[1] = h->g.single_sweep_cutoff;
}

// field: 0 wrapped, 1 count, 2 cumulative, 3 structure_indices,
// 4 neighbor_species, 5 rel, 6 image
void ref_geom_get(void *p, int field, void *out) {
  auto *h = (GeomHandle *)p;
  switch (field) {
    case 0: std::memcpy(out, h->g.wrapped.data(), h->g.wrapped.size() * 8); break;
    case 1: std::memcpy(out, h->nl.count.data(), h->nl.count.size() * 4); break;
    case 2: std::memcpy(out, h->nl.cumulative.data(), h->nl.cumulative.size() * 4); break;
    case 3: std::memcpy(out, h->nl.structure_indices.data(), h->nl.structure_indices.size() * 4); break;
    case 4: std::memcpy(out, h->nl.neighbor_species.data(), h->nl.neighbor_species.size() * 4); break;
    case 5: std::memcpy(out, h->nl.rel.data(), h->nl.rel.size() * 8); break;
    case 6: std::memcpy(out, h->nl.image.data(), h->nl.image.size() * 4); break;
  }
}

void *ref_desc_new(void *geom, int nos, int N, int lmax, int cutoff_kind,
                   double r1, const double *cutoffs, int with_dervs) {
  auto *gh = (GeomHandle *)geom;
  auto *h = new DescHandle();
  B2Settings s = make_settings(nos, N, lmax, cutoff_kind, r1, cutoffs);
  compute_b2_struc(gh->g, gh->nl, s, with_dervs != 0, h->d);
  return h;
}
// The same DescriptorValues for the B3 descriptor (b3.cpp:80-289).
void *ref_desc_new_b3(void *geom, int nos, int N, int lmax, int cutoff_kind, double r1,
                      const double *cutoffs, int with_dervs) {
  auto *gh = (GeomHandle *)geom;
  auto *h = new DescHandle();
  B2Settings s = make_settings(nos, N, lmax, cutoff_kind, r1, cutoffs);
  compute_b3_struc(gh->g, gh->nl, s, with_dervs != 0, h->d);
  return h;
}
void ref_desc_free(void *p) { delete (DescHandle *)p; }

// out: n_descriptors, n_types, n_atoms, n_clusters; per species s:
// n_clusters_by_type[s], n_neighbors_by_type[s], cumulative_type_count[s]
void ref_desc_sizes(void *p, int s, int *out) {
  auto &d = ((DescHandle *)p)->d;
  out[0] = d.n_descriptors;
  out[1] = d.n_types;
  out[2] = d.n_atoms;
  out[3] = d.n_clusters;
  out[4] = d.n_clusters_by_type[s];
  out[5] = d.n_neighbors_by_type[s];
  out[6] = d.cumulative_type_count[s];
  out[7] = d.has_force_dervs ? 1 : 0;
}

// field: 0 descriptors, 1 force_dervs, 2 neighbor_coordinates, 3 norms,
// 4 force_dots, 5 cutoff_values, 6 neighbor_counts, 7 cumulative counts,
// 8 atom_indices, 9 neighbor_indices
void ref_desc_get(void *p, int field, int s, void *out) {
  auto &d = ((DescHandle *)p)->d;
  auto cpd = [&](const std::vector<double> &v) { std::memcpy(out, v.data(), v.size() * 8); };
  auto cpi = [&](const std::vector<int> &v) { std::memcpy(out, v.data(), v.size() * 4); };
  switch (field) {
    case 0: cpd(d.descriptors[s]); break;
    case 1: cpd(d.descriptor_force_dervs[s]); break;
    case 2: cpd(d.neighbor_coordinates[s]); break;
    case 3: cpd(d.descriptor_norms[s]); break;
    case 4: cpd(d.descriptor_force_dots[s]); break;
    case 5: cpd(d.cutoff_values[s]); break;
    case 6: cpi(d.neighbor_counts[s]); break;
    case 7: cpi(d.cumulative_neighbor_counts[s]); break;
    case 8: cpi(d.atom_indices[s]); break;
    case 9: cpi(d.neighbor_indices[s]); break;
  }
}

// Sparse rows are given per species, concatenated in species order.
void *ref_model_new(double sigma, double power, int nos, int nd,
                    const int *n_by_type, const double *descriptors,
                    const double *norms, const double *alpha, const double *L_inv,
                    const double *Kuu_inverse, const double *Sigma,
                    const double *R_inv) {
  auto *m = new Model();
  m->kernel.sigma = sigma;
  m->kernel.power = power;
  SparseSet &sp = m->sparse;
  sp.n_types = nos;
  sp.n_descriptors = nd;
  sp.descriptors.resize(nos);
  sp.descriptor_norms.resize(nos);
  sp.n_clusters_by_type.assign(n_by_type, n_by_type + nos);
  sp.cumulative_type_count.assign(nos, 0);
  int off = 0;
  for (int t = 0; t < nos; t++) {
    sp.cumulative_type_count[t] = off;
    sp.descriptors[t].assign(descriptors + (size_t)off * nd, descriptors + (size_t)(off + n_by_type[t]) * nd);
    sp.descriptor_norms[t].assign(norms + off, norms + off + n_by_type[t]);
    off += n_by_type[t];
  }
  sp.n_clusters = off;
  size_t M = off;
  m->alpha.assign(alpha, alpha + M);
  if (L_inv) m->L_inv.assign(L_inv, L_inv + M * M);
  if (Kuu_inverse) m->Kuu_inverse.assign(Kuu_inverse, Kuu_inverse + M * M);
  if (Sigma) m->Sigma.assign(Sigma, Sigma + M * M);
  if (R_inv) m->R_inv.assign(R_inv, R_inv + M * M);
  return m;
}
void ref_model_free(void *p) { delete (Model *)p; }

int ref_envs_struc(void *model, void *desc, double *out) {
  try {
    std::vector<double> K;
    envs_struc(((Model *)model)->sparse, ((DescHandle *)desc)->d, ((Model *)model)->kernel, K);
    std::memcpy(out, K.data(), K.size() * 8);
    return 0;
  } catch (std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Kernel between all clusters of desc and the model's sparse set.
int ref_envs_envs(void *model, void *desc, double *out) {
  auto *m = (Model *)model;
  SparseSet cl = cluster_from_values(((DescHandle *)desc)->d);
  std::vector<double> K;
  envs_envs(cl, m->sparse, m->kernel, K);
  std::memcpy(out, K.data(), K.size() * 8);
  return 0;
}

int ref_self_kernel(void *model, void *desc, double *out) {
  try {
    std::vector<double> v;
    self_kernel_struc(((DescHandle *)desc)->d, ((Model *)model)->kernel, v);
    std::memcpy(out, v.data(), v.size() * 8);
    return 0;
  } catch (std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

int ref_predict(void *model, void *desc, int kind, double *mean, double *var) {
  try {
    auto *m = (Model *)model;
    auto &d = ((DescHandle *)desc)->d;
    std::vector<double> me, va;
    if (kind == 0) {
      predict_mean(*m, d, me);
    } else if (kind == 1) {
      predict_mean(*m, d, me);
      local_uncertainties(*m, d, va);
    } else if (kind == 2) {
      predict_dtc(*m, d, me, va);
    } else {
      predict_sor(*m, d, me, va);
    }
    std::memcpy(mean, me.data(), me.size() * 8);
    if (var && !va.empty()) std::memcpy(var, va.data(), va.size() * 8);
    return 0;
  } catch (std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

int ref_predict_scalable(void *model, void *geom, int nos, int N, int lmax,
                         int cutoff_kind, double r1, const double *cutoffs,
                         double *mean, double *var_atom_order,
                         double *local_energy) {
  auto *m = (Model *)model;
  auto *gh = (GeomHandle *)geom;
  B2Settings s = make_settings(nos, N, lmax, cutoff_kind, r1, cutoffs);
  std::vector<double> me, va, le;
  predict_scalable(*m, gh->g, gh->nl, s, me, va, &le);
  std::memcpy(mean, me.data(), me.size() * 8);
  std::memcpy(var_atom_order, va.data(), va.size() * 8);
  if (local_energy) std::memcpy(local_energy, le.data(), le.size() * 8);
  return 0;
}

int ref_mapping(void *model, double *beta, int *beta_size) {
  try {
    std::vector<double> b;
    mapping_coefficients(*(Model *)model, b, *beta_size);
    if (beta) std::memcpy(beta, b.data(), b.size() * 8);
    return 0;
  } catch (std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Writes the text into buf when buflen is large enough; returns its length.
long ref_coeff_text(void *model, int nos, int N, int lmax, int cutoff_kind,
                    const double *cutoffs, const char *date, const char *contributor,
                    int which, char *buf, long buflen) {
  auto *m = (Model *)model;
  B2Settings s = make_settings(nos, N, lmax, cutoff_kind, 0.0, cutoffs);
  std::string t;
  if (which == 0) t = coefficient_file_text(*m, s, date, contributor);
  else if (which == 1) t = L_inverse_file_text(*m, s, date, contributor);
  else t = sparse_desc_file_text(*m, date, contributor);
  if (buf && buflen > (long)t.size()) {
    std::memcpy(buf, t.data(), t.size());
    buf[t.size()] = 0;
  }
  return (long)t.size();
}

int ref_pair_flare(const char *text, void *geom, double *atom_energy,
                   double *forces, double *virial) {
  MappedPotential p;
  std::string err;
  if (!parse_coefficient_file(text, p, err)) {
    g_err = err;
    return 1;
  }
  auto *gh = (GeomHandle *)geom;
  std::vector<double> e, f;
  pair_flare_eval(p, gh->g, gh->nl, e, f, virial);
  std::memcpy(atom_energy, e.data(), e.size() * 8);
  std::memcpy(forces, f.data(), f.size() * 8);
  return 0;
}

double ref_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  return wigner3j(j1, j2, j3, m1, m2, m3);
}
int ref_b3_n_ls(int lmax) { return b3_n_ls(lmax); }

// Complex Y_l^m and gradients, interleaved (re, im): out[8 H].
void ref_complex_sph_harm(double x, double y, double z, int lmax, double *out) {
  int H = (lmax + 1) * (lmax + 1);
  std::vector<std::complex<double>> Y(H), Yx(H), Yy(H), Yz(H);
  complex_sph_harm(Y.data(), Yx.data(), Yy.data(), Yz.data(), x, y, z, lmax);
  const std::vector<std::complex<double>> *parts[4] = {&Y, &Yx, &Yy, &Yz};
  for (int p = 0; p < 4; p++)
    for (int k = 0; k < H; k++) {
      out[2 * (p * H + k)] = (*parts[p])[k].real();
      out[2 * (p * H + k) + 1] = (*parts[p])[k].imag();
    }
}

void ref_sph_harm(double x, double y, double z, int lmax, double *out) {
  int H = (lmax + 1) * (lmax + 1);
  real_sph_harm(out, out + H, out + 2 * H, out + 3 * H, x, y, z, lmax);
}

void ref_radial(double x, double y, double z, double r, double rc, int N,
                double r1, int kind, double *out) {
  radial_basis(out, out + N, out + 2 * N, out + 3 * N, x, y, z, r, rc, N, r1, kind);
}

void ref_gemm_abt(int m, int n, int k, const double *A, const double *B, double *C) {
  gemm_abt(m, n, k, A, B, C);
}

}  // extern "C"
