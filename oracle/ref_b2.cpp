// Oracle: B2 descriptor and the species-sorted DescriptorValues packing
// (test infrastructure only).
#include <cmath>
#include <vector>

#include "ref_cpu.hpp"

namespace ref {

// The descriptor-cutoff filter of b2.cpp:262-280 (r <= rc(s_i, s_j)); returns
// the indices into the structure neighbour list, in list order.
static void filtered_bonds(const Geometry &g, const NeighborList &nl,
                           const B2Settings &s, int i, std::vector<int> &out) {
  out.clear();
  int si = g.species[i];
  for (int q = nl.cumulative[i]; q < nl.cumulative[i + 1]; q++) {
    double rc = s.cutoffs[si * s.nos + nl.neighbor_species[q]];
    if (nl.rel[4 * (size_t)q] <= rc) out.push_back(q);
  }
}

// b2.cpp:305-398 for one atom: c[s_j*N*H + n*H + lm] += g_n Y_lm and, when
// dc != nullptr, the dense derivative rows 3*bond + comp.
static void single_bond_dense(const Geometry &g, const NeighborList &nl,
                              const B2Settings &s, int i,
                              const std::vector<int> &bonds, double *c,
                              double *dc) {
  int N = s.N, H = (s.lmax + 1) * (s.lmax + 1), nb = s.n_bond();
  int si = g.species[i];
  for (int k = 0; k < nb; k++) c[k] = 0;
  if (dc)
    for (size_t k = 0; k < 3 * bonds.size() * (size_t)nb; k++) dc[k] = 0;
  std::vector<double> gv(N), gx(N), gy(N), gz(N), h(H), hx(H), hy(H), hz(H);
  for (size_t b = 0; b < bonds.size(); b++) {
    int q = bonds[b];
    int sj = nl.neighbor_species[q];
    double rc = s.cutoffs[si * s.nos + sj];
    const double *rel = &nl.rel[4 * (size_t)q];
    double r = rel[0], x = rel[1], y = rel[2], z = rel[3];
    radial_basis(gv.data(), gx.data(), gy.data(), gz.data(), x, y, z, r, rc, N,
                 s.r1, s.cutoff_kind);
    real_sph_harm(h.data(), hx.data(), hy.data(), hz.data(), x, y, z, s.lmax);
    int base = sj * N * H;
    for (int n = 0; n < N; n++) {
      for (int lm = 0; lm < H; lm++) {
        int k = base + n * H + lm;
        c[k] += gv[n] * h[lm];
        if (dc) {
          dc[(3 * b + 0) * nb + k] += gx[n] * h[lm] + gv[n] * hx[lm];
          dc[(3 * b + 1) * nb + k] += gy[n] * h[lm] + gv[n] * hy[lm];
          dc[(3 * b + 2) * nb + k] += gz[n] * h[lm] + gv[n] * hz[lm];
        }
      }
    }
  }
}

void single_bond_atom(const Geometry &g, const NeighborList &nl,
                      const B2Settings &s, int i, double *c) {
  std::vector<int> bonds;
  filtered_bonds(g, nl, s, i, bonds);
  single_bond_dense(g, nl, s, i, bonds, c, nullptr);
}

// b2.cpp:209-231: k runs over n1, n2 >= n1, l; B_k = sum_m c[n1,lm] c[n2,lm].
void b2_from_c(const B2Settings &s, const double *c, double *B) {
  int nr = s.nos * s.N, H = (s.lmax + 1) * (s.lmax + 1);
  int k = 0;
  for (int n1 = 0; n1 < nr; n1++)
    for (int n2 = n1; n2 < nr; n2++)
      for (int l = 0; l <= s.lmax; l++) {
        double acc = 0;
        for (int m = 0; m < 2 * l + 1; m++)
          acc += c[n1 * H + l * l + m] * c[n2 * H + l * l + m];
        B[k++] = acc;
      }
}

void compute_b2_struc(const Geometry &g, const NeighborList &nl,
                      const B2Settings &s, bool with_dervs,
                      DescriptorValues &out) {
  int noa = g.noa, nos = s.nos, nd = s.n_descriptors(), nbd = s.n_bond();
  int nr = s.nos * s.N, H = (s.lmax + 1) * (s.lmax + 1);
  std::vector<std::vector<int>> bonds(noa);
  for (int i = 0; i < noa; i++) filtered_bonds(g, nl, s, i, bonds[i]);

  // Per-atom results in atom order.
  std::vector<std::vector<double>> B(noa), dB(noa), fdots(noa);
  std::vector<double> norms(noa);
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < noa; i++) {
    size_t nbi = bonds[i].size();
    std::vector<double> c(nbd), dc;
    if (with_dervs) dc.resize(3 * nbi * nbd);
    single_bond_dense(g, nl, s, i, bonds[i], c.data(), with_dervs ? dc.data() : nullptr);
    B[i].assign(nd, 0.0);
    b2_from_c(s, c.data(), B[i].data());
    if (with_dervs) {
      // b2.cpp:219-228 product rule.
      dB[i].assign(3 * nbi * nd, 0.0);
      int k = 0;
      for (int n1 = 0; n1 < nr; n1++)
        for (int n2 = n1; n2 < nr; n2++)
          for (int l = 0; l <= s.lmax; l++) {
            for (int m = 0; m < 2 * l + 1; m++) {
              int a1 = n1 * H + l * l + m, a2 = n2 * H + l * l + m;
              for (size_t row = 0; row < 3 * nbi; row++)
                dB[i][row * nd + k] += c[a1] * dc[row * nbd + a2] + dc[row * nbd + a1] * c[a2];
            }
            k++;
          }
      fdots[i].assign(3 * nbi, 0.0);
      for (size_t row = 0; row < 3 * nbi; row++) {
        double acc = 0;
        for (int q = 0; q < nd; q++) acc += dB[i][row * nd + q] * B[i][q];
        fdots[i][row] = acc;
      }
    }
    double nn = 0;
    for (int q = 0; q < nd; q++) nn += B[i][q] * B[i][q];
    norms[i] = std::sqrt(nn);
  }

  pack_descriptor_values(g, nl, s, nd, bonds, B, dB, fdots, norms, with_dervs, out);
}

// b2.cpp:104-176 (b3.cpp:106-180): the species-sorted DescriptorValues
// repack of per-atom rows, derivative blocks and bond data (atom order kept
// within a species).
void pack_descriptor_values(const Geometry &g, const NeighborList &nl, const B2Settings &s,
                            int nd, const std::vector<std::vector<int>> &bonds,
                            const std::vector<std::vector<double>> &B,
                            const std::vector<std::vector<double>> &dB,
                            const std::vector<std::vector<double>> &fdots,
                            const std::vector<double> &norms, bool with_dervs,
                            DescriptorValues &out) {
  int noa = g.noa, nos = s.nos;
  out = DescriptorValues();
  out.n_descriptors = nd;
  out.n_types = nos;
  out.n_atoms = noa;
  out.volume = g.volume;
  out.has_force_dervs = with_dervs;
  std::vector<int> sc(nos, 0), nc(nos, 0);
  for (int i = 0; i < noa; i++) {
    sc[g.species[i]]++;
    nc[g.species[i]] += (int)bonds[i].size();
  }
  out.cumulative_type_count.push_back(0);
  out.descriptors.resize(nos);
  out.descriptor_force_dervs.resize(nos);
  out.neighbor_coordinates.resize(nos);
  out.descriptor_norms.resize(nos);
  out.descriptor_force_dots.resize(nos);
  out.cutoff_values.resize(nos);
  out.neighbor_counts.resize(nos);
  out.cumulative_neighbor_counts.resize(nos);
  out.atom_indices.resize(nos);
  out.neighbor_indices.resize(nos);
  for (int t = 0; t < nos; t++) {
    out.n_clusters_by_type.push_back(sc[t]);
    out.cumulative_type_count.push_back(out.cumulative_type_count[t] + sc[t]);
    out.n_clusters += sc[t];
    out.n_neighbors_by_type.push_back(nc[t]);
    out.descriptors[t].assign((size_t)sc[t] * nd, 0.0);
    if (with_dervs) {
      out.descriptor_force_dervs[t].assign((size_t)nc[t] * 3 * nd, 0.0);
      out.descriptor_force_dots[t].assign((size_t)nc[t] * 3, 0.0);
    }
    out.neighbor_coordinates[t].assign((size_t)nc[t] * 3, 0.0);
    out.descriptor_norms[t].assign(sc[t], 0.0);
    out.cutoff_values[t].assign(sc[t], 1.0);
    out.neighbor_counts[t].assign(sc[t], 0);
    out.cumulative_neighbor_counts[t].assign(sc[t], 0);
    out.atom_indices[t].assign(sc[t], 0);
    out.neighbor_indices[t].assign(nc[t], 0);
  }
  std::vector<int> scount(nos, 0), ncount(nos, 0);
  for (int i = 0; i < noa; i++) {
    int t = g.species[i];
    int si = scount[t], ni = ncount[t];
    int nbi = (int)bonds[i].size();
    for (int q = 0; q < nd; q++) out.descriptors[t][(size_t)si * nd + q] = B[i][q];
    if (with_dervs) {
      for (size_t q = 0; q < (size_t)3 * nbi * nd; q++)
        out.descriptor_force_dervs[t][(size_t)ni * 3 * nd + q] = dB[i][q];
      for (int q = 0; q < 3 * nbi; q++) out.descriptor_force_dots[t][3 * ni + q] = fdots[i][q];
    }
    for (int b = 0; b < nbi; b++) {
      int q = bonds[i][b];
      for (int k = 0; k < 3; k++)
        out.neighbor_coordinates[t][3 * (size_t)(ni + b) + k] = nl.rel[4 * (size_t)q + 1 + k];
      out.neighbor_indices[t][ni + b] = nl.structure_indices[q];
    }
    out.descriptor_norms[t][si] = norms[i];
    out.neighbor_counts[t][si] = nbi;
    out.cumulative_neighbor_counts[t][si] = ni;
    out.atom_indices[t][si] = i;
    scount[t]++;
    ncount[t] += nbi;
  }
}

}  // namespace ref
