// Oracle: B3 descriptor (descriptors/b3.cpp) with the Wigner 3j symbols and
// complex spherical harmonics it uses (test infrastructure only).
#include <cmath>
#include <complex>
#include <vector>

#include "ref_cpu.hpp"

namespace ref {

using cd = std::complex<double>;

static double fact(int n) {
  double f = 1;
  for (int k = 2; k <= n; k++) f *= k;
  return f;
}

// Wigner 3j symbol (j1 j2 j3; m1 m2 m3) by the Racah formula (the values
// dev/derivations/compute_wigner.py tabulates with sympy for
// wigner3j.cpp:5-1988).
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0;
  double tri = fact(j1 + j2 - j3) * fact(j1 - j2 + j3) * fact(-j1 + j2 + j3) /
               fact(j1 + j2 + j3 + 1);
  double pre = std::sqrt(tri * fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) *
                         fact(j2 - m2) * fact(j3 + m3) * fact(j3 - m3));
  double sum = 0;
  for (int k = 0; k <= j1 + j2 + j3; k++) {
    int a[6] = {k, j3 - j2 + k + m1, j3 - j1 + k - m2, j1 + j2 - j3 - k, j1 - k - m1,
                j2 - k + m2};
    bool ok = true;
    for (int v : a) ok = ok && v >= 0;
    if (!ok) continue;
    double den = 1;
    for (int v : a) den *= fact(v);
    sum += ((k % 2) ? -1.0 : 1.0) / den;
  }
  int ph = j1 - j2 - m3;
  return ((ph % 2 + 2) % 2 ? -1.0 : 1.0) * pre * sum;
}

int b3_n_ls(int lmax) {
  int n = 0;
  for (int l1 = 0; l1 <= lmax; l1++)
    for (int l2 = 0; l2 <= lmax; l2++)
      for (int l3 = 0; l3 <= lmax; l3++)
        if (std::abs(l1 - l2) <= l3 && l3 <= l1 + l2) n++;
  return n;
}

// Complex Y_l^m with the Condon-Shortley phase, index l*l + (m + l), as
// get_complex_Y (y_grad.cpp:24-665; index 1 is Y_1^-1 ~ (x - iy)), with
// gradients. From the real harmonics Y^r_lm (no Condon-Shortley phase) of
// real_sph_harm: for m > 0, Y_l^m = (-1)^m (Y^r_lm + i Y^r_l,-m) / sqrt 2 and
// Y_l^-m = (Y^r_lm - i Y^r_l,-m) / sqrt 2; Y_l^0 = Y^r_l0.
void complex_sph_harm(cd *Y, cd *Yx, cd *Yy, cd *Yz, double x, double y, double z,
                      int lmax) {
  int H = (lmax + 1) * (lmax + 1);
  std::vector<double> R(H), Rx(H), Ry(H), Rz(H);
  real_sph_harm(R.data(), Rx.data(), Ry.data(), Rz.data(), x, y, z, lmax);
  const double s2 = 1.0 / std::sqrt(2.0);
  for (int l = 0; l <= lmax; l++) {
    int b = l * l + l;
    Y[b] = R[b];
    Yx[b] = Rx[b];
    Yy[b] = Ry[b];
    Yz[b] = Rz[b];
    for (int m = 1; m <= l; m++) {
      double sg = (m % 2) ? -1.0 : 1.0;
      int p = b + m, q = b - m;
      Y[p] = sg * s2 * cd(R[p], R[q]);
      Yx[p] = sg * s2 * cd(Rx[p], Rx[q]);
      Yy[p] = sg * s2 * cd(Ry[p], Ry[q]);
      Yz[p] = sg * s2 * cd(Rz[p], Rz[q]);
      Y[q] = s2 * cd(R[p], -R[q]);
      Yx[q] = s2 * cd(Rx[p], -Rx[q]);
      Yy[q] = s2 * cd(Ry[p], -Ry[q]);
      Yz[q] = s2 * cd(Rz[p], -Rz[q]);
    }
  }
}

// b3.cpp:291-443 for one atom: complex c[s_j*N*H + n*H + lm] += g_n Y_lm and,
// when dc != nullptr, the dense derivative rows 3*bond + comp.
static void complex_single_bond(const Geometry &g, const NeighborList &nl,
                                const B2Settings &s, int i, const std::vector<int> &bonds,
                                cd *c, cd *dc) {
  int N = s.N, H = (s.lmax + 1) * (s.lmax + 1), nb = s.n_bond();
  int si = g.species[i];
  for (int k = 0; k < nb; k++) c[k] = 0;
  if (dc)
    for (size_t k = 0; k < 3 * bonds.size() * (size_t)nb; k++) dc[k] = 0;
  std::vector<double> gv(N), gx(N), gy(N), gz(N);
  std::vector<cd> h(H), hx(H), hy(H), hz(H);
  for (size_t b = 0; b < bonds.size(); b++) {
    int q = bonds[b];
    int sj = nl.neighbor_species[q];
    double rc = s.cutoffs[si * s.nos + sj];
    const double *rel = &nl.rel[4 * (size_t)q];
    double r = rel[0], x = rel[1], y = rel[2], z = rel[3];
    radial_basis(gv.data(), gx.data(), gy.data(), gz.data(), x, y, z, r, rc, N, s.r1,
                 s.cutoff_kind);
    complex_sph_harm(h.data(), hx.data(), hy.data(), hz.data(), x, y, z, s.lmax);
    int base = sj * N * H;
    for (int n = 0; n < N; n++)
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

// b3.cpp:185-289. k runs over n1 <= n2 <= n3, then the triangle-valid
// (l1, l2, l3) in nested order; B3_k = Re sum_{m1 + m2 + m3 = 0}
// w3j(l1 l2 l3; m1 m2 m3) c[n1, l1 m1] c[n2, l2 m2] c[n3, l3 m3].
void compute_b3_struc(const Geometry &g, const NeighborList &nl, const B2Settings &s,
                      bool with_dervs, DescriptorValues &out) {
  int noa = g.noa, nos = s.nos, nbd = s.n_bond();
  int nr = s.nos * s.N, H = (s.lmax + 1) * (s.lmax + 1), L = s.lmax;
  int nd = nr * (nr + 1) * (nr + 2) / 6 * b3_n_ls(L);
  // The (l, m) terms of each valid (l1, l2, l3), in the reference's order.
  struct Term {
    int a1, a2, a3;
    double w;
  };
  std::vector<std::vector<Term>> terms;
  for (int l1 = 0; l1 <= L; l1++)
    for (int l2 = 0; l2 <= L; l2++)
      for (int l3 = 0; l3 <= L; l3++) {
        if (std::abs(l1 - l2) > l3 || l3 > l1 + l2) continue;
        std::vector<Term> t;
        for (int m1 = 0; m1 < 2 * l1 + 1; m1++)
          for (int m2 = 0; m2 < 2 * l2 + 1; m2++)
            for (int m3 = 0; m3 < 2 * l3 + 1; m3++) {
              if (m1 + m2 + m3 - l1 - l2 - l3 != 0) continue;
              double w = wigner3j(l1, l2, l3, m1 - l1, m2 - l2, m3 - l3);
              t.push_back({l1 * l1 + m1, l2 * l2 + m2, l3 * l3 + m3, w});
            }
        terms.push_back(t);
      }
  std::vector<std::vector<int>> bonds(noa);
  for (int i = 0; i < noa; i++) {
    int si = g.species[i];
    for (int q = nl.cumulative[i]; q < nl.cumulative[i + 1]; q++)
      if (nl.rel[4 * (size_t)q] <= s.cutoffs[si * s.nos + nl.neighbor_species[q]])
        bonds[i].push_back(q);
  }
  std::vector<std::vector<double>> B(noa), dB(noa), fdots(noa);
  std::vector<double> norms(noa);
#pragma omp parallel for schedule(dynamic, 2)
  for (int i = 0; i < noa; i++) {
    size_t nbi = bonds[i].size();
    std::vector<cd> c(nbd), dc;
    if (with_dervs) dc.resize(3 * nbi * nbd);
    complex_single_bond(g, nl, s, i, bonds[i], c.data(), with_dervs ? dc.data() : nullptr);
    B[i].assign(nd, 0.0);
    if (with_dervs) dB[i].assign(3 * nbi * nd, 0.0);
    int k = 0;
    for (int n1 = 0; n1 < nr; n1++)
      for (int n2 = n1; n2 < nr; n2++)
        for (int n3 = n2; n3 < nr; n3++)
          for (const auto &ts : terms) {
            double acc = 0;
            for (const Term &t : ts) {
              int a1 = n1 * H + t.a1, a2 = n2 * H + t.a2, a3 = n3 * H + t.a3;
              acc += std::real(c[a1] * c[a2] * c[a3] * t.w);
              if (with_dervs)
                for (size_t row = 0; row < 3 * nbi; row++)
                  dB[i][row * nd + k] +=
                      std::real(t.w * (dc[row * nbd + a1] * c[a2] * c[a3] +
                                       c[a1] * dc[row * nbd + a2] * c[a3] +
                                       c[a1] * c[a2] * dc[row * nbd + a3]));
            }
            B[i][k++] = acc;
          }
    if (with_dervs) {
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

}  // namespace ref
