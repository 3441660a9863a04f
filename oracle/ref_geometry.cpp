// Oracle: periodic cell algebra and neighbour lists (test infrastructure only).
// Compiled with -ffp-contract=off so that every distance is the IEEE result of
// the expression order of src/flare_pp/structure.cpp:72-75.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

#include "ref_cpu.hpp"

namespace ref {

double det3(const double *m) {
  // Expansion along the first row.
  return m[0] * (m[4] * m[8] - m[5] * m[7]) -
         m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

void inverse3(const double *m, double *o) {
  double c00 = m[4] * m[8] - m[5] * m[7];
  double c01 = m[5] * m[6] - m[3] * m[8];
  double c02 = m[3] * m[7] - m[4] * m[6];
  double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  double inv = 1.0 / det;
  o[0] = c00 * inv;
  o[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
  o[3] = c01 * inv;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
  o[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
  o[6] = c02 * inv;
  o[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
}

// structure.cpp:135-165: the smallest in-plane height of the three cell faces.
static double single_sweep_cutoff(const double *c) {
  const double *v1 = c, *v2 = c + 3, *v3 = c + 6;
  double ab = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
  double ac = v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2];
  double bc = v2[0] * v3[0] + v2[1] * v3[1] + v2[2] * v3[2];
  double a = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  double b = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  double cc = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
  double cand[6];
  cand[0] = a * std::sqrt(1 - std::pow(ab / (a * b), 2));
  cand[1] = b * std::sqrt(1 - std::pow(ab / (a * b), 2));
  cand[2] = a * std::sqrt(1 - std::pow(ac / (a * cc), 2));
  cand[3] = cc * std::sqrt(1 - std::pow(ac / (a * cc), 2));
  cand[4] = b * std::sqrt(1 - std::pow(bc / (b * cc), 2));
  cand[5] = cc * std::sqrt(1 - std::pow(bc / (b * cc), 2));
  double best = cand[0];
  for (int i = 0; i < 6; i++)
    if (cand[i] < best) best = cand[i];
  return best;
}

// out = A(n x 3) * B(3 x 3), sums in index order.
static inline void row_times3(const double *a, const double *B, double *o) {
  for (int k = 0; k < 3; k++) o[k] = (a[0] * B[k] + a[1] * B[3 + k]) + a[2] * B[6 + k];
}

void make_geometry(Geometry &g, const double *cell, int noa, const int *species,
                   const double *positions, double cutoff) {
  g.noa = noa;
  for (int i = 0; i < 9; i++) g.cell[i] = cell[i];
  double ct[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) ct[3 * r + c] = cell[3 * c + r];
  inverse3(ct, g.cell_t_inv);
  // cell_dot = cell * cell^T
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      g.cell_dot[3 * r + c] = (cell[3 * r] * cell[3 * c] + cell[3 * r + 1] * cell[3 * c + 1]) +
                              cell[3 * r + 2] * cell[3 * c + 2];
  inverse3(g.cell_dot, g.cell_dot_inv);
  g.volume = std::fabs(det3(cell));
  g.single_sweep_cutoff = single_sweep_cutoff(cell);
  g.cutoff = cutoff;
  g.sweep = (int)std::ceil(cutoff / g.single_sweep_cutoff);
  g.species.assign(species, species + noa);
  g.positions.assign(positions, positions + 3 * (size_t)noa);
  g.wrapped.resize(3 * (size_t)noa);
  // structure.cpp:120-133: frac = (P * cell^T) * (cell cell^T)^-1, wrapped by
  // floor, then mapped back with the same two products.
  for (int i = 0; i < noa; i++) {
    double r1[3], fr[3], w1[3];
    row_times3(positions + 3 * i, ct, r1);
    row_times3(r1, g.cell_dot_inv, fr);
    for (int k = 0; k < 3; k++) fr[k] = fr[k] - std::floor(fr[k]);
    row_times3(fr, g.cell_dot, w1);
    row_times3(w1, g.cell_t_inv, &g.wrapped[3 * i]);
  }
}

// The image vector and distance of structure.cpp:72-75.
static inline double image_dist(const Geometry &g, int i, int j, int s1, int s2,
                                int s3, double im[3]) {
  const double *wi = &g.wrapped[3 * i], *wj = &g.wrapped[3 * j];
  for (int k = 0; k < 3; k++) {
    double d = wj[k] - wi[k];
    im[k] = ((d + (double)s1 * g.cell[k]) + (double)s2 * g.cell[3 + k]) +
            (double)s3 * g.cell[6 + k];
  }
  return std::sqrt(im[0] * im[0] + im[1] * im[1] + im[2] * im[2]);
}

static void finish_list(const Geometry &g, NeighborList &nl,
                        std::vector<std::vector<int>> &per_j,
                        std::vector<std::vector<int>> &per_s) {
  int n = g.noa;
  nl.count.assign(n, 0);
  nl.cumulative.assign(n + 1, 0);
  for (int i = 0; i < n; i++) nl.count[i] = (int)per_j[i].size();
  for (int i = 0; i < n; i++) nl.cumulative[i + 1] = nl.cumulative[i] + nl.count[i];
  size_t nnb = nl.cumulative[n];
  nl.structure_indices.resize(nnb);
  nl.neighbor_species.resize(nnb);
  nl.rel.resize(4 * nnb);
  nl.image.resize(3 * nnb);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    size_t base = nl.cumulative[i];
    for (size_t q = 0; q < per_j[i].size(); q++) {
      int j = per_j[i][q];
      int s1 = per_s[i][3 * q], s2 = per_s[i][3 * q + 1], s3 = per_s[i][3 * q + 2];
      double im[3];
      double dist = image_dist(g, i, j, s1, s2, s3, im);
      size_t o = base + q;
      nl.structure_indices[o] = j;
      nl.neighbor_species[o] = g.species[j];
      nl.rel[4 * o] = dist;
      nl.rel[4 * o + 1] = im[0];
      nl.rel[4 * o + 2] = im[1];
      nl.rel[4 * o + 3] = im[2];
      nl.image[3 * o] = s1;
      nl.image[3 * o + 1] = s2;
      nl.image[3 * o + 2] = s3;
    }
  }
}

// structure.cpp:52-118: every atom j, every image in [-sweep, sweep]^3 in the
// loop order j, s1, s2, s3; keep dist < cutoff and dist != 0.
void neighbors_literal(const Geometry &g, NeighborList &nl) {
  int n = g.noa, sw = g.sweep;
  std::vector<std::vector<int>> per_j(n), per_s(n);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++)
      for (int s1 = -sw; s1 <= sw; s1++)
        for (int s2 = -sw; s2 <= sw; s2++)
          for (int s3 = -sw; s3 <= sw; s3++) {
            double im[3];
            double dist = image_dist(g, i, j, s1, s2, s3, im);
            if ((dist < g.cutoff) && (dist != 0)) {
              per_j[i].push_back(j);
              per_s[i].push_back(s1);
              per_s[i].push_back(s2);
              per_s[i].push_back(s3);
            }
          }
  }
  finish_list(g, nl, per_j, per_s);
}

static inline int floor_div(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
  return q;
}

// Cell-list build whose output equals neighbors_literal: candidates come from
// bins along the fractional axes (bin width >= cutoff perpendicular to the
// face), every candidate is tested with the literal image/distance
// expression, images outside [-sweep, sweep]^3 are dropped as the literal
// loop never visits them, and each list is sorted by (j, s1, s2, s3).
void neighbors_cells(const Geometry &g, NeighborList &nl) {
  int n = g.noa, sw = g.sweep;
  const double *a = g.cell, *b = g.cell + 3, *c = g.cell + 6;
  double cr[3][3];
  auto cross = [](const double *u, const double *v, double *o) {
    o[0] = u[1] * v[2] - u[2] * v[1];
    o[1] = u[2] * v[0] - u[0] * v[2];
    o[2] = u[0] * v[1] - u[1] * v[0];
  };
  cross(b, c, cr[0]);
  cross(c, a, cr[1]);
  cross(a, b, cr[2]);
  double rc = g.cutoff * (1 + 1e-9) + 1e-9;
  int nb[3], K[3];
  for (int ax = 0; ax < 3; ax++) {
    double h = g.volume / std::sqrt(cr[ax][0] * cr[ax][0] + cr[ax][1] * cr[ax][1] +
                                    cr[ax][2] * cr[ax][2]);
    nb[ax] = std::max(1, (int)std::floor(h / rc));
    nb[ax] = std::min(nb[ax], 1024);
    double w = h / nb[ax];
    K[ax] = (int)std::ceil(rc / w);
  }
  // Fractional coordinates of the wrapped positions.
  std::vector<int> bin(3 * (size_t)n);
  for (int i = 0; i < n; i++) {
    double r1[3], fr[3];
    double ct[9];
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++) ct[3 * r + cc] = g.cell[3 * cc + r];
    row_times3(&g.wrapped[3 * i], ct, r1);
    row_times3(r1, g.cell_dot_inv, fr);
    for (int ax = 0; ax < 3; ax++) {
      int q = (int)std::floor(fr[ax] * nb[ax]);
      q = std::min(std::max(q, 0), nb[ax] - 1);
      bin[3 * i + ax] = q;
    }
  }
  int ncell = nb[0] * nb[1] * nb[2];
  std::vector<int> head(ncell + 1, 0), order(n);
  for (int i = 0; i < n; i++)
    head[(bin[3 * i] * nb[1] + bin[3 * i + 1]) * nb[2] + bin[3 * i + 2] + 1]++;
  for (int q = 0; q < ncell; q++) head[q + 1] += head[q];
  {
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (int i = 0; i < n; i++) {
      int q = (bin[3 * i] * nb[1] + bin[3 * i + 1]) * nb[2] + bin[3 * i + 2];
      order[fill[q]++] = i;
    }
  }
  std::vector<std::vector<int>> per_j(n), per_s(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    std::vector<std::pair<long long, int>> found;  // key, j
    std::vector<int> simg;
    for (int o0 = -K[0]; o0 <= K[0]; o0++) {
      int t0 = bin[3 * i] + o0, s1 = floor_div(t0, nb[0]), b0 = t0 - s1 * nb[0];
      if (s1 < -sw || s1 > sw) continue;
      for (int o1 = -K[1]; o1 <= K[1]; o1++) {
        int t1 = bin[3 * i + 1] + o1, s2 = floor_div(t1, nb[1]), b1 = t1 - s2 * nb[1];
        if (s2 < -sw || s2 > sw) continue;
        for (int o2 = -K[2]; o2 <= K[2]; o2++) {
          int t2 = bin[3 * i + 2] + o2, s3 = floor_div(t2, nb[2]), b2 = t2 - s3 * nb[2];
          if (s3 < -sw || s3 > sw) continue;
          int q = (b0 * nb[1] + b1) * nb[2] + b2;
          for (int p = head[q]; p < head[q + 1]; p++) {
            int j = order[p];
            double im[3];
            double dist = image_dist(g, i, j, s1, s2, s3, im);
            if ((dist < g.cutoff) && (dist != 0)) {
              int span = 2 * sw + 1;
              long long key = ((long long)j * span + (s1 + sw)) * span;
              key = (key + (s2 + sw)) * span + (s3 + sw);
              found.push_back({key, (int)(simg.size() / 3)});
              simg.push_back(s1);
              simg.push_back(s2);
              simg.push_back(s3);
              (void)j;
            }
          }
        }
      }
    }
    std::sort(found.begin(), found.end());
    int span = 2 * sw + 1;
    for (auto &f : found) {
      long long key = f.first;
      int j = (int)(key / ((long long)span * span * span));
      per_j[i].push_back(j);
      per_s[i].push_back(simg[3 * f.second]);
      per_s[i].push_back(simg[3 * f.second + 1]);
      per_s[i].push_back(simg[3 * f.second + 2]);
    }
  }
  finish_list(g, nl, per_j, per_s);
}

}  // namespace ref
