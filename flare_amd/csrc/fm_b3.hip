// B3 descriptor (descriptors/b3.cpp:80-443) on top of the B2 bond machinery.
//
// The complex single-bond coefficients of b3.cpp:291-443 (complex Y_l^m with
// the Condon-Shortley phase) are a fixed linear map of the real ones the
// descriptor kernel already accumulates: for m > 0,
//   c_l^m  = (-1)^m (c^r_lm + i c^r_l,-m) / sqrt 2,
//   c_l^-m = (c^r_lm - i c^r_l,-m) / sqrt 2,   c_l^0 = c^r_l0.
// So B3 runs the bond pass of k_b2 (c only, kept in HBM), then per atom
//   B3_k = Re sum_{m1 + m2 + m3 = 0} w3j(l1 l2 l3; m1 m2 m3) c[n1, l1 m1]
//          c[n2, l2 m2] c[n3, l3 m3]
// over n1 <= n2 <= n3 and the triangle-valid (l1, l2, l3) in the
// reference's order (b3.cpp:185-289). The adjoint maps dE/dB3 to the complex
// gradient dE/dRe c + i dE/dIm c = sum conj(w3j w_k c c) over the two other
// factors, then through the transpose of the linear map to dE/dc^r, so the
// bond-force kernel of B2 produces the forces unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <type_traits>
#include <vector>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

namespace {

double fact(int n) {
  double f = 1;
  for (int k = 2; k <= n; k++) f *= k;
  return f;
}

// Wigner 3j symbol by the Racah formula (the values wigner3j.cpp tabulates).
double w3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0;
  double tri = fact(j1 + j2 - j3) * fact(j1 - j2 + j3) * fact(-j1 + j2 + j3) /
               fact(j1 + j2 + j3 + 1);
  double pre = std::sqrt(tri * fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) * fact(j2 - m2) *
                         fact(j3 + m3) * fact(j3 - m3));
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
  int ph = ((j1 - j2 - m3) % 2 + 2) % 2;
  return (ph ? -1.0 : 1.0) * pre * sum;
}

}  // namespace

// Host tables of a B3 setting (shared by the forward and adjoint kernels).
struct B3Tables {
  int nr = 0, H = 0, n_ls = 0, n_tri = 0, nd = 0;
  std::vector<int> tri;        // n_tri x 3 (n1, n2, n3), n1 <= n2 <= n3
  std::vector<int> tri_base;   // nr: index of the first triple with this n1
  std::vector<int> ls_l;       // n_ls x 3
  std::vector<int> term_off;   // n_ls + 1
  std::vector<int> term_lm;    // n_terms x 3 (lm of each factor)
  std::vector<double> term_w;  // n_terms
  // Gradient lists: for each l-set ls, harmonic lm and position p, the
  // terms of ls whose factor p has that lm; entries grouped by
  // (ls H + lm) 3 + p, terms in increasing order.
  std::vector<int> gls_off;    // n_ls 3 H + 1
  std::vector<int> gls_term;
  std::vector<int> tri_pk;     // n_tri: n1 | n2 << 8 | n3 << 16
  // Radial pairs n1 <= n2, ordered by n2 then n1 (pair n2 (n2 + 1) / 2 + n1,
  // so the pairs with n2 <= n form a prefix): n1 | n2 << 8, and the index
  // of the triple (n1, n2, n2) (the triples (n1, n2, n3 >= n2) follow it).
  int npair = 0;
  std::vector<int> pair_pk, pair_tb;
  // Couplings of the real coefficients per l-set: with c = U r the complex
  // coefficients of the real ones r (the map in the header comment),
  //   B3 = sum C[mu1, mu2, mu3] r[n1, l1 mu1] r[n2, l2 mu2] r[n3, l3 mu3],
  //   C = Re sum_terms w3j U_l1[m1, mu1] U_l2[m2, mu2] U_l3[m3, mu3]
  // (B3 is the real part and r is real). Nonzero entries of l-set ls are
  // [rt_off[ls], rt_off[ls + 1]): mu1 | mu2 << 4 | mu3 << 8 (0-based within
  // each l block) and the value.
  std::vector<int> rt_off, rt_pk;
  std::vector<double> rt_w;
};

// Largest l block (2 l_max + 1, l_max <= 4 for B3).
constexpr int kLM = 9;
// Triples per thread of one l-set in the pair kernels: n_tri / n_pair =
// (n + 2) / 3 <= 11 for the n <= 31 radial channels they take.
constexpr int kPre = 12;

static void build_b3_tables(int nr, int lmax, B3Tables &t) {
  t.nr = nr;
  t.H = (lmax + 1) * (lmax + 1);
  t.term_off.push_back(0);
  for (int l1 = 0; l1 <= lmax; l1++)
    for (int l2 = 0; l2 <= lmax; l2++)
      for (int l3 = 0; l3 <= lmax; l3++) {
        if (std::abs(l1 - l2) > l3 || l3 > l1 + l2) continue;
        t.ls_l.insert(t.ls_l.end(), {l1, l2, l3});
        for (int m1 = 0; m1 < 2 * l1 + 1; m1++)
          for (int m2 = 0; m2 < 2 * l2 + 1; m2++) {
            int m3 = l1 + l2 + l3 - m1 - m2;
            if (m3 < 0 || m3 > 2 * l3) continue;
            double w = w3j(l1, l2, l3, m1 - l1, m2 - l2, m3 - l3);
            if (w == 0.0) continue;
            t.term_lm.insert(t.term_lm.end(), {l1 * l1 + m1, l2 * l2 + m2, l3 * l3 + m3});
            t.term_w.push_back(w);
          }
        t.term_off.push_back((int)t.term_w.size());
      }
  t.n_ls = (int)t.ls_l.size() / 3;
  for (int n1 = 0; n1 < nr; n1++) {
    t.tri_base.push_back((int)t.tri.size() / 3);
    for (int n2 = n1; n2 < nr; n2++)
      for (int n3 = n2; n3 < nr; n3++) t.tri.insert(t.tri.end(), {n1, n2, n3});
  }
  t.n_tri = (int)t.tri.size() / 3;
  t.nd = t.n_tri * t.n_ls;
  for (int k = 0; k < t.n_tri; k++)
    t.tri_pk.push_back(t.tri[3 * k] | t.tri[3 * k + 1] << 8 | t.tri[3 * k + 2] << 16);
  t.npair = nr * (nr + 1) / 2;
  for (int n2 = 0; n2 < nr; n2++)
    for (int n1 = 0; n1 <= n2; n1++) {
      t.pair_pk.push_back(n1 | n2 << 8);
      // Triples of n1 are (n1, a, b), n1 <= a <= b < nr, a-major.
      int k = t.tri_base[n1];
      for (int a = n1; a < n2; a++) k += nr - a;
      t.pair_tb.push_back(k);
    }
  // Row k = m + l of U_l: the complex c_l^m as a combination of the real
  // r_l,mu (at most two entries; index mu + l).
  auto urow = [](int l, int k, int *mu, std::complex<double> *v) {
    const double s2 = 0.70710678118654752440;
    int m = k - l;
    if (m == 0) {
      mu[0] = l, v[0] = 1.0, mu[1] = -1;
    } else if (m > 0) {
      double sg = (m & 1) ? -1.0 : 1.0;
      mu[0] = m + l, v[0] = sg * s2;
      mu[1] = -m + l, v[1] = std::complex<double>(0.0, sg * s2);
    } else {
      mu[0] = -m + l, v[0] = s2;
      mu[1] = m + l, v[1] = std::complex<double>(0.0, -s2);
    }
  };
  t.rt_off.push_back(0);
  for (int ls = 0; ls < t.n_ls; ls++) {
    int l[3] = {t.ls_l[3 * ls], t.ls_l[3 * ls + 1], t.ls_l[3 * ls + 2]};
    std::vector<std::complex<double>> T((size_t)kLM * kLM * kLM, 0.0);
    for (int q = t.term_off[ls]; q < t.term_off[ls + 1]; q++) {
      int mu[3][2];
      std::complex<double> v[3][2];
      for (int p = 0; p < 3; p++) urow(l[p], t.term_lm[3 * q + p] - l[p] * l[p], mu[p], v[p]);
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
          for (int k = 0; k < 2; k++) {
            if (mu[0][i] < 0 || mu[1][j] < 0 || mu[2][k] < 0) continue;
            T[((size_t)mu[0][i] * kLM + mu[1][j]) * kLM + mu[2][k]] +=
                t.term_w[q] * v[0][i] * v[1][j] * v[2][k];
          }
    }
    for (int a = 0; a < kLM; a++)
      for (int b = 0; b < kLM; b++)
        for (int c = 0; c < kLM; c++) {
          double w = T[((size_t)a * kLM + b) * kLM + c].real();
          if (std::abs(w) < 1e-13) continue;  // cancelled to rounding
          t.rt_pk.push_back(a | b << 4 | c << 8);
          t.rt_w.push_back(w);
        }
    t.rt_off.push_back((int)t.rt_w.size());
  }
  t.gls_off.push_back(0);
  for (int ls = 0; ls < t.n_ls; ls++) {
    std::vector<std::vector<int>> lists(3 * t.H);
    for (int q = t.term_off[ls]; q < t.term_off[ls + 1]; q++)
      for (int p = 0; p < 3; p++) lists[t.term_lm[3 * q + p] * 3 + p].push_back(q);
    for (auto &v : lists) {
      t.gls_term.insert(t.gls_term.end(), v.begin(), v.end());
      t.gls_off.push_back((int)t.gls_term.size());
    }
  }
}

int b3_size(int nos, int N, int lmax) {
  B3Tables t;
  build_b3_tables(nos * N, lmax, t);
  return t.nd;
}

// Device copies of the tables.
//
// Order of the descriptor on the device: l-set major, k' = ls n_tri + tri
// (the reference's index is k = tri n_ls + ls, b3.cpp:230-242). Dot
// products do not depend on the order, so the contractions run on rows and
// sparse sets stored this way (the model's sparse set is permuted once, on
// its first use with a B3 structure) and the descriptor read-back restores
// the reference's order. With the l-set outermost, the forward kernel's
// lanes share one term list, and the adjoint reads one contiguous slice of
// dE/dB per l-set.
struct B3Dev {
  const int *tri, *ls_l, *term_off, *term_lm, *gls_off, *gls_term, *tri_base, *tri_pk;
  const double *term_w;
  int nr, H, n_ls, n_tri, nd;
  const int *pair_pk, *pair_tb;
  int npair;
};

// Complex c of one atom in LDS from the real coefficients (interleaved
// re, im; index n H + lm with lm = l^2 + m + l).
template <typename T>
__device__ void complex_c(const T *cr, T *cc, int nr, int H, int lane) {
  const T s2 = T(0.70710678118654752440);
  for (int k = lane; k < nr * H; k += 64) {
    int n = k / H, lm = k - n * H;
    int l = 0;
    while ((l + 1) * (l + 1) <= lm) l++;
    int m = lm - l * l - l;
    const T *row = cr + n * H + l * l + l;
    T re, im;
    if (m == 0) {
      re = row[0];
      im = 0;
    } else if (m > 0) {
      T sg = (m & 1) ? T(-1) : T(1);
      re = sg * s2 * row[m];
      im = sg * s2 * row[-m];
    } else {
      re = s2 * row[-m];
      im = -s2 * row[m];
    }
    cc[2 * k] = re;
    cc[2 * k + 1] = im;
  }
}

// B3 per atom: one wavefront, complex c in LDS. For each l-set (uniform
// across the wave, so its term list is read once per wave), lanes take the
// radial triples and write the row in l-set-major order.
// Rows of one call: every atom (aor null: wave i takes atom i and writes its
// species-sorted row), or the species-sorted rows [r0, r0 + n) (aor =
// atom_of_row: wave k takes row r0 + k and writes row k of B, a row tile).
struct RowRange {
  int r0;
  const int *aor;
  __device__ __forceinline__ void map(int k, const int *row_of_atom, int &atom, int &row,
                                      int &brow) const {
    if (aor) {
      row = r0 + k;
      atom = aor[row];
      brow = k;
    } else {
      atom = k;
      row = row_of_atom[k];
      brow = row;
    }
  }
};

template <typename T>
__global__ __launch_bounds__(256) void k_b3(int n, B3Dev t, const T *Cg, const int *row_of_atom,
                                             RowRange rr, T *B, double *norm, double *nnv) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int k = blockIdx.x * (blockDim.x >> 6) + w;
  if (k >= n) return;
  int i, row, brow;
  rr.map(k, row_of_atom, i, row, brow);
  int nc = t.nr * t.H, H = t.H;
  T *cc = reinterpret_cast<T *>(smem) + (size_t)w * 2 * nc;
  complex_c<T>(Cg + (size_t)i * nc, cc, t.nr, t.H, lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  T *Brow = B + (size_t)brow * t.nd;
  T s2 = 0;
  for (int ls = 0; ls < t.n_ls; ls++) {
    int q0 = t.term_off[ls], q1 = t.term_off[ls + 1];
    for (int tr = lane; tr < t.n_tri; tr += 64) {
      int pk = t.tri_pk[tr];
      const T *c1 = cc + 2 * (pk & 255) * H, *c2 = cc + 2 * ((pk >> 8) & 255) * H;
      const T *c3 = cc + 2 * (pk >> 16) * H;
      T acc = 0;
      for (int q = q0; q < q1; q++) {
        int a1 = t.term_lm[3 * q], a2 = t.term_lm[3 * q + 1], a3 = t.term_lm[3 * q + 2];
        T xr = c1[2 * a1], xi = c1[2 * a1 + 1], yr = c2[2 * a2], yi = c2[2 * a2 + 1];
        T zr = c3[2 * a3], zi = c3[2 * a3 + 1];
        T pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
        acc += (T)t.term_w[q] * (pr * zr - pi * zi);
      }
      int kd = ls * t.n_tri + tr;
      Brow[kd] = acc;
      s2 += acc * acc;
    }
  }
  s2 = wave_sum(s2);
  if (lane == 0) {
    norm[row] = (double)sqrt(s2);
    nnv[row] = (double)s2;
  }
}

__device__ __forceinline__ int tri_index(const B3Dev &t, int n1, int n2, int n3) {
  int nr = t.nr;
  return t.tri_base[n1] + (n2 - n1) * nr - (n2 * (n2 - 1) - n1 * (n1 - 1)) / 2 + (n3 - n2);
}

// dE/dc for B3 per atom. With `project`, w_k = (G_k - (G . B_hat) B_k / |B|)
// / |B| (as for B2). For each l-set the slice w[ls n_tri, (ls + 1) n_tri)
// is formed from contiguous G and B reads into LDS; lanes then take the
// targets (n, lm) and add, over each factor position p holding (n, lm) and
// the terms of that l-set, w w3j conj(product of the other two factors)
// (radial triples enumerated in a fixed order, so results are
// deterministic). The real gradient u = dE/dc^r follows by the transpose of
// the complex map.
template <typename T>
__global__ __launch_bounds__(256) void k_dEdc_b3(int n, B3Dev t, const T *Cg,
                                                  const int *row_of_atom, RowRange rr,
                                                  const T *B, const double *norm,
                                                  const double *nnv, const T *G, long long g_ld,
                                                  int project, double empty_norm,
                                                  double empty_nn, T *Ug) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int kw = blockIdx.x * (blockDim.x >> 6) + w;
  int nc = t.nr * t.H, nr = t.nr, H = t.H;
  int per = 4 * nc + t.n_tri;
  // Block-shared: index of the first triple of each n1.
  int *tb = reinterpret_cast<int *>(smem);
  for (int k = threadIdx.x; k < nr; k += blockDim.x) tb[k] = t.tri_base[k];
  __syncthreads();
  if (kw >= n) return;
  int i, row, brow;
  rr.map(kw, row_of_atom, i, row, brow);
  T *cc = reinterpret_cast<T *>(smem + 256) + (size_t)w * per;
  T *gc = cc + 2 * nc;  // complex gradient, one (re, im) per target
  T *wl = gc + 2 * nc;  // w slice of one l-set
  double nrm = norm[row];
  if (nrm < empty_norm || nnv[row] < empty_nn) return;
  complex_c<T>(Cg + (size_t)i * nc, cc, t.nr, t.H, lane);
  for (int k = lane; k < 2 * nc; k += 64) gc[k] = 0;
  const T *Brow = B + (size_t)brow * t.nd;
  const T *Grow = G + (size_t)brow * g_ld;
  T ninv = T(1) / (T)nrm, gb = 0;
  if (project) {
#pragma unroll 8
    for (int k = lane; k < t.nd; k += 64) gb += Grow[k] * Brow[k];
    gb = wave_sum(gb) * ninv;
  }
  for (int ls = 0; ls < t.n_ls; ls++) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const T *gs = Grow + (size_t)ls * t.n_tri, *bs = Brow + (size_t)ls * t.n_tri;
    for (int k = lane; k < t.n_tri; k += 64)
      wl[k] = project ? (gs[k] - gb * bs[k] * ninv) * ninv : gs[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Per factor position p, only the targets whose lm lies in l_p's block
    // have terms in this l-set: lanes take those (n, m) pairs. A target met
    // at several positions is updated in consecutive passes (one wave, LDS
    // operations in program order).
    for (int p = 0; p < 3; p++) {
      int lp = t.ls_l[3 * ls + p], nm = 2 * lp + 1;
      for (int idx = lane; idx < nr * nm; idx += 64) {
        int nn = idx / nm, lm = lp * lp + (idx - nn * nm);
        int g = nn * H + lm;
        T xr = 0, xi = 0;
        int base = ((ls * H + lm) * 3 + p);
        int e0 = t.gls_off[base], e1 = t.gls_off[base + 1];
        for (int e = e0; e < e1; e++) {
          int q = t.gls_term[e];
          T wq = (T)t.term_w[q];
          int a0 = t.term_lm[3 * q], a1 = t.term_lm[3 * q + 1], a2 = t.term_lm[3 * q + 2];
          // The two other factors' lm.
          int oa = p == 0 ? a1 : a0, ob = p == 2 ? a1 : a2;
          // Triples with factor p = nn: (nn, x, y), (x, nn, y) or (x, y, nn).
          int lo_x = p == 0 ? nn : 0, hi_x = p == 0 ? nr - 1 : nn;
          for (int x = lo_x; x <= hi_x; x++) {
            int lo_y = p == 0 ? x : (p == 1 ? nn : x);
            int hi_y = p == 2 ? nn : nr - 1;
            const T *u = cc + 2 * (x * H + oa);
            for (int y = lo_y; y <= hi_y; y++) {
              int n1 = p == 0 ? nn : x, n2 = p == 0 ? x : (p == 1 ? nn : y);
              int n3 = p == 2 ? nn : y;
              int tri = tb[n1] + (n2 - n1) * nr - (n2 * (n2 - 1) - n1 * (n1 - 1)) / 2 + (n3 - n2);
              T wkv = wl[tri] * wq;
              const T *v = cc + 2 * (y * H + ob);
              T pr = u[0] * v[0] - u[1] * v[1], pi = u[0] * v[1] + u[1] * v[0];
              xr += wkv * pr;
              xi -= wkv * pi;
            }
          }
        }
        gc[2 * g] += xr;
        gc[2 * g + 1] += xi;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // u[n, l, m] = d E / d c^r: transpose of the complex map.
  const T s2 = T(0.70710678118654752440);
  T *urow = Ug + (size_t)i * nc;
  for (int g = lane; g < nc; g += 64) {
    int nn = g / H, lm = g - nn * H;
    int l = 0;
    while ((l + 1) * (l + 1) <= lm) l++;
    int m = lm - l * l - l;
    const T *gr = gc + 2 * (nn * H + l * l + l);
    T u;
    if (m == 0) {
      u = gr[0];
    } else {
      int mm = m > 0 ? m : -m;
      T sg = (mm & 1) ? T(-1) : T(1);
      if (m > 0)  // c^r_lm enters Re c_l^m and Re c_l^-m
        u = s2 * (sg * gr[2 * mm] + gr[-2 * mm]);
      else        // c^r_l,-m enters Im c_l^m and Im c_l^-m
        u = s2 * (sg * gr[2 * mm + 1] - gr[-2 * mm + 1]);
    }
    urow[g] = u;
  }
}

// ---------------------------------------------------------------------------
// Pair-factorized B3 on the real coefficients (one workgroup per atom, one
// thread per radial pair n1 <= n2).
//
// With C the real couplings of the l-set (B3Tables::rt_*), for each pair
//   P[mu3] = sum_{(mu1, mu2, mu3) in C} C r[n1, l1 mu1] r[n2, l2 mu2]
// and B(n1, n2, n3) = sum_mu3 r[n3, l3 mu3] P[mu3] for n3 >= n2: the triple
// sum over couplings becomes a pair sum plus one dot product per triple.
// The adjoint runs the same factors backwards: with w = dE/dB of the l-set,
//   D[mu3] = sum_{n3 >= n2} w(n1, n2, n3) r[n3, l3 mu3],
//   u[n3, l3 mu3] += sum_{n1 <= n2 <= n3} w(n1, n2, n3) P(n1, n2)[mu3],
//   u[n1, l1 mu1] += sum_{n2 >= n1} sum_C C r[n2, l2 mu2] D[mu3],
//   u[n2, l2 mu2] += sum_{n1 <= n2} sum_C C r[n1, l1 mu1] D[mu3].
// The coupling list is read at uniform addresses and indexes small register
// arrays (uniform indices, so no scratch); the sums over pairs that end on
// one target (n, mu) run with one thread per target after a barrier.
// Summation orders are fixed, so results are deterministic.
// ---------------------------------------------------------------------------

// Sum over the block in a fixed order (red: one slot per wave).
template <typename T>
__device__ T block_sum(T v, T *red) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  for (int k = 0; k < nw; k++) s += red[k];
  __syncthreads();
  return s;
}

// One l block of r (2 l + 1 values) into registers.
template <typename T>
__device__ __forceinline__ void load_block(const T *r, int nm, T *v) {
#pragma unroll
  for (int k = 0; k < kLM; k++) v[k] = k < nm ? r[k] : T(0);
}

template <typename T>
__global__ __launch_bounds__(512) void k_b3p(int n, B3Dev t, const int *__restrict__ rt_off,
                                              const int *__restrict__ rt_pk,
                                              const T *__restrict__ rt_w, const T *Cg,
                                              const int *row_of_atom, RowRange rr, T *B,
                                              double *norm, double *nnv) {
  extern __shared__ __align__(16) unsigned char smem[];
  int k = blockIdx.x;
  if (k >= n) return;
  int tid = threadIdx.x, nt = blockDim.x;
  int i, row, brow;
  rr.map(k, row_of_atom, i, row, brow);
  int nr = t.nr, H = t.H, nc = nr * H, n_tri = t.n_tri;
  T *rc = reinterpret_cast<T *>(smem);  // r of the atom
  T *bl = rc + nc;                        // one l-set's slice of the row
  T *red = bl + n_tri;                    // 16 wave slots
  for (int g = tid; g < nc; g += nt) rc[g] = Cg[(size_t)i * nc + g];
  bool act = tid < t.npair;
  int n1 = 0, n2 = 0, tb = 0;
  if (act) {
    int pk = t.pair_pk[tid];
    n1 = pk & 255;
    n2 = pk >> 8;
    tb = t.pair_tb[tid];
  }
  __syncthreads();
  T s2 = 0;
  T *Brow = B + (size_t)brow * t.nd;
  for (int ls = 0; ls < t.n_ls; ls++) {
    int l1 = t.ls_l[3 * ls], l2 = t.ls_l[3 * ls + 1], l3 = t.ls_l[3 * ls + 2];
    if (act) {
      T x[kLM], y[kLM], p[kLM];
      load_block<T>(rc + n1 * H + l1 * l1, 2 * l1 + 1, x);
      load_block<T>(rc + n2 * H + l2 * l2, 2 * l2 + 1, y);
#pragma unroll
      for (int m = 0; m < kLM; m++) p[m] = 0;
      // Couplings in order, their entries loaded four at a time (one wait
      // for the scalar loads per four terms).
#define FM_B3_FWD_TERM(c, w) p[(c) >> 8] += (w) * x[(c) & 15] * y[((c) >> 4) & 15]
      int q = rt_off[ls], q1 = rt_off[ls + 1];
      for (; q + 4 <= q1; q += 4) {
        int c0 = rt_pk[q], c1 = rt_pk[q + 1], c2 = rt_pk[q + 2], c3 = rt_pk[q + 3];
        T w0 = rt_w[q], w1 = rt_w[q + 1], w2 = rt_w[q + 2], w3 = rt_w[q + 3];
        FM_B3_FWD_TERM(c0, w0);
        FM_B3_FWD_TERM(c1, w1);
        FM_B3_FWD_TERM(c2, w2);
        FM_B3_FWD_TERM(c3, w3);
      }
      for (; q < q1; q++) FM_B3_FWD_TERM(rt_pk[q], rt_w[q]);
#undef FM_B3_FWD_TERM
      int n3m = 2 * l3 + 1;
      for (int n3 = n2; n3 < nr; n3++) {
        const T *z = rc + n3 * H + l3 * l3;
        T acc = 0;
#pragma unroll
        for (int m = 0; m < kLM; m++)
          if (m < n3m) acc += z[m] * p[m];
        bl[tb + n3 - n2] = acc;
      }
    }
    __syncthreads();
    for (int kk = tid; kk < n_tri; kk += nt) {
      T v = bl[kk];
      size_t kd = (size_t)ls * n_tri + kk;
      Brow[kd] = v;
      s2 += v * v;
    }
    __syncthreads();
  }
  s2 = block_sum(s2, red);
  if (tid == 0) {
    norm[row] = (double)sqrt(s2);
    nnv[row] = (double)s2;
  }
}

template <typename T>
__global__ __launch_bounds__(512) void k_dEdc_b3p(int n, B3Dev t, const int *__restrict__ rt_off,
                                                   const int *__restrict__ rt_pk,
                                                   const T *__restrict__ rt_w, const T *Cg,
                                                   const int *row_of_atom, RowRange rr,
                                                   const T *B, const double *norm,
                                                   const double *nnv, const T *G,
                                                   long long g_ld, int project,
                                                   double empty_norm, double empty_nn, T *Ug) {
  extern __shared__ __align__(16) unsigned char smem[];
  int k = blockIdx.x;
  if (k >= n) return;
  int tid = threadIdx.x, nt = blockDim.x;
  int i, row, brow;
  rr.map(k, row_of_atom, i, row, brow);
  int nr = t.nr, H = t.H, nc = nr * H, n_tri = t.n_tri, npair = t.npair;
  T *urow = Ug + (size_t)i * nc;
  double nrm = norm[row];
  if (nrm < empty_norm || nnv[row] < empty_nn) {
    for (int g = tid; g < nc; g += nt) urow[g] = 0;
    return;
  }
  T *rc = reinterpret_cast<T *>(smem);  // r of the atom
  T *gc = rc + nc;                        // dE/dr
  T *wl = gc + nc;                        // w of one l-set
  T *Pl = wl + n_tri;                     // P, then Q: [mu][pair]
  T *El = Pl + (size_t)kLM * npair;       // factor-1, then factor-2 partials
  T *red = El + (size_t)kLM * npair;      // 16 wave slots
  for (int g = tid; g < nc; g += nt) {
    rc[g] = Cg[(size_t)i * nc + g];
    gc[g] = 0;
  }
  const T *Brow = B + (size_t)brow * t.nd;
  const T *Grow = G + (size_t)brow * g_ld;
  T ninv = T(1) / (T)nrm, gb = 0;
  if (project) {
    T part = 0;
#pragma unroll 16
    for (int kk = tid; kk < t.nd; kk += nt) part += Grow[kk] * Brow[kk];
    gb = block_sum(part, red) * ninv;
  }
  // FP32: the next l-set's slices of dE/dB and B are loaded into registers
  // while the current l-set is processed (each thread holds at most kPre).
  // FP64 reads them when staging (its registers go to the coupling sweep).
  constexpr int NQ = sizeof(T) == 4 ? kPre : 1;
  T gq[NQ], bq[NQ];
  auto fetch = [&](int ls) {
    if constexpr (sizeof(T) == 4) {
      const T *gs = Grow + (size_t)ls * n_tri, *bs = Brow + (size_t)ls * n_tri;
#pragma unroll
      for (int j = 0; j < NQ; j++) {
        int kk = tid + j * nt;
        gq[j] = kk < n_tri ? gs[kk] : T(0);
        bq[j] = (project && kk < n_tri) ? bs[kk] : T(0);
      }
    }
  };
  fetch(0);
  bool act = tid < npair;
  int n1 = 0, n2 = 0, tb = 0;
  if (act) {
    int pk = t.pair_pk[tid];
    n1 = pk & 255;
    n2 = pk >> 8;
    tb = t.pair_tb[tid];
  }
  for (int ls = 0; ls < t.n_ls; ls++) {
    int l1 = t.ls_l[3 * ls], l2 = t.ls_l[3 * ls + 1], l3 = t.ls_l[3 * ls + 2];
    int n1m = 2 * l1 + 1, n2m = 2 * l2 + 1, n3m = 2 * l3 + 1;
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int j = 0; j < NQ; j++) {
        int kk = tid + j * nt;
        if (kk < n_tri) wl[kk] = project ? (gq[j] - gb * bq[j] * ninv) * ninv : gq[j];
      }
    } else {
      const T *gs = Grow + (size_t)ls * n_tri, *bs = Brow + (size_t)ls * n_tri;
#pragma unroll 4
      for (int kk = tid; kk < n_tri; kk += nt)
        wl[kk] = project ? (gs[kk] - gb * bs[kk] * ninv) * ninv : gs[kk];
    }
    __syncthreads();
    if (ls + 1 < t.n_ls) fetch(ls + 1);
    T e2[kLM];
    if (act) {
      T x[kLM], y[kLM], d[kLM], p[kLM], e1[kLM];
      load_block<T>(rc + n1 * H + l1 * l1, n1m, x);
      load_block<T>(rc + n2 * H + l2 * l2, n2m, y);
#pragma unroll
      for (int m = 0; m < kLM; m++) d[m] = p[m] = e1[m] = e2[m] = 0;
      for (int n3 = n2; n3 < nr; n3++) {
        T wv = wl[tb + n3 - n2];
        const T *z = rc + n3 * H + l3 * l3;
#pragma unroll
        for (int m = 0; m < kLM; m++)
          if (m < n3m) d[m] += wv * z[m];
      }
      // Couplings in order, their entries loaded four at a time (one wait
      // for the scalar loads per four terms).
#define FM_B3_ADJ_TERM(c, w)                              \
  {                                                       \
    int a_ = (c) & 15, b_ = ((c) >> 4) & 15, e_ = (c) >> 8; \
    T wx = (w) * x[a_], wy = (w) * y[b_];                 \
    p[e_] += wx * y[b_];                                  \
    e1[a_] += wy * d[e_];                                 \
    e2[b_] += wx * d[e_];                                 \
  }
      int q = rt_off[ls], q1 = rt_off[ls + 1];
      for (; q + 4 <= q1; q += 4) {
        int c0 = rt_pk[q], c1 = rt_pk[q + 1], c2 = rt_pk[q + 2], c3 = rt_pk[q + 3];
        T w0 = rt_w[q], w1 = rt_w[q + 1], w2 = rt_w[q + 2], w3 = rt_w[q + 3];
        FM_B3_ADJ_TERM(c0, w0);
        FM_B3_ADJ_TERM(c1, w1);
        FM_B3_ADJ_TERM(c2, w2);
        FM_B3_ADJ_TERM(c3, w3);
      }
      for (; q < q1; q++) FM_B3_ADJ_TERM(rt_pk[q], rt_w[q]);
#undef FM_B3_ADJ_TERM
#pragma unroll
      for (int m = 0; m < kLM; m++) {
        if (m < n3m) Pl[m * npair + tid] = p[m];
        if (m < n1m) El[m * npair + tid] = e1[m];
      }
    }
    __syncthreads();
    // Factor 1 targets (n, l1 mu1): pairs (n, n2 >= n).
    for (int tg = tid; tg < nr * n1m; tg += nt) {
      int nn = tg / n1m, m1 = tg - nn * n1m;
      T acc = 0;
      for (int b = nn; b < nr; b++) acc += El[m1 * npair + b * (b + 1) / 2 + nn];
      gc[nn * H + l1 * l1 + m1] += acc;
    }
    // Factor 3, first stage: the thread of pair (a, b) forms
    // Q[mu3] = sum_{n1 <= a} w(n1, a, b) P(n1, a)[mu3]; the triple
    // (n1, a, b) and the pair (n1, a) are stepped along n1.
    T qv[kLM];
    if (act) {
#pragma unroll
      for (int m = 0; m < kLM; m++) qv[m] = 0;
      int a = n1, b = n2;
      int idx = a * nr - a * (a - 1) / 2 + b - a;
      const T *pp = Pl + a * (a + 1) / 2;
      for (int c = 0; c <= a; c++) {
        T wv = wl[idx];
#pragma unroll
        for (int m = 0; m < kLM; m++)
          if (m < n3m) qv[m] += wv * pp[(size_t)m * npair];
        idx += (nr - c) * (nr - c - 1) / 2;
        pp += 1;
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int m = 0; m < kLM; m++) {
        if (m < n3m) Pl[m * npair + tid] = qv[m];
        if (m < n2m) El[m * npair + tid] = e2[m];
      }
    }
    __syncthreads();
    // Factor 3 targets (n, l3 mu3): Q over the pairs (n2 <= n, n), and factor
    // 2 targets (n, l2 mu2): E2 over the pairs (n1 <= n, n); both contiguous.
    // With l2 == l3 the targets coincide and one thread adds both.
    for (int tg = tid; tg < nr * n3m; tg += nt) {
      int nn = tg / n3m, m3 = tg - nn * n3m;
      const T *e = Pl + m3 * npair + nn * (nn + 1) / 2;
      T acc = 0;
      for (int a = 0; a <= nn; a++) acc += e[a];
      if (l2 == l3) {
        const T *f = El + m3 * npair + nn * (nn + 1) / 2;
        for (int a = 0; a <= nn; a++) acc += f[a];
      }
      gc[nn * H + l3 * l3 + m3] += acc;
    }
    if (l2 != l3)
      for (int tg = tid; tg < nr * n2m; tg += nt) {
        int nn = tg / n2m, m2 = tg - nn * n2m;
        const T *e = El + m2 * npair + nn * (nn + 1) / 2;
        T acc = 0;
        for (int a = 0; a <= nn; a++) acc += e[a];
        gc[nn * H + l2 * l2 + m2] += acc;
      }
    __syncthreads();
  }
  for (int g = tid; g < nc; g += nt) urow[g] = gc[g];
}

// Column / row permutations between the reference's B3 order and the
// l-set-major one: out[r][k'] = in[r][src(k')] for k' < nd (columns past nd,
// the padding, are copied), or out row k' = in row src(k'), with
// src(ls n_tri + tri) = tri n_ls + ls (to_device) or its inverse.
__device__ __forceinline__ long long b3_src(long long kp, int n_tri, int n_ls, bool to_device) {
  if (to_device) {
    long long ls = kp / n_tri, tri = kp - ls * n_tri;
    return tri * n_ls + ls;
  }
  long long tri = kp / n_ls, ls = kp - tri * n_ls;
  return ls * n_tri + tri;
}

template <typename E>
__global__ void k_b3_permute_cols(long long rows, long long ld, int nd, int n_tri, int n_ls,
                                  int to_device, const E *in, E *out) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= rows * ld) return;
  long long r = idx / ld, k = idx - r * ld;
  long long src = k < nd ? b3_src(k, n_tri, n_ls, to_device) : k;
  out[idx] = in[r * ld + src];
}

template <typename E>
__global__ void k_b3_permute_rows(int nd, long long ld, int n_tri, int n_ls, int to_device,
                                  const E *in, E *out) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nd * ld) return;
  long long k = idx / ld, c = idx - k * ld;
  out[idx] = in[b3_src(k, n_tri, n_ls, to_device) * ld + c];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct B3State {
  B3Tables t;
  DevBuf tri, tri_base, ls_l, term_off, term_lm, term_w, gls_off, gls_term, tri_pk;
  DevBuf pair_pk, pair_tb, rt_off, rt_pk, rt_w64, rt_w32;
  B3Dev dev() const {
    return B3Dev{tri.as<int>(),      ls_l.as<int>(),     term_off.as<int>(),
                 term_lm.as<int>(),  gls_off.as<int>(),  gls_term.as<int>(),
                 tri_base.as<int>(), tri_pk.as<int>(),   term_w.as<double>(),
                 t.nr,               t.H,                t.n_ls,
                 t.n_tri,            t.nd,               pair_pk.as<int>(),
                 pair_tb.as<int>(),  t.npair};
  }
  template <typename T>
  const T *rt_w() const {
    return std::is_same<T, double>::value ? (const T *)rt_w64.as<double>()
                                          : (const T *)rt_w32.as<float>();
  }};

static B3State *b3_state(fm_structure s) {
  if (!s->b3) {
    auto *st = new B3State();
    build_b3_tables(s->nos * s->N, s->lmax, st->t);
    hipStream_t stream = s->ctx->stream;
    Pool *pool = &s->ctx->pool;
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
      b.alloc(pool, std::max<size_t>(1, bytes));
      if (bytes) FM_HIP(hipMemcpyAsync(b.as<void>(), src, bytes, hipMemcpyHostToDevice, stream));
    };
    const B3Tables &t = st->t;
    up(st->tri, t.tri.data(), t.tri.size() * 4);
    up(st->tri_base, t.tri_base.data(), t.tri_base.size() * 4);
    up(st->ls_l, t.ls_l.data(), t.ls_l.size() * 4);
    up(st->term_off, t.term_off.data(), t.term_off.size() * 4);
    up(st->term_lm, t.term_lm.data(), t.term_lm.size() * 4);
    up(st->term_w, t.term_w.data(), t.term_w.size() * 8);
    up(st->gls_off, t.gls_off.data(), t.gls_off.size() * 4);
    up(st->gls_term, t.gls_term.data(), t.gls_term.size() * 4);
    up(st->tri_pk, t.tri_pk.data(), t.tri_pk.size() * 4);
    up(st->pair_pk, t.pair_pk.data(), t.pair_pk.size() * 4);
    up(st->pair_tb, t.pair_tb.data(), t.pair_tb.size() * 4);
    up(st->rt_off, t.rt_off.data(), t.rt_off.size() * 4);
    up(st->rt_pk, t.rt_pk.data(), t.rt_pk.size() * 4);
    up(st->rt_w64, t.rt_w.data(), t.rt_w.size() * 8);
    std::vector<float> w32(t.rt_w.begin(), t.rt_w.end());
    up(st->rt_w32, w32.data(), w32.size() * 4);
    FM_HIP(hipStreamSynchronize(stream));
    s->b3 = st;
  }
  return s->b3;
}

namespace {

template <typename E>
void permute_cols(hipStream_t st, DevBuf &buf, Pool *pool, long long rows, long long ld, int nd,
                  int n_tri, int n_ls, int to_device) {
  long long total = rows * ld;
  if (total == 0) return;
  DevBuf out(pool, (size_t)total * sizeof(E));
  hipLaunchKernelGGL(k_b3_permute_cols<E>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     st, rows, ld, nd, n_tri, n_ls, to_device, buf.as<E>(), out.as<E>());
  FM_HIP(hipGetLastError());
  std::swap(buf, out);
}

template <typename E>
void permute_rows(hipStream_t st, DevBuf &buf, Pool *pool, int nd, long long ld, int n_tri,
                  int n_ls, int to_device) {
  long long total = (long long)nd * ld;
  if (total == 0) return;
  DevBuf out(pool, (size_t)total * sizeof(E));
  hipLaunchKernelGGL(k_b3_permute_rows<E>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     st, nd, ld, n_tri, n_ls, to_device, buf.as<E>(), out.as<E>());
  FM_HIP(hipGetLastError());
  std::swap(buf, out);
}

}  // namespace

void b3_order_model(fm_model m, fm_structure s) {
  B3State *b = b3_state(s);
  const B3Tables &t = b->t;
  if (m->b3_ntri == t.n_tri && m->b3_nls == t.n_ls) return;
  FM_CHECK(m->b3_ntri == 0, FM_ERR_INVALID,
           "the model's sparse set is in the order of another B3 setting");
  FM_CHECK(m->nd == t.nd, FM_ERR_INVALID, "model and B3 descriptor sizes disagree");
  // The model lives on its own context's stream.
  hipStream_t st = m->ctx->stream;
  Pool *pool = &m->ctx->pool;
  int nd = t.nd;
  for (int sp = 0; sp < m->nos; sp++) {
    long long Ms = m->M_by_type[sp];
    if (Ms == 0) continue;
    permute_cols<double>(st, m->X[sp], pool, Ms, nd, nd, t.n_tri, t.n_ls, 1);
    permute_rows<double>(st, m->XT[sp], pool, nd, Ms, t.n_tri, t.n_ls, 1);
    permute_cols<uint16_t>(st, m->Xs0[sp], pool, Ms, split_ld(nd), nd, t.n_tri, t.n_ls, 1);
    permute_cols<uint16_t>(st, m->Xs1[sp], pool, Ms, split_ld(nd), nd, t.n_tri, t.n_ls, 1);
    permute_rows<uint16_t>(st, m->XTs0[sp], pool, nd, split_ld(Ms), t.n_tri, t.n_ls, 1);
    permute_rows<uint16_t>(st, m->XTs1[sp], pool, nd, split_ld(Ms), t.n_tri, t.n_ls, 1);
  }
  FM_HIP(hipStreamSynchronize(st));
  m->b3_ntri = t.n_tri;
  m->b3_nls = t.n_ls;
  m->xq_ready = false;  // slices follow X's new column order
}

void b3_rows_to_reference(fm_structure s, int r0, int rows, void *out) {
  B3State *b = b3_state(s);
  const B3Tables &t = b->t;
  hipStream_t st = s->ctx->stream;
  long long total = (long long)rows * t.nd;
  if (total == 0) return;
  dim3 grid((unsigned)((total + 255) / 256));
  if (s->form == FM_FORM_FP64)
    hipLaunchKernelGGL(k_b3_permute_cols<double>, grid, dim3(256), 0, st, (long long)rows,
                       (long long)t.nd, t.nd, t.n_tri, t.n_ls, 0,
                       s->B.as<double>() + (size_t)r0 * t.nd, (double *)out);
  else
    hipLaunchKernelGGL(k_b3_permute_cols<float>, grid, dim3(256), 0, st, (long long)rows,
                       (long long)t.nd, t.nd, t.n_tri, t.n_ls, 0,
                       s->B.as<float>() + (size_t)r0 * t.nd, (float *)out);
  FM_HIP(hipGetLastError());
}

void free_b3_state(fm_structure s) {
  delete s->b3;
  s->b3 = nullptr;
}

static int waves_per_block(size_t lds_per_atom) {
  const size_t cap = 160 * 1024;
  FM_CHECK(lds_per_atom <= cap, FM_ERR_CAPACITY, "per-atom B3 workspace exceeds LDS");
  int apb = 4;
  while (apb > 1 && apb * lds_per_atom > cap) apb >>= 1;
  return apb;
}

// Workgroup size and LDS bytes of the pair kernels, or 0 when the pairs do
// not fit one workgroup of 512 threads or the workspace does not fit LDS
// (then the wave-per-atom kernels run).
static size_t b3_pair_lds(const B3Dev &d, size_t esz, bool adjoint, int &nt) {
  if (d.npair > 512) return 0;
  nt = std::max(64, (d.npair + 63) / 64 * 64);
  if (d.n_tri > kPre * nt) return 0;
  size_t nc = (size_t)d.nr * d.H;
  size_t el = nc + d.n_tri + 16;
  if (adjoint) el += nc + 2 * (size_t)kLM * d.npair;
  size_t bytes = el * esz;
  return bytes <= 160 * 1024 ? bytes : 0;
}

void launch_b3(fm_structure s) {
  launch_b3_rows(s, 0, s->n, false, s->B.as<void>());
}

void launch_b3_rows(fm_structure s, int r0, int n, bool tile, void *Bout) {
  hipStream_t st = s->ctx->stream;
  B3State *b = b3_state(s);
  B3Dev d = b->dev();
  if (n == 0) return;
  RowRange rr{r0, tile ? s->atom_of_row.as<int>() : nullptr};
  bool f64 = s->form == FM_FORM_FP64;
  size_t esz = f64 ? 8 : 4;
  int nt = 0;
  size_t lds = b3_pair_lds(d, esz, false, nt);
  if (lds) {
    if (f64)
      hipLaunchKernelGGL(k_b3p<double>, dim3(n), dim3(nt), lds, st, n, d, b->rt_off.as<int>(),
                         b->rt_pk.as<int>(), b->rt_w<double>(), s->C.as<double>(),
                         s->row_of_atom.as<int>(), rr, (double *)Bout, s->norm.as<double>(),
                         s->nn.as<double>());
    else
      hipLaunchKernelGGL(k_b3p<float>, dim3(n), dim3(nt), lds, st, n, d, b->rt_off.as<int>(),
                         b->rt_pk.as<int>(), b->rt_w<float>(), s->C.as<float>(),
                         s->row_of_atom.as<int>(), rr, (float *)Bout, s->norm.as<double>(),
                         s->nn.as<double>());
    FM_HIP(hipGetLastError());
    return;
  }
  size_t per = 2 * (size_t)d.nr * d.H * esz;
  int apb = waves_per_block(per);
  int nblk = (n + apb - 1) / apb;
  if (f64)
    hipLaunchKernelGGL(k_b3<double>, dim3(nblk), dim3(64 * apb), apb * per, st, n, d,
                       s->C.as<double>(), s->row_of_atom.as<int>(), rr, (double *)Bout,
                       s->norm.as<double>(), s->nn.as<double>());
  else
    hipLaunchKernelGGL(k_b3<float>, dim3(nblk), dim3(64 * apb), apb * per, st, n, d,
                       s->C.as<float>(), s->row_of_atom.as<int>(), rr, (float *)Bout,
                       s->norm.as<double>(), s->nn.as<double>());
  FM_HIP(hipGetLastError());
}

void launch_dEdc_b3(fm_structure s, const void *G, int64_t g_ld, int project,
                    double empty_norm, double empty_nn, void *U) {
  launch_dEdc_b3_rows(s, 0, s->n, false, s->B.as<void>(), G, g_ld, project, empty_norm,
                      empty_nn, U);
}

void launch_dEdc_b3_rows(fm_structure s, int r0, int n, bool tile, const void *B,
                         const void *G, int64_t g_ld, int project, double empty_norm,
                         double empty_nn, void *U) {
  hipStream_t st = s->ctx->stream;
  B3State *b = b3_state(s);
  B3Dev d = b->dev();
  if (n == 0) return;
  RowRange rr{r0, tile ? s->atom_of_row.as<int>() : nullptr};
  bool f64 = s->form == FM_FORM_FP64;
  size_t esz = f64 ? 8 : 4;
  int nt = 0;
  size_t plds = b3_pair_lds(d, esz, true, nt);
  if (plds) {
    if (f64)
      hipLaunchKernelGGL(k_dEdc_b3p<double>, dim3(n), dim3(nt), plds, st, n, d,
                         b->rt_off.as<int>(), b->rt_pk.as<int>(), b->rt_w<double>(),
                         s->C.as<double>(),
                         s->row_of_atom.as<int>(), rr, (const double *)B, s->norm.as<double>(),
                         s->nn.as<double>(), (const double *)G, (long long)g_ld, project,
                         empty_norm, empty_nn, (double *)U);
    else
      hipLaunchKernelGGL(k_dEdc_b3p<float>, dim3(n), dim3(nt), plds, st, n, d,
                         b->rt_off.as<int>(), b->rt_pk.as<int>(), b->rt_w<float>(),
                         s->C.as<float>(), s->row_of_atom.as<int>(), rr,
                         (const float *)B, s->norm.as<double>(), s->nn.as<double>(),
                         (const float *)G, (long long)g_ld, project, empty_norm, empty_nn,
                         (float *)U);
    FM_HIP(hipGetLastError());
    return;
  }
  int nc = d.nr * d.H;
  FM_CHECK(d.nr <= 64, FM_ERR_INVALID, "B3 adjoint: n_species n_max above 64");
  size_t per = (4 * (size_t)nc + (size_t)d.n_tri) * esz;
  int apb = waves_per_block(per + 256);
  int nblk = (n + apb - 1) / apb;
  size_t lds = 256 + apb * per;
  if (f64)
    hipLaunchKernelGGL(k_dEdc_b3<double>, dim3(nblk), dim3(64 * apb), lds, st, n, d,
                       s->C.as<double>(), s->row_of_atom.as<int>(), rr, (const double *)B,
                       s->norm.as<double>(), s->nn.as<double>(), (const double *)G,
                       (long long)g_ld, project, empty_norm, empty_nn, (double *)U);
  else
    hipLaunchKernelGGL(k_dEdc_b3<float>, dim3(nblk), dim3(64 * apb), lds, st, n, d,
                       s->C.as<float>(), s->row_of_atom.as<int>(), rr, (const float *)B,
                       s->norm.as<double>(), s->nn.as<double>(), (const float *)G,
                       (long long)g_ld, project, empty_norm, empty_nn, (float *)U);
  FM_HIP(hipGetLastError());
}

}  // namespace fm
