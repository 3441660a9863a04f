// Full kernel matrices for the DTC and SOR variances of small structures
// (sparse_gp.cpp:667-729): dense descriptor derivatives dB/dr
// (b2.cpp:181-242), the kernel matrix K = envs_struc(sparse, structure)
// (normalized_dot_product.cpp:117-214), the self kernel
// (normalized_dot_product.cpp:450-586), and diag(K^T Kuu^-1 K) and
// diag(K^T Sigma K) as |L_inv K|^2 and |R_inv^T K|^2 on the matrix cores.
//
// These follow the reference's dense formulation on purpose (the DTC
// variance needs the force-force products); memory grows as nb_s^2 per
// species, so they are meant for structures of up to a few thousand atoms,
// as in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

constexpr int kMaxHCd = 2;

__constant__ double c_knorm_d[kMaxH];

// ---------------------------------------------------------------------------
// Dense dB/dr per bond and component (b2.cpp:219-228), written at the
// bond's species-sorted position f (rows 3 f + m: the layout of
// DescriptorValues::descriptor_force_dervs), and the force dots dB . B
// (b2.cpp:236-240). One wavefront per atom.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_b2_dervs(
    int n, int nos, int N, int lmax, int H, int nrad, int nd, int cutkind, double r1,
    const int *species, const int *bond_off, const int *bond_q, const int *nspec,
    const double *rel, const double *cutoffs, const int *row_of_atom, const int *sb_off,
    const int2 *pairs, const double *B, double *DB, double *fdots) {
  extern __shared__ __align__(16) unsigned char smem[];
  int i = blockIdx.x, lane = threadIdx.x;
  if (i >= n) return;
  double *c = reinterpret_cast<double *>(smem);  // nrad x H
  double *dc = c + (size_t)nrad * H;              // N x H x 3 (current bond)
  int si = species[i];
  int b0 = bond_off[i], b1 = bond_off[i + 1];
  int row = row_of_atom[i];
  int f0 = sb_off[row];
  int ls[kMaxHCd], ms[kMaxHCd];
  double kn[kMaxHCd];
  for (int ch = 0; ch < kMaxHCd; ch++) {
    int lm = lane + 64 * ch;
    int lmc = lm < H ? lm : 0;
    int l = 0;
    while ((l + 1) * (l + 1) <= lmc) l++;
    ls[ch] = l;
    ms[ch] = lmc - l * l - l;
    kn[ch] = c_knorm_d[lmc];
  }
  // Single-bond coefficients c (b2.cpp:367-395), every neighbour species.
  for (int q = lane; q < nrad * H; q += 64) c[q] = 0;
  __syncthreads();
  for (int b = b0; b < b1; b++) {
    int qn = bond_q[b];
    int sj = nspec[qn];
    const double *rp = rel + 4 * (size_t)qn;
    double rc = cutoffs[si * nos + sj];
    double g[kMaxN], dg[kMaxN];
    radial_basis<double>(rp[0], rc, r1, N, cutkind, g, dg);
    double rinv = 1.0 / rp[0];
    double ux = rp[1] * rinv, uy = rp[2] * rinv, uz = rp[3] * rinv;
    for (int ch = 0; ch < kMaxHCd; ch++) {
      int lm = lane + 64 * ch;
      if (lm >= H) continue;
      double Y, d0, d1, d2;
      sph_harm_lane<double, false>(ls[ch], ms[ch], kn[ch], ux, uy, uz, rinv, Y, d0, d1, d2);
      for (int nn = 0; nn < N; nn++) c[(sj * N + nn) * H + lm] += g[nn] * Y;
    }
  }
  __syncthreads();
  const double *Brow = B + (size_t)row * nd;
  int L1 = lmax + 1;
  for (int b = b0; b < b1; b++) {
    int qn = bond_q[b];
    int sj = nspec[qn];
    const double *rp = rel + 4 * (size_t)qn;
    double rc = cutoffs[si * nos + sj];
    double g[kMaxN], dg[kMaxN];
    radial_basis<double>(rp[0], rc, r1, N, cutkind, g, dg);
    double r = rp[0], rinv = 1.0 / r;
    double ux = rp[1] * rinv, uy = rp[2] * rinv, uz = rp[3] * rinv;
    for (int ch = 0; ch < kMaxHCd; ch++) {
      int lm = lane + 64 * ch;
      if (lm >= H) continue;
      double Y, Yx, Yy, Yz;
      sph_harm_lane<double, true>(ls[ch], ms[ch], kn[ch], ux, uy, uz, rinv, Y, Yx, Yy, Yz);
      for (int nn = 0; nn < N; nn++) {
        double *o = dc + ((size_t)nn * H + lm) * 3;
        o[0] = dg[nn] * ux * Y + g[nn] * Yx;
        o[1] = dg[nn] * uy * Y + g[nn] * Yy;
        o[2] = dg[nn] * uz * Y + g[nn] * Yz;
      }
    }
    __syncthreads();
    int f = f0 + (b - b0);
    double fd[3] = {0, 0, 0};
    for (int k = lane; k < nd; k += 64) {
      int p = k / L1, l = k - p * L1;
      int2 pr = pairs[p];
      int n1 = pr.x, n2 = pr.y;
      bool in1 = n1 / N == sj, in2 = n2 / N == sj;
      double acc[3] = {0, 0, 0};
      for (int mm = 0; mm < 2 * l + 1; mm++) {
        int lm = l * l + mm;
        double c1v = c[n1 * H + lm], c2v = c[n2 * H + lm];
        for (int comp = 0; comp < 3; comp++) {
          double d1 = in1 ? dc[((size_t)(n1 - sj * N) * H + lm) * 3 + comp] : 0.0;
          double d2 = in2 ? dc[((size_t)(n2 - sj * N) * H + lm) * 3 + comp] : 0.0;
          acc[comp] += c1v * d2 + d1 * c2v;
        }
      }
      for (int comp = 0; comp < 3; comp++) {
        DB[((size_t)3 * f + comp) * nd + k] = acc[comp];
        fd[comp] += acc[comp] * Brow[k];
      }
    }
    for (int comp = 0; comp < 3; comp++) {
      double v = wave_sum(fd[comp]);
      if (lane == 0) fdots[3 * (size_t)f + comp] = v;
    }
    __syncthreads();
  }
}

__global__ void k_bond_owner(int n, const int *bond_off, int *owner) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int b = bond_off[i]; b < bond_off[i + 1]; b++) owner[b] = i;
}

__device__ __forceinline__ double dpow(double x, double p) {
  if (p == 2.0) return x * x;
  if (p == 1.0) return x;
  return pow(x, p);
}

// ---------------------------------------------------------------------------
// K = envs_struc for the sparse environments of one species t, written
// transposed (Kt: n_out x M, column cs + j).
// ---------------------------------------------------------------------------
struct EnvsArgs {
  int n, nd, M, cs, Ms, t, ro, nt, sbo, nbt;
  double sig2, power, vol_inv;
  const int *species, *row_of_atom, *bond_off, *in_off, *in_bond, *sb_pos, *sb_row, *sb_bond,
      *bond_q;
  const double *norm, *xnorm, *rel;
  const double *S;   // Ms x nt: x_hat_j . B_r (raw B)
  const double *FD;  // Ms x 3 nbt: x_hat_j . dB_f,m
  const double *fdots;
  double *Kt;
};

// fk(j, f, m) = sigma^2 p S^(p-1) (x_hat . dB / |B| - S fdot / |B|^2), the
// force kernel of normalized_dot_product.cpp:187-195 with S = x.B/(|x||B|).
__device__ __forceinline__ double fk_of(const EnvsArgs &e, int j, int f, int m) {
  int r = e.sb_row[f];
  double nb = e.norm[r];
  if (nb < kEmptyThresh) return 0.0;
  double s = e.S[(size_t)j * e.nt + (r - e.ro)] / nb;
  double dval = e.power * dpow(s, e.power - 1);
  double f1 = e.FD[(size_t)j * 3 * e.nbt + 3 * (size_t)(f - e.sbo) + m] / nb;
  double f2 = s * e.fdots[3 * (size_t)f + m] / (nb * nb);
  return e.sig2 * dval * (f1 - f2);
}

// Force columns: + bonds owned by a, - bonds pointing at a
// (normalized_dot_product.cpp:197-199), owners of species t only.
__global__ void k_envs_forces(EnvsArgs e) {
  int j = blockIdx.y;
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= e.Ms || a >= e.n) return;
  size_t col = (size_t)e.cs + j;
  double acc[3] = {0, 0, 0};
  if (e.xnorm[j] >= kEmptyThresh) {
    if (e.species[a] == e.t)
      for (int b = e.bond_off[a]; b < e.bond_off[a + 1]; b++)
        for (int m = 0; m < 3; m++) acc[m] += fk_of(e, j, e.sb_pos[b], m);
    for (int q = e.in_off[a]; q < e.in_off[a + 1]; q++) {
      int f = e.sb_pos[e.in_bond[q]];
      if (f < e.sbo || f >= e.sbo + e.nbt) continue;  // owner of another species
      for (int m = 0; m < 3; m++) acc[m] -= fk_of(e, j, f, m);
    }
  }
  for (int m = 0; m < 3; m++) e.Kt[(size_t)(1 + 3 * a + m) * e.M + col] = acc[m];
}

// Energy and stress columns of sparse env j (one block of 256 per j):
// E = sum_r sigma^2 S^p; S_mn = -1/V sum_f fk(m) coord_f(n), n >= m
// (normalized_dot_product.cpp:175-177, 201-206).
__global__ __launch_bounds__(256) void k_envs_energy_stress(EnvsArgs e) {
  __shared__ double sh[7][256];
  int j = blockIdx.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (e.xnorm[j] >= kEmptyThresh) {
    for (int r = threadIdx.x; r < e.nt; r += blockDim.x) {
      double nb = e.norm[e.ro + r];
      if (nb < kEmptyThresh) continue;
      double s = e.S[(size_t)j * e.nt + r] / nb;
      acc[0] += e.sig2 * dpow(s, e.power);
    }
    for (int fl = threadIdx.x; fl < e.nbt; fl += blockDim.x) {
      int f = e.sbo + fl;
      const double *co = e.rel + 4 * (size_t)e.bond_q[e.sb_bond[f]] + 1;
      int sc = 0;
      for (int m = 0; m < 3; m++) {
        double fk = fk_of(e, j, f, m);
        for (int nn = m; nn < 3; nn++) acc[1 + sc++] -= fk * co[nn] * e.vol_inv;
      }
    }
  }
  for (int q = 0; q < 7; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 7; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 7) {
    size_t col = (size_t)e.cs + j;
    int c = threadIdx.x == 0 ? 0 : 1 + 3 * e.n + (threadIdx.x - 1);
    e.Kt[(size_t)c * e.M + col] = sh[threadIdx.x][0];
  }
}

// ---------------------------------------------------------------------------
// Self kernel of one species (normalized_dot_product.cpp:450-586), literally:
// pairs j >= i with factor 2 off the diagonal, bond pairs (f1 of i, f2 of j)
// and components m, scattered by the atom-index coincidences of the
// reference. One wavefront per row i; lanes take bonds f1 of i.
// ---------------------------------------------------------------------------
struct SelfArgs {
  int ro, nt, sbo, nbt, n;
  double sig2, power, vol_inv_sq;
  const int *atom_of_row, *sb_off, *nbr_of_f, *sb_bond, *bond_q;
  const double *norm, *rel;
  const double *BB;   // nt x nt raw dots
  const double *FD2;  // 3 nbt x nt: dB_f,m . B_r
  const double *FF;   // 3 x nbt x nbt: dB_f1,m . dB_f2,m
  const double *fdots;
  double *Pn;  // 3 nb: per bond (species-sorted position)
  double *Pc;  // 3 n rows
  double *Ps;  // 6 n rows
  double *Pe;  // n rows
  int *bad;
};

__global__ __launch_bounds__(64) void k_self_pairs(SelfArgs a) {
  int il = blockIdx.x, lane = threadIdx.x;
  if (il >= a.nt) return;
  int ri = a.ro + il;
  double norm_i = a.norm[ri];
  double ec = 0, cc[3] = {0, 0, 0}, sc6[6] = {0, 0, 0, 0, 0, 0};
  int fb0 = a.sb_off[ri], fb1 = a.sb_off[ri + 1];
  if (norm_i >= kEmptyThresh) {
    double norm_i2 = norm_i * norm_i, norm_i3 = norm_i2 * norm_i;
    int ci1 = a.atom_of_row[ri];
    for (int fc = fb0; fc < fb1 || fc == fb0; fc += 64) {
      int f1 = fc + lane;
      bool own = f1 < fb1;
      int ni1 = own ? a.nbr_of_f[f1] : -1;
      const double *co1 = own ? a.rel + 4 * (size_t)a.bond_q[a.sb_bond[f1]] + 1 : nullptr;
      double nacc[3] = {0, 0, 0};
      for (int jl = il; jl < a.nt; jl++) {
        int rj = a.ro + jl;
        double norm_j = a.norm[rj];
        if (norm_j < kEmptyThresh) continue;
        double mult = (jl == il) ? 1 : 2;
        double norm_j2 = norm_j * norm_j, norm_j3 = norm_j2 * norm_j;
        double norm_ij = norm_i * norm_j;
        double norm_dot = a.BB[(size_t)il * a.nt + jl] / norm_ij;
        double c1 = (a.power - 1) * a.power * dpow(norm_dot, a.power - 2);
        if (fabs(norm_dot) < kEmptyThresh && a.power < 2) a.bad[0] = 1;
        double c2 = a.power * dpow(norm_dot, a.power - 1);
        if (fc == fb0 && lane == 0) ec += a.sig2 * mult * dpow(norm_dot, a.power);
        if (!own) continue;
        int ci2 = a.atom_of_row[rj];
        int fj0 = a.sb_off[rj], fj1 = a.sb_off[rj + 1];
        int fl1 = f1 - a.sbo;
        for (int f2 = fj0; f2 < fj1; f2++) {
          int ni2 = a.nbr_of_f[f2];
          int fl2 = f2 - a.sbo;
          const double *co2 = a.rel + 4 * (size_t)a.bond_q[a.sb_bond[f2]] + 1;
          int s_c = 0;
          for (int m = 0; m < 3; m++) {
            double fd1 = a.fdots[3 * (size_t)f1 + m], fd2 = a.fdots[3 * (size_t)f2 + m];
            double fdj = a.FD2[(3 * (size_t)fl1 + m) * a.nt + jl];
            double fdi = a.FD2[(3 * (size_t)fl2 + m) * a.nt + il];
            double v1 = fdj / norm_ij - norm_dot * fd1 / norm_i2;
            double v2 = fdi / norm_ij - norm_dot * fd2 / norm_j2;
            double v3 = a.FF[((size_t)m * a.nbt + fl1) * a.nbt + fl2] / norm_ij;
            double v4 = fd1 * fdi / (norm_i3 * norm_j);
            double v5 = fd2 * fdj / (norm_i * norm_j3);
            double v6 = fd1 * fd2 * norm_dot / (norm_i2 * norm_j2);
            double kv = a.sig2 * mult * (c1 * v1 * v2 + c2 * (v3 - v4 - v5 + v6));
            if (ci1 == ci2) cc[m] += kv;
            if (ci1 == ni2) cc[m] -= kv;
            if (ni1 == ci2) nacc[m] -= kv;
            if (ni1 == ni2) nacc[m] += kv;
            for (int nn = m; nn < 3; nn++) {
              sc6[s_c] += kv * co1[nn] * co2[nn] * a.vol_inv_sq;
              s_c++;
            }
          }
        }
      }
      if (own)
        for (int m = 0; m < 3; m++) a.Pn[3 * (size_t)f1 + m] = nacc[m];
      if (fb1 == fb0) break;
    }
  } else {
    for (int f1 = fb0 + lane; f1 < fb1; f1 += 64)
      for (int m = 0; m < 3; m++) a.Pn[3 * (size_t)f1 + m] = 0;
  }
  for (int m = 0; m < 3; m++) cc[m] = wave_sum(cc[m]);
  for (int q = 0; q < 6; q++) sc6[q] = wave_sum(sc6[q]);
  ec = wave_sum(ec);
  if (lane == 0) {
    for (int m = 0; m < 3; m++) a.Pc[3 * (size_t)ri + m] = cc[m];
    for (int q = 0; q < 6; q++) a.Ps[6 * (size_t)ri + q] = sc6[q];
    a.Pe[ri] = ec;
  }
}

// K_self columns: energy and stress as sums over rows, forces as the own
// row's partial plus the partials of bonds pointing at the atom.
__global__ void k_self_forces(int n, const int *row_of_atom, const int *in_off,
                              const int *in_bond, const int *sb_pos, const double *Pc,
                              const double *Pn, double *kself) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  int r = row_of_atom[a];
  double acc[3] = {Pc[3 * (size_t)r], Pc[3 * (size_t)r + 1], Pc[3 * (size_t)r + 2]};
  for (int q = in_off[a]; q < in_off[a + 1]; q++) {
    int f = sb_pos[in_bond[q]];
    for (int m = 0; m < 3; m++) acc[m] += Pn[3 * (size_t)f + m];
  }
  for (int m = 0; m < 3; m++) kself[1 + 3 * a + m] = acc[m];
}

__global__ __launch_bounds__(256) void k_self_sums(int n, const double *Pe, const double *Ps,
                                                   double *kself) {
  __shared__ double sh[7][256];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    acc[0] += Pe[r];
    for (int q = 0; q < 6; q++) acc[1 + q] += Ps[6 * (size_t)r + q];
  }
  for (int q = 0; q < 7; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 7; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) kself[0] = sh[0][0];
  if (threadIdx.x >= 1 && threadIdx.x < 7) kself[1 + 3 * n + threadIdx.x - 1] = sh[threadIdx.x][0];
}

// var = kself - sum Q partials + sum V partials (DTC) or the V partials (SOR).
__global__ void k_var_efs(int n_out, int dtc, const double *kself, const double *Qp, int nq,
                          const double *Vp, int nv, double *var) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_out) return;
  double q = 0, v = 0;
  for (int t = 0; t < nq; t++) q += Qp[(size_t)t * n_out + c];
  for (int t = 0; t < nv; t++) v += Vp[(size_t)t * n_out + c];
  var[c] = dtc ? kself[c] - q + v : v;
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
static void upload_knorm_d(hipStream_t st, int) {
  static thread_local double host[kMaxH];
  knorm_values(kMaxL, host);
  FM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_knorm_d), host, sizeof(host), 0,
                                hipMemcpyHostToDevice, st));
}

void compute_dervs(fm_structure s) {
  if (s->has_dervs) return;
  FM_CHECK(s->has_b2 && s->form == FM_FORM_FP64, FM_ERR_STATE,
           "dense descriptor derivatives need FP64 descriptors");
  FM_CHECK(s->desc_kind == 2, FM_ERR_INVALID,
           "dense descriptor derivatives are built for the B2 descriptor");
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  int n = s->n, nd = s->nd, H = (s->lmax + 1) * (s->lmax + 1), nrad = s->nos * s->N;
  size_t nb = (size_t)s->n_bonds;
  FM_CHECK((double)nb * 3 * nd * 8 < 64e9, FM_ERR_CAPACITY,
           "dense descriptor derivatives exceed 64 GB; the adjoint path needs none");
  // Species-sorted bond layout: rows in species order, bonds of a row
  // contiguous (b2.cpp:150-154).
  std::vector<int> bo(n + 1), aor(n), sp(n);
  FM_HIP(hipMemcpyAsync(bo.data(), s->bond_off.as<int>(), (n + 1) * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  FM_HIP(hipMemcpyAsync(aor.data(), s->atom_of_row.as<int>(), n * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  FM_HIP(hipMemcpyAsync(sp.data(), s->species.as<int>(), n * sizeof(int), hipMemcpyDeviceToHost,
                        st));
  FM_HIP(hipStreamSynchronize(st));
  std::vector<int> sb_off(n + 1, 0), sb_pos(nb), sb_row(nb), sb_bond(nb);
  s->nbt_off.assign(s->nos + 1, 0);
  for (int r = 0; r < n; r++) {
    int i = aor[r];
    int cnt = bo[i + 1] - bo[i];
    sb_off[r + 1] = sb_off[r] + cnt;
    for (int k = 0; k < cnt; k++) {
      int f = sb_off[r] + k, b = bo[i] + k;
      sb_pos[b] = f;
      sb_row[f] = r;
      sb_bond[f] = b;
    }
    s->nbt_off[sp[i] + 1] += cnt;
  }
  for (int t = 0; t < s->nos; t++) s->nbt_off[t + 1] += s->nbt_off[t];
  auto up = [&](DevBuf &buf, const void *src, size_t bytes) {
    buf.alloc(pool, std::max<size_t>(1, bytes));
    if (bytes) FM_HIP(hipMemcpyAsync(buf.as<void>(), src, bytes, hipMemcpyHostToDevice, st));
  };
  up(s->sb_off, sb_off.data(), sb_off.size() * sizeof(int));
  up(s->sb_pos, sb_pos.data(), sb_pos.size() * sizeof(int));
  up(s->sb_row, sb_row.data(), sb_row.size() * sizeof(int));
  up(s->sb_bond, sb_bond.data(), sb_bond.size() * sizeof(int));
  s->DB.alloc(pool, std::max<size_t>(1, nb * 3 * nd * sizeof(double)));
  s->fdots.alloc(pool, std::max<size_t>(1, nb * 3 * sizeof(double)));
  upload_knorm_d(st, s->lmax);
  size_t lds = ((size_t)nrad * H + (size_t)s->N * H * 3) * sizeof(double);
  FM_CHECK(lds <= 160 * 1024, FM_ERR_CAPACITY, "derivative workspace exceeds LDS");
  if (n > 0)
    hipLaunchKernelGGL(k_b2_dervs, dim3(n), dim3(64), lds, st, n, s->nos, s->N, s->lmax, H, nrad,
                       nd, s->cutkind, s->r1, s->species.as<int>(), s->bond_off.as<int>(),
                       s->bond_q.as<int>(), s->nspecies.as<int>(), s->rel.as<double>(),
                       s->d_cutoffs.as<double>(), s->row_of_atom.as<int>(), s->sb_off.as<int>(),
                       s->pair_table.as<int2>(), s->B.as<double>(), s->DB.as<double>(),
                       s->fdots.as<double>());
  FM_HIP(hipGetLastError());
  FM_HIP(hipStreamSynchronize(st));
  s->has_dervs = true;
}

// Kt (n_out x M) = envs_struc(sparse, structure)^T.
static void envs_struc_t(fm_model m, fm_structure s, DevBuf &Kt, DevBuf &owner) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  int n = s->n, nd = s->nd, M = m->M;
  size_t n_out = 1 + 3 * (size_t)n + 6;
  Kt.alloc(pool, std::max<size_t>(1, n_out * M * sizeof(double)));
  FM_HIP(hipMemsetAsync(Kt.as<double>(), 0, n_out * M * sizeof(double), st));
  owner.alloc(pool, std::max<size_t>(1, (size_t)s->n_bonds * sizeof(int)));
  if (n > 0)
    hipLaunchKernelGGL(k_bond_owner, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       s->bond_off.as<int>(), owner.as<int>());
  for (int t = 0; t < s->nos; t++) {
    int Ms = m->M_by_type[t], nt = s->n_by_type[t], ro = s->type_off[t];
    int sbo = s->nbt_off[t], nbt = s->nbt_off[t + 1] - s->nbt_off[t];
    if (Ms == 0) continue;
    FM_CHECK(Ms < 65536, FM_ERR_CAPACITY, "more than 65535 sparse environments of one species");
    DevBuf S(pool, std::max<size_t>(1, (size_t)Ms * std::max(nt, 1) * sizeof(double)));
    DevBuf FD(pool, std::max<size_t>(1, (size_t)Ms * 3 * std::max(nbt, 1) * sizeof(double)));
    DevBuf xn(pool, (size_t)Ms * sizeof(double));
    FM_HIP(hipMemcpyAsync(xn.as<double>(), m->xnorm[t].data(), Ms * sizeof(double),
                          hipMemcpyHostToDevice, st));
    if (nt > 0) {
      EpArgs e1;
      e1.W = S.as<void>();
      e1.ldw = nt;
      gemm_f64(st, EP_STORE, Ms, nt, nd, m->X[t].as<double>(), nd,
               s->B.as<double>() + (size_t)ro * nd, nd, 0, e1);
    }
    if (nbt > 0) {
      EpArgs e2;
      e2.W = FD.as<void>();
      e2.ldw = 3 * (long long)nbt;
      gemm_f64(st, EP_STORE, Ms, 3 * nbt, nd, m->X[t].as<double>(), nd,
               s->DB.as<double>() + (size_t)3 * sbo * nd, nd, 0, e2);
    }
    EnvsArgs e;
    e.n = n;
    e.nd = nd;
    e.M = M;
    e.cs = m->off[t];
    e.Ms = Ms;
    e.t = t;
    e.ro = ro;
    e.nt = nt;
    e.sbo = sbo;
    e.nbt = nbt;
    e.sig2 = m->sigma * m->sigma;
    e.power = m->power;
    e.vol_inv = 1.0 / s->volume;
    e.species = s->species.as<int>();
    e.row_of_atom = s->row_of_atom.as<int>();
    e.bond_off = s->bond_off.as<int>();
    e.in_off = s->in_off.as<int>();
    e.in_bond = s->in_bond.as<int>();
    e.sb_pos = s->sb_pos.as<int>();
    e.sb_row = s->sb_row.as<int>();
    e.sb_bond = s->sb_bond.as<int>();
    e.bond_q = s->bond_q.as<int>();
    e.norm = s->norm.as<double>();
    e.xnorm = xn.as<double>();
    e.rel = s->rel.as<double>();
    e.S = S.as<double>();
    e.FD = FD.as<double>();
    e.fdots = s->fdots.as<double>();
    e.Kt = Kt.as<double>();
    if (n > 0)
      hipLaunchKernelGGL(k_envs_forces, dim3((n + 127) / 128, Ms), dim3(128), 0, st, e);
    hipLaunchKernelGGL(k_envs_energy_stress, dim3(Ms), dim3(256), 0, st, e);
    FM_HIP(hipGetLastError());
    FM_HIP(hipStreamSynchronize(st));
  }
}

static void self_kernel(fm_model m, fm_structure s, DevBuf &kself) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  int n = s->n, nd = s->nd;
  size_t n_out = 1 + 3 * (size_t)n + 6;
  kself.alloc(pool, n_out * sizeof(double));
  FM_HIP(hipMemsetAsync(kself.as<double>(), 0, n_out * sizeof(double), st));
  if (n == 0) return;
  DevBuf Pn(pool, std::max<size_t>(1, (size_t)3 * s->n_bonds * sizeof(double)));
  DevBuf Pc(pool, (size_t)3 * n * sizeof(double)), Ps(pool, (size_t)6 * n * sizeof(double)),
      Pe(pool, (size_t)n * sizeof(double)), bad(pool, sizeof(int));
  FM_HIP(hipMemsetAsync(bad.as<int>(), 0, sizeof(int), st));
  // Neighbour (structure index) of every species-sorted bond.
  std::vector<int> sbb(s->n_bonds), bq(s->n_bonds), sidx(s->n_nb);
  if (s->n_bonds) {
    FM_HIP(hipMemcpyAsync(sbb.data(), s->sb_bond.as<int>(), sbb.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    FM_HIP(hipMemcpyAsync(bq.data(), s->bond_q.as<int>(), bq.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
  }
  if (s->n_nb)
    FM_HIP(hipMemcpyAsync(sidx.data(), s->sidx.as<int>(), sidx.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  std::vector<int> nbr(s->n_bonds);
  for (size_t f = 0; f < nbr.size(); f++) nbr[f] = sidx[bq[sbb[f]]];
  DevBuf d_nbr(pool, std::max<size_t>(1, nbr.size() * sizeof(int)));
  if (!nbr.empty())
    FM_HIP(hipMemcpyAsync(d_nbr.as<int>(), nbr.data(), nbr.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
  for (int t = 0; t < s->nos; t++) {
    int nt = s->n_by_type[t], ro = s->type_off[t];
    int sbo = s->nbt_off[t], nbt = s->nbt_off[t + 1] - s->nbt_off[t];
    if (nt == 0) continue;
    FM_CHECK((double)nbt * nbt * 3 * 8 < 64e9, FM_ERR_CAPACITY,
             "the self kernel's force-force products exceed 64 GB");
    DevBuf BB(pool, (size_t)nt * nt * sizeof(double));
    DevBuf FD2(pool, std::max<size_t>(1, (size_t)3 * nbt * nt * sizeof(double)));
    DevBuf FF(pool, std::max<size_t>(1, (size_t)3 * nbt * nbt * sizeof(double)));
    const double *Bs = s->B.as<double>() + (size_t)ro * nd;
    const double *DBs = s->DB.as<double>() + (size_t)3 * sbo * nd;
    EpArgs e;
    e.W = BB.as<void>();
    e.ldw = nt;
    gemm_f64(st, EP_STORE, nt, nt, nd, Bs, nd, Bs, nd, 0, e);
    if (nbt > 0) {
      e.W = FD2.as<void>();
      e.ldw = nt;
      gemm_f64(st, EP_STORE, 3 * nbt, nt, nd, DBs, nd, Bs, nd, 0, e);
      for (int mm = 0; mm < 3; mm++) {
        e.W = FF.as<double>() + (size_t)mm * nbt * nbt;
        e.ldw = nbt;
        gemm_f64(st, EP_STORE, nbt, nbt, nd, DBs + (size_t)mm * nd, 3 * (long long)nd,
                 DBs + (size_t)mm * nd, 3 * (long long)nd, 0, e);
      }
    }
    SelfArgs a;
    a.ro = ro;
    a.nt = nt;
    a.sbo = sbo;
    a.nbt = nbt;
    a.n = n;
    a.sig2 = m->sigma * m->sigma;
    a.power = m->power;
    a.vol_inv_sq = 1.0 / (s->volume * s->volume);
    a.atom_of_row = s->atom_of_row.as<int>();
    a.sb_off = s->sb_off.as<int>();
    a.nbr_of_f = d_nbr.as<int>();
    a.sb_bond = s->sb_bond.as<int>();
    a.bond_q = s->bond_q.as<int>();
    a.norm = s->norm.as<double>();
    a.rel = s->rel.as<double>();
    a.BB = BB.as<double>();
    a.FD2 = FD2.as<double>();
    a.FF = FF.as<double>();
    a.fdots = s->fdots.as<double>();
    a.Pn = Pn.as<double>();
    a.Pc = Pc.as<double>();
    a.Ps = Ps.as<double>();
    a.Pe = Pe.as<double>();
    a.bad = bad.as<int>();
    hipLaunchKernelGGL(k_self_pairs, dim3(nt), dim3(64), 0, st, a);
    FM_HIP(hipGetLastError());
    FM_HIP(hipStreamSynchronize(st));
  }
  int hb = 0;
  FM_HIP(hipMemcpyAsync(&hb, bad.as<int>(), sizeof(int), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  FM_CHECK(!hb, FM_ERR_INVALID,
           "Dot product of descriptors is 0, and the negative power function in force-force "
           "kernel diverges.");
  hipLaunchKernelGGL(k_self_forces, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     s->row_of_atom.as<int>(), s->in_off.as<int>(), s->in_bond.as<int>(),
                     s->sb_pos.as<int>(), Pc.as<double>(), Pn.as<double>(), kself.as<double>());
  hipLaunchKernelGGL(k_self_sums, dim3(1), dim3(256), 0, st, n, Pe.as<double>(), Ps.as<double>(),
                     kself.as<double>());
  FM_HIP(hipGetLastError());
  FM_HIP(hipStreamSynchronize(st));
}

void envs_struc(fm_model m, fm_structure s, double *K_out) {
  FM_CHECK(s->n_struct == 1, FM_ERR_INVALID, "kernel matrices are per single structure");
  compute_dervs(s);
  DevBuf Kt, owner;
  envs_struc_t(m, s, Kt, owner);
  size_t n_out = 1 + 3 * (size_t)s->n + 6;
  std::vector<double> kt(n_out * m->M);
  FM_HIP(hipMemcpyAsync(kt.data(), Kt.as<double>(), kt.size() * sizeof(double),
                        hipMemcpyDeviceToHost, s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  for (int j = 0; j < m->M; j++)
    for (size_t c = 0; c < n_out; c++) K_out[(size_t)j * n_out + c] = kt[c * m->M + j];
}

void self_kernel_struc(fm_model m, fm_structure s, double *kself_out) {
  FM_CHECK(s->n_struct == 1, FM_ERR_INVALID, "kernel matrices are per single structure");
  compute_dervs(s);
  DevBuf ks;
  self_kernel(m, s, ks);
  size_t n_out = 1 + 3 * (size_t)s->n + 6;
  FM_HIP(hipMemcpyAsync(kself_out, ks.as<double>(), n_out * sizeof(double), hipMemcpyDeviceToHost,
                        s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
}

void predict_dense(fm_model m, fm_structure s, int kind, double *mean_efs, double *var_efs) {
  FM_CHECK(s->n_struct == 1, FM_ERR_INVALID, "DTC/SOR variances are per single structure");
  FM_CHECK(m->has_R, FM_ERR_STATE, "R_inv is needed for the DTC and SOR variances");
  FM_CHECK(kind != 1 || m->has_L, FM_ERR_STATE,
           "L_inv must be up to date to make predictions. Please call update_matrices_QR and try "
           "again.");
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  compute_dervs(s);
  int M = m->M;
  size_t n_out = 1 + 3 * (size_t)s->n + 6;
  DevBuf Kt, owner, ks;
  envs_struc_t(m, s, Kt, owner);
  // mean = K^T alpha
  DevBuf mean(pool, n_out * sizeof(double));
  {
    EpArgs e;
    e.W = mean.as<void>();
    e.ldw = 1;
    gemm_f64(st, EP_STORE, (int)n_out, 1, M, Kt.as<double>(), M, m->alpha_full.as<double>(), M, 0,
             e);
  }
  int dtc = kind == 1;
  if (dtc) self_kernel(m, s, ks);
  // |L_inv K_c|^2 per species block and |R_inv^T K_c|^2.
  int nq = 0;
  for (int t = 0; t < m->nos; t++) nq += dtc ? gemm_row_tiles(FM_FORM_FP64, m->L_rows[t]) : 0;
  int nv = gemm_row_tiles(FM_FORM_FP64, M);
  DevBuf Qp(pool, std::max<size_t>(1, (size_t)nq * n_out * sizeof(double)));
  DevBuf Vp(pool, std::max<size_t>(1, (size_t)nv * n_out * sizeof(double)));
  int qo = 0;
  if (dtc)
    for (int t = 0; t < m->nos; t++) {
      if (m->M_by_type[t] == 0 || m->L_rows[t] == 0) continue;
      EpArgs eq;
      eq.part = Qp.as<double>() + (size_t)qo * n_out;
      eq.part_ld = n_out;
      gemm_f64(st, EP_SUMSQ, m->L_rows[t], (int)n_out, m->M_by_type[t], m->L[t].as<double>(),
               m->M_by_type[t], Kt.as<double>() + m->off[t], M, m->L_lower[t], eq);
      qo += gemm_row_tiles(FM_FORM_FP64, m->L_rows[t]);
    }
  {
    EpArgs ev;
    ev.part = Vp.as<double>();
    ev.part_ld = n_out;
    gemm_f64(st, EP_SUMSQ, M, (int)n_out, M, m->RinvT.as<double>(), M, Kt.as<double>(), M,
             m->R_lower, ev);
  }
  DevBuf var(pool, n_out * sizeof(double));
  hipLaunchKernelGGL(k_var_efs, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, st,
                     (int)n_out, dtc, dtc ? ks.as<double>() : nullptr, Qp.as<double>(), qo,
                     Vp.as<double>(), nv, var.as<double>());
  FM_HIP(hipGetLastError());
  if (mean_efs)
    FM_HIP(hipMemcpyAsync(mean_efs, mean.as<double>(), n_out * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  if (var_efs)
    FM_HIP(hipMemcpyAsync(var_efs, var.as<double>(), n_out * sizeof(double), hipMemcpyDeviceToHost,
                          st));
  FM_HIP(hipStreamSynchronize(st));
}

}  // namespace fm
