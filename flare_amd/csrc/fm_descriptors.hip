// Fused bond-basis kernels for the B2 descriptor (descriptors/b2.cpp:181-399)
// and its adjoint (the arithmetic of lammps_descriptor.cpp:250-315 and
// pair_flare.cpp:141-174, driven by dE/dB from the kernel contraction).
//
// Descriptor (k_b2): an atom's bonds are taken in chunks of 32; lanes write
// the radial values g_n and all harmonics Y_lm of the chunk into LDS tables
// grouped by neighbour species, and the density coefficients accumulate as
// c_t += G_t^T Y_t on the matrix cores (16x16x4, FP32 or FP64) over that
// species' rows only. The power spectrum B_l = C_l C_l^T is a second MFMA
// product, stored straight to the atom's species-sorted row with |B| and
// B.B; c goes to HBM for the adjoint. The mixed variant computes all of
// this in FP64 and writes the FP32 rows and c the adjoint reads plus the
// int8 slices of the FP64 row that the exact kernel contraction consumes.
// Adjoint: u = dE/dc per atom (w = dE/dB and c staged in LDS, u_l = W_l C_l
// on the matrix cores), then the bond forces with lanes over bonds, each
// lane sweeping all (l, m) of its bond by recursion. Dense dc/dr and dB/dr
// are never formed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

constexpr int kAtomsPerBlock = 4;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct B2Args {
  int n, nos, N, lmax, H, nrad, nd, cutkind;
  double r1;
  const int *species, *bond_off, *bond_q, *nspec, *row_of_atom;
  const double *rel, *cutoffs;
  const int2 *pairs;  // n1, n2 of each radial pair in B2 order
};

__constant__ double c_knorm[kMaxH];

// 1 / k for the associated-Legendre recursions (k <= lmax + 1).
__constant__ double c_rcp[kMaxL + 2];

// Every real Y_lm (l <= L) of one unit vector u by an (m, l) sweep:
// Y_lm = K_lm Q_l^|m|(u_z) {Re, Im}(u_x + i u_y)^|m| (no Condon-Shortley
// phase; l = 1 gives y, z, x as y_grad.cpp:718-736) with Q_l^m = d^m P_l/dt^m
// from the upward recursion in l. Calls f(lm, Y, px, py, pz); with kGrad,
// (px, py, pz) is the gradient in u of the degree-0 extension
// (dQ_l^m/du_z = Q_l^(m+1)), so grad Y_lm(r) = (p - (u . p) u) / r.
// With LC >= 0 the degree is the compile-time LC (L is ignored) and the
// sweep is fully unrolled: every index and constant-table offset is known
// at compile time.
template <typename T, bool kGrad>
struct YlmState {
  T cm = 1, sm = 0, cm1 = 0, sm1 = 0;  // (u_x + i u_y)^mm and ^(mm - 1)
  T dfact = 1;                         // (2 mm - 1)!!
  T q1 = 0, q2 = 0, p1 = 0, p2 = 0;
};

template <typename T, bool kGrad>
__device__ __forceinline__ void ylm_m_begin(YlmState<T, kGrad> &y, int mm, T ux, T uy) {
  if (mm > 0) {
    y.cm1 = y.cm;
    y.sm1 = y.sm;
    y.cm = y.cm1 * ux - y.sm1 * uy;
    y.sm = y.sm1 * ux + y.cm1 * uy;
    y.dfact *= (T)(2 * mm - 1);
  }
  y.q1 = y.dfact;
  y.q2 = 0;
  y.p1 = 0;
  y.p2 = 0;
}

template <typename T, bool kGrad, typename F>
__device__ __forceinline__ void ylm_step(YlmState<T, kGrad> &y, int mm, int l, T uz, F &f) {
  T ql, pl = 0;
  if (l == mm) {
    ql = y.dfact;
  } else {
    ql = ((T)(2 * l - 1) * uz * y.q1 - (T)(l + mm - 1) * y.q2) * (T)c_rcp[l - mm];
    y.q2 = y.q1;
    y.q1 = ql;
    if (kGrad) {
      if (l == mm + 1)
        pl = y.dfact * (T)(2 * mm + 1);
      else
        pl = ((T)(2 * l - 1) * uz * y.p1 - (T)(l + mm) * y.p2) * (T)c_rcp[l - mm - 1];
      y.p2 = y.p1;
      y.p1 = pl;
    }
  }
  int base = l * l + l;
#pragma unroll
  for (int sgn = 0; sgn < 2; sgn++) {
    if (sgn == 1 && mm == 0) break;
    int lm = sgn ? base - mm : base + mm;
    T K = (T)c_knorm[lm];
    T trig, dtx = 0, dty = 0;
    if (mm == 0) {
      trig = 1;
    } else if (sgn == 0) {
      trig = y.cm;
      dtx = (T)mm * y.cm1;
      dty = -(T)mm * y.sm1;
    } else {
      trig = y.sm;
      dtx = (T)mm * y.sm1;
      dty = (T)mm * y.cm1;
    }
    T Kq = K * ql;
    if (kGrad)
      f(lm, Kq * trig, Kq * dtx, Kq * dty, K * pl * trig);
    else
      f(lm, Kq * trig, T(0), T(0), T(0));
  }
}

template <typename T, bool kGrad, int LC = -1, typename F>
__device__ __forceinline__ void ylm_sweep(int L, T ux, T uy, T uz, F &&f) {
  YlmState<T, kGrad> y;
  if constexpr (LC >= 0) {
#pragma unroll
    for (int mm = 0; mm <= LC; mm++) {
      ylm_m_begin<T, kGrad>(y, mm, ux, uy);
#pragma unroll
      for (int l = mm; l <= LC; l++) ylm_step<T, kGrad>(y, mm, l, uz, f);
    }
  } else {
    for (int mm = 0; mm <= L; mm++) {
      ylm_m_begin<T, kGrad>(y, mm, ux, uy);
      for (int l = mm; l <= L; l++) ylm_step<T, kGrad>(y, mm, l, uz, f);
    }
  }
}


// 16 x 16 x 4 MFMA in the descriptor precision (FP32 accumulators hold
// rows 4 (lane / 16) + q, FP64 ones rows lane / 16 + 4 q; both take
// A[lane % 16][lane / 16] and B[lane / 16][lane % 16]).
template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct MfmaAcc<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};
__device__ __forceinline__ MfmaAcc<float>::type mfma_4(float a, float b,
                                                       MfmaAcc<float>::type c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ MfmaAcc<double>::type mfma_4(double a, double b,
                                                        MfmaAcc<double>::type c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <typename T>
__device__ __forceinline__ int mfma_row(int lane, int q) {
  return sizeof(T) == 8 ? (lane >> 4) + 4 * q : 4 * (lane >> 4) + q;
}

// Bonds per table chunk of the descriptor kernel, CH (a template parameter):
// lanes [0, CH) write the tables, so a 32-bond chunk leaves half the wave
// idle in the harmonic sweep and a 64-bond chunk doubles the tables' LDS
// (fewer atoms resident per CU). launch_b2 picks 64 where that costs at
// most a third more LDS per atom, else 32.

// Workspace (elements of T) of one atom of the descriptor kernel: c, then the
// bond tables of one chunk (Y ch x H, g ch x N).
__host__ __device__ inline size_t b2_work(int nc, int H, int N, int ch) {
  return (size_t)nc + (size_t)ch * ((size_t)H + (size_t)N);
}

// c[(t N + n) H + lm] of atom i for all neighbour species t into LDS
// (b2.cpp:367-395). For each chunk of up to CH bonds, lane k takes
// bond k and writes its radial basis g_n and every Y_lm to LDS tables, the
// rows grouped by neighbour species (list order kept within a species); then,
// per species t and 16 harmonics, c_t += G_t^T Y_t on the matrix cores over
// that species' rows only (A = g_n of the bonds, B = their Y_lm).
template <typename T, int LC, int CH>
__device__ void accumulate_c(const B2Args &a, int i, int lane, T *c, T *tab) {
  int si = a.species[i];
  int b0 = a.bond_off[i], b1 = a.bond_off[i + 1];
  int nc = a.nrad * a.H, H = a.H, N = a.N;
  T *Yt = tab, *gt = tab + CH * (size_t)H;
  for (int k = lane; k < nc; k += 64) c[k] = 0;
  int r16 = lane & 15, k4 = lane >> 4;
  unsigned long long below = (1ull << lane) - 1ull;
  for (int bb = b0; bb < b1; bb += CH) {
    int nk = min(CH, b1 - bb);
    int q = 0, sj = -1;
    if (lane < nk) {
      q = a.bond_q[bb + lane];
      sj = a.nspec[q];
    }
    // Table row of this lane's bond: its rank among the chunk's bonds of
    // its species, after the rows of the lower species.
    int pos = 0, off = 0;
    for (int t = 0; t < a.nos; t++) {
      unsigned long long m = __ballot(sj == t);
      if (sj == t) pos = off + __popcll(m & below);
      off += __popcll(m);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < nk) {
      const double *rp = a.rel + 4 * (size_t)q;
      T g[kMaxN], dg[kMaxN];
      radial_basis<T>((T)rp[0], (T)a.cutoffs[si * a.nos + sj], (T)a.r1, N, a.cutkind, g, dg);
#pragma unroll
      for (int nn = 0; nn < kMaxN; nn++)
        if (nn < N) gt[pos * N + nn] = g[nn];
      T rinv = T(1) / (T)rp[0];
      T *yrow = Yt + (size_t)pos * H;
      ylm_sweep<T, false, LC>(a.lmax, (T)rp[1] * rinv, (T)rp[2] * rinv, (T)rp[3] * rinv,
                          [&](int lm, T Y, T, T, T) { yrow[lm] = Y; });
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int r0 = 0;
    for (int t = 0; t < a.nos; t++) {
      int r1 = r0 + __popcll(__ballot(sj == t));  // rows [r0, r1) are species t
      if (r1 > r0) {
        for (int ct = 0; ct < H; ct += 16) {
          int lm = ct + r16;
          typename MfmaAcc<T>::type acc;
#pragma unroll
          for (int qq = 0; qq < 4; qq++) {
            int n = mfma_row<T>(lane, qq);
            acc[qq] = (n < N && lm < H) ? c[(t * N + n) * H + lm] : T(0);
          }
          // Four k-steps per iteration: the loads (clamped in bounds, then
          // zeroed outside the species' rows) all issue before the MFMAs.
          int rc = r16 < N ? r16 : N - 1, lc = lm < H ? lm : H - 1;
          for (int k0 = r0; k0 < r1; k0 += 16) {
            T av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              int b = k0 + 4 * u + k4;
              int bc = b < CH ? b : CH - 1;
              T x = gt[bc * N + rc], y = Yt[bc * H + lc];
              av[u] = (b < r1 && r16 < N) ? x : T(0);
              bv[u] = (b < r1 && lm < H) ? y : T(0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
              if (k0 + 4 * u < r1) acc = mfma_4(av[u], bv[u], acc);
          }
#pragma unroll
          for (int qq = 0; qq < 4; qq++) {
            int n = mfma_row<T>(lane, qq);
            if (n < N && lm < H) c[(t * N + n) * H + lm] = acc[qq];
          }
        }
      }
      r0 = r1;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// B2 (b2.cpp:209-231) per atom: B[k] = sum_m c[n1, lm] c[n2, lm], written at
// the atom's species-sorted row, with |B| and B.B. The FP32 form also writes
// each row as two BF16 pieces (b0 = bf16(B), b1 = bf16(B - b0)) with leading
// dimension ldp and zero padding, the operand of the kernel contraction. The
// coefficients c also go to HBM (Cg) for the adjoint.
struct SplitOut {
  __bf16 *p[2];
};

// The chosen form's outputs of the FP64 descriptor pass (k_b2 with MIXED
// set; 1 stages the FP64 row in LDS and slices it after B_l, 2 slices each
// value as it is formed and stages only the slices):
// the B row and c in FP32 (for the adjoint) and the FP64 row as kI8Slices
// int8 slices with a power-of-two scale (the kernel contraction's operand,
// see k_slice_rows in fm_gemm.hip for the slicing).
struct MixedOut {
  float *B, *C;
  int8_t *q;  // tiled slice layout, kI8TileA rows per tile
  double *scale;
  long long nk;        // k steps of kI8Pad (the padded reduction length / kI8Pad)
  const int *shift;    // per species: slice row = species-sorted row + shift
};

template <typename T, int LC, int CH, int MIXED = 0>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_b2(B2Args a, T *B, double *norm,
                                                              double *nn, SplitOut Bs, int ldp,
                                                              T *Cg, int want_b2, int work,
                                                              MixedOut mo) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * (blockDim.x >> 6) + w;
  if (i >= a.n) return;
  int nc = a.nrad * a.H, H = a.H, nr = a.nrad;
  T *c = reinterpret_cast<T *>(smem) + (size_t)w * work;  // work: elements per atom
  T *tab = c + nc;
  accumulate_c<T, LC, CH>(a, i, lane, c, tab);
  if constexpr (MIXED) {
    for (int k = lane; k < nc; k += 64) mo.C[(size_t)i * nc + k] = (float)c[k];
  } else {
    for (int k = lane; k < nc; k += 64) Cg[(size_t)i * nc + k] = c[k];
  }
  if (!want_b2) return;
  // B_l = C_l C_l^T on the matrix cores (C_l[n][m] = c[n, l^2 + m]); the
  // upper triangle n2 >= n1 belongs at p(n1, n2) (L + 1) + l of the row. FP64
  // and MIXED 2's FP32 row store each element there directly (the L2 merges
  // a row's scattered stores; staging would cost occupancy); FP32 and MIXED 1
  // stage the row in LDS (over the bond tables, no longer needed) and write
  // it with coalesced stores.
  int L1 = a.lmax + 1;
  int r16 = lane & 15, k4 = lane >> 4;
  int row = a.row_of_atom[i];
  T *Brow = B + (size_t)row * a.nd;
  T *bst = tab;
  T s2 = 0;
  // MIXED 2: the row's power-of-two scale comes before the row itself: |B|'s
  // largest element is a diagonal one, max_{n,l} sum_m c[n, lm]^2 (Cauchy-
  // Schwarz within each l), so the values are sliced as they are formed and
  // only the int8 slices (rows of Kq bytes, zero past nd) are staged in LDS
  // over the bond tables; the FP32 row goes straight to HBM.
  double q_inv = 0;
  long long Kq = 0;
  int8_t *qst = reinterpret_cast<int8_t *>(tab);
  if constexpr (MIXED) Kq = mo.nk * kI8Pad;
  if constexpr (MIXED == 2) {
    double mx = 0;
    for (int q = lane; q < nr * L1; q += 64) {
      int n = q / L1, l = q % L1;
      const T *cn = c + n * H + l * l;
      double sq = 0;
      for (int m = 0; m < 2 * l + 1; m++) sq += (double)cn[m] * (double)cn[m];
      mx = fmax(mx, sq);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    int E = 0;
    if (mx > 0) frexp(mx, &E);  // mx = m 2^E, m in [0.5, 1)
    q_inv = ldexp(1.0, -E);
    if (lane == 0) mo.scale[row] = ldexp(1.0, E);
    for (int q = a.nd + lane; q < Kq; q += 64)
#pragma unroll
      for (int p = 0; p < kI8Slices; p++) qst[p * Kq + q] = 0;
  }
  for (int l = 0; l <= a.lmax; l++) {
    int nm = 2 * l + 1, lm0 = l * l;
    for (int rt = 0; rt < nr; rt += 16) {
      for (int cc = rt; cc < nr; cc += 16) {
        typename MfmaAcc<T>::type acc = {};
        int n1a = rt + r16, n2b = cc + r16;
        const T *ca = c + (n1a < nr ? n1a : nr - 1) * H + lm0;
        const T *cb = c + (n2b < nr ? n2b : nr - 1) * H + lm0;
        for (int k0 = 0; k0 < nm; k0 += 8) {
          T av[2], bv[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            int m = k0 + 4 * u + k4, mc = m < nm ? m : nm - 1;
            T x = ca[mc], y = cb[mc];
            av[u] = (n1a < nr && m < nm) ? x : T(0);
            bv[u] = (n2b < nr && m < nm) ? y : T(0);
          }
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (k0 + 4 * u < nm) acc = mfma_4(av[u], bv[u], acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int n1 = rt + mfma_row<T>(lane, q), n2 = n2b;
          if (n1 < nr && n2 < nr && n2 >= n1) {
            int k = (n1 * nr - (n1 * (n1 - 1)) / 2 + (n2 - n1)) * L1 + l;
            T v = acc[q];
            if constexpr (MIXED == 2) {
              mo.B[(size_t)row * a.nd + k] = (float)v;
              double r = (double)v * q_inv;
#pragma unroll
              for (int p = 0; p < kI8Slices; p++) {
                double sl = rint(ldexp(r, 6 + 7 * p));
                r -= ldexp(sl, -(6 + 7 * p));
                qst[p * Kq + k] = (int8_t)(int)sl;
              }
            } else if constexpr (sizeof(T) == 4 || MIXED == 1) {
              bst[k] = v;
            } else {
              Brow[k] = v;
            }
            s2 += v * v;
          }
        }
      }
    }
  }
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int k = lane; k < a.nd; k += 64) Brow[k] = bst[k];
    // Pieces: 8 consecutive values per lane, one 16-byte store per piece
    // (ldp is a multiple of 32; the padding past nd is zero).
    for (int k8 = lane; k8 < ldp / 8; k8 += 64) {
      bf16x8 h0, h1;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        int k = 8 * k8 + e;
        float r = k < a.nd ? (float)bst[k] : 0.0f;
        __bf16 h = (__bf16)r;
        h0[e] = h;
        h1[e] = (__bf16)(r - (float)h);
      }
      *reinterpret_cast<bf16x8 *>(Bs.p[0] + (size_t)row * ldp + 8 * k8) = h0;
      *reinterpret_cast<bf16x8 *>(Bs.p[1] + (size_t)row * ldp + 8 * k8) = h1;
    }
  }
  if constexpr (MIXED) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    long long qrow = row + mo.shift[a.species[i]];
    if constexpr (MIXED == 1) {
      float *B32 = mo.B + (size_t)row * a.nd;
      double mx = 0;
      for (int k = lane; k < a.nd; k += 64) {
        B32[k] = (float)bst[k];
        mx = fmax(mx, fabs((double)bst[k]));
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
      int E = 0;
      if (mx > 0) frexp(mx, &E);  // mx = m 2^E, m in [0.5, 1)
      double inv = ldexp(1.0, -E);
      if (lane == 0) mo.scale[row] = ldexp(1.0, E);
      for (int k0 = 4 * lane; k0 < Kq; k0 += 256) {
        int packed[kI8Slices];
#pragma unroll
        for (int p = 0; p < kI8Slices; p++) packed[p] = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          int k = k0 + j;
          double r = k < a.nd ? (double)bst[k] * inv : 0.0;
#pragma unroll
          for (int p = 0; p < kI8Slices; p++) {
            double v = rint(ldexp(r, 6 + 7 * p));
            r -= ldexp(v, -(6 + 7 * p));
            packed[p] |= ((int)v & 0xff) << (8 * j);
          }
        }
#pragma unroll
        for (int p = 0; p < kI8Slices; p++)
          *reinterpret_cast<int *>(mo.q + i8_tile_offset(kI8TileA, mo.nk, qrow, k0, p)) =
              packed[p];
      }
    } else {
      for (int k0 = 4 * lane; k0 < Kq; k0 += 256)
#pragma unroll
        for (int p = 0; p < kI8Slices; p++)
          *reinterpret_cast<int *>(mo.q + i8_tile_offset(kI8TileA, mo.nk, qrow, k0, p)) =
              *reinterpret_cast<const int *>(&qst[p * Kq + k0]);
    }
  }
  s2 = wave_sum(s2);
  if (lane == 0) {
    norm[row] = (double)sqrt(s2);
    nn[row] = (double)s2;
  }
}

// Adjoint, first half, per atom (one wavefront): with `project` set,
// w = dE/dB = (G - (G.B_hat) B_hat)/|B| from G = dE/dB_hat |B| (for the
// sparse GP, G = sum_j sigma^2 alpha_j p S_j^(p-1) x_hat_j, the closed form of
// normalized_dot_product.cpp:191-195); without it w = G (unnormalised
// kernels). Then u = dE/dc (lammps_descriptor.cpp:286-314):
// u[n1, lm] = sum_n2 (1 + [n1 = n2]) w[(min, max)(n1, n2), l] c[n2, lm]. An
// atom with |B| < empty_norm or B.B < empty_nn contributes nothing (its u is
// not read).
struct AdjArgs {
  int project;
  double empty_norm, empty_nn;
  // With project: G . B per row from the local energies when given,
  // G . B = p |B| E (G = sum_j sigma^2 alpha_j p S_j^(p-1) x_hat_j and
  // E = sum_j sigma^2 alpha_j S_j^p), else a pass over the G and B rows.
  const double *energy;
  double power;
};

__device__ __forceinline__ int pair_index(int lo, int hi, int nr) {
  return lo * nr - (lo * (lo - 1)) / 2 + (hi - lo);
}

__host__ __device__ inline int pad4(int n) { return (n + 3) & ~3; }

// n elements from global memory to LDS: 16-byte vectors when the source is
// 16-byte aligned (the LDS destination always is), four in flight per lane.
template <typename T>
__device__ __forceinline__ void stage_row(const T *src, T *dst, int n, int lane, int nt = 64) {
  constexpr int E = 16 / sizeof(T);
  int done = 0;
  if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    const V *sv = reinterpret_cast<const V *>(src);
    V *dv = reinterpret_cast<V *>(dst);
    int nv = n / E;
#pragma unroll 4
    for (int v = lane; v < nv; v += nt) dv[v] = sv[v];
    done = nv * E;
  }
  for (int k = done + lane; k < n; k += nt) dst[k] = src[k];
}

// G . B over the elements of one row that this lane's wave takes (nt lanes
// stride the row; nt = 64: the whole row), summed over the wave: 16-byte
// loads of both rows when both are 16-byte aligned (all of a lane's vectors
// in flight together), a fixed order per lane and a fixed wave reduction.
template <typename T>
__device__ __forceinline__ T row_dot(const T *G, const T *B, int n, int lane, int nt = 64) {
  // Lane l takes the E-element groups v = l, l + nt, ... and the tail
  // elements nv E + l, ...: the same elements in the same order whether a
  // group arrives as one 16-byte load or as E scalar ones, so the sum does
  // not depend on where the rows sit in memory (an atom's row is placed
  // differently alone and in a batch).
  constexpr int E = 16 / sizeof(T);
  using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  T acc = 0;
  int nv = n / E;
  if (((reinterpret_cast<uintptr_t>(G) | reinterpret_cast<uintptr_t>(B)) & 15) == 0) {
    const V *gv = reinterpret_cast<const V *>(G), *bv = reinterpret_cast<const V *>(B);
#pragma unroll 8
    for (int v = lane; v < nv; v += nt) {
      V x = gv[v], y = bv[v];
      const T *xs = reinterpret_cast<const T *>(&x), *ys = reinterpret_cast<const T *>(&y);
#pragma unroll
      for (int e = 0; e < E; e++) acc = fma(xs[e], ys[e], acc);
    }
  } else {
#pragma unroll 4
    for (int v = lane; v < nv; v += nt)
#pragma unroll
      for (int e = 0; e < E; e++) acc = fma(G[E * v + e], B[E * v + e], acc);
  }
  for (int k = nv * E + lane; k < n; k += nt) acc = fma(G[k], B[k], acc);
  return wave_sum(acc);
}

// w = sg G - sb B (the atom's dE/dB row) into LDS, n elements: 16-byte
// loads of G and B when both rows are 16-byte aligned, four vectors of each
// in flight per lane; B is not read when sb is zero. Every element is the
// same fma(sg, G, -sb B) on either path.
template <typename T>
__device__ __forceinline__ T w_elem(T sg, T g, T sb, T b) {
  return fma(sg, g, -(sb * b));
}

template <typename T>
__device__ __forceinline__ void stage_w(const T *G, const T *B, T sg, T sb, T *w, int n,
                                        int lane, int nt = 64) {
  constexpr int E = 16 / sizeof(T);
  int done = 0;
  bool withb = sb != T(0);
  if (((reinterpret_cast<uintptr_t>(G) | (withb ? reinterpret_cast<uintptr_t>(B) : 0)) & 15) ==
      0) {
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    const V *gv = reinterpret_cast<const V *>(G), *bv = reinterpret_cast<const V *>(B);
    V *wv = reinterpret_cast<V *>(w);
    int nv = n / E;
    if (withb) {
#pragma unroll 4
      for (int v = lane; v < nv; v += nt) {
        V x = gv[v], y = bv[v], r;
        const T *xs = reinterpret_cast<const T *>(&x), *ys = reinterpret_cast<const T *>(&y);
        T *rs = reinterpret_cast<T *>(&r);
#pragma unroll
        for (int e = 0; e < E; e++) rs[e] = w_elem(sg, xs[e], sb, ys[e]);
        wv[v] = r;
      }
    } else {
#pragma unroll 4
      for (int v = lane; v < nv; v += nt) {
        V x = gv[v], r;
        const T *xs = reinterpret_cast<const T *>(&x);
        T *rs = reinterpret_cast<T *>(&r);
#pragma unroll
        for (int e = 0; e < E; e++) rs[e] = sg * xs[e];
        wv[v] = r;
      }
    }
    done = nv * E;
  }
  for (int k = done + lane; k < n; k += nt)
    w[k] = withb ? w_elem(sg, G[k], sb, B[k]) : sg * G[k];
}

// FP32 rows: G . B and w = sg G - sb B (sb = (G . B) s3) in one pass over
// both rows. Each of the atom's 128 threads stores its V-vectors of G to LDS
// as they arrive and keeps those of B (at most R, rows of at most 128 R E
// elements) in registers until sb is known; the partition and summation
// order of the dot are those of row_dot. red holds the two wavefronts'
// partial dots; the barrier inside is reached by every thread of the block
// (dead atoms load nothing).
template <typename V, int R>
__device__ __forceinline__ void stage_w_onepass(const float *G, const float *B, float sg,
                                                float s3, float *wl, int n, bool live, int t2,
                                                int lane, int w, int slot, float *red) {
  constexpr int E = sizeof(V) / sizeof(float);
  const V *gv = reinterpret_cast<const V *>(G), *bv = reinterpret_cast<const V *>(B);
  V *wv = reinterpret_cast<V *>(wl);
  int nv = live ? n / E : 0;
  V b[R];
  float acc = 0;
#pragma unroll
  for (int j = 0; j < R; j++) {
    int v = t2 + 128 * j;
    if (v < nv) {
      V g = gv[v];
      b[j] = bv[v];
      const float *xs = reinterpret_cast<const float *>(&g);
      const float *ys = reinterpret_cast<const float *>(&b[j]);
#pragma unroll
      for (int e = 0; e < E; e++) acc = fma(xs[e], ys[e], acc);
      wv[v] = g;
    }
  }
  int kt = nv * E + t2;
  bool tail = live && kt < n;
  float gt = tail ? G[kt] : 0.f, bt = tail ? B[kt] : 0.f;
  acc = fma(gt, bt, acc);
  acc = wave_sum(acc);
  if (lane == 0) red[w] = acc;
  __syncthreads();
  float sb = (red[2 * slot] + red[2 * slot + 1]) * s3;
#pragma unroll
  for (int j = 0; j < R; j++) {
    int v = t2 + 128 * j;
    if (v < nv) {
      V r = wv[v];
      float *rs = reinterpret_cast<float *>(&r);
      const float *ys = reinterpret_cast<const float *>(&b[j]);
#pragma unroll
      for (int e = 0; e < E; e++) rs[e] = w_elem(sg, rs[e], sb, ys[e]);
      wv[v] = r;
    }
  }
  if (tail) wl[kt] = w_elem(sg, gt, sb, bt);
}

// Longest row for the one-pass staging of k_dEdc<float, LC, true>.
constexpr int kOnePassRow = 128 * 32;

// LC >= 0: l_max known at compile time (one accumulator per l up to it, in
// place of kMaxL + 1); LC = -1 for any l_max. OP (FP32 only):
// projection from the rows with one pass over G and B (more registers;
// chosen where LDS, not registers, bounds the atoms per CU). Two wavefronts per atom share
// its LDS rows: both stage (twice the loads in flight per atom for the same
// LDS) and each takes half of the 16-row output tiles.
template <typename T, int LC, bool OP = false>
__global__ __launch_bounds__(128 * kAtomsPerBlock) void k_dEdc(B2Args a, const T *B,
                                                                 const double *norm,
                                                                 const double *nnv, const T *G,
                                                                 long long g_ld, AdjArgs aa,
                                                                 const T *Cg, T *Ug) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ T red[2 * kAtomsPerBlock];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, slot = w >> 1, sub = w & 1;
  int t2 = threadIdx.x & 127;  // thread of the atom's two wavefronts
  int i = blockIdx.x * (blockDim.x >> 7) + slot;
  int nc = a.nrad * a.H, ncp = pad4(nc), ndp = pad4(a.nd);
  T *c = reinterpret_cast<T *>(smem) + (size_t)slot * ((size_t)ncp + (size_t)ndp);
  T *wl = c + ncp;  // the w = dE/dB row
  // Atoms past the end or empty skip the work but stay for the barriers.
  bool live = i < a.n;
  int row = live ? a.row_of_atom[i] : 0;
  double nrm = live ? norm[row] : 1.0;
  live = live && !(nrm < aa.empty_norm || nnv[row] < aa.empty_nn);
  const T *Brow = B + (size_t)row * a.nd;
  const T *Grow = G + (size_t)row * g_ld;
  // w = dE/dB = (G - (G . B_hat) B_hat) / |B| = sg G - sb B.
  T ninv = T(1) / (T)nrm;
  T sg = aa.project ? ninv : T(1), sb = 0;
  if constexpr (OP && sizeof(T) == 4) {
    // The launch guarantees project, no energies and nd <= kOnePassRow.
    if (live) stage_row(Cg + (size_t)i * nc, c, nc, t2, 128);
    // Every row 16-byte aligned: the same for all atoms of the grid, so the
    // whole block takes one branch and one barrier.
    bool al = ((reinterpret_cast<uintptr_t>(G) | reinterpret_cast<uintptr_t>(B)) & 15) == 0 &&
              (g_ld & 3) == 0 && (a.nd & 3) == 0;
    if (al)
      stage_w_onepass<float4, kOnePassRow / 512>(Grow, Brow, sg, ninv * ninv * ninv, wl, a.nd,
                                                 live, t2, lane, w, slot, red);
    else
      stage_w_onepass<float, kOnePassRow / 128>(Grow, Brow, sg, ninv * ninv * ninv, wl, a.nd,
                                                live, t2, lane, w, slot, red);
  } else {
    if (aa.project) {
      if (aa.energy) {
        if (live) sb = (T)(aa.power * aa.energy[row] / (nrm * nrm));
      } else {
        T part = live ? row_dot(Grow, Brow, a.nd, t2, 128) : T(0);
        if (lane == 0) red[w] = part;
        __syncthreads();
        sb = (red[2 * slot] + red[2 * slot + 1]) * ninv * ninv * ninv;
      }
    }
    // One pass over HBM with 16-byte loads: c and w to LDS.
    if (live) {
      stage_row(Cg + (size_t)i * nc, c, nc, t2, 128);
      stage_w(Grow, Brow, sg, sb, wl, a.nd, t2, 128);
    }
  }
  __syncthreads();
  if (!live) return;
  // u_l = W_l C_l per l on the matrix cores: W_l[n1][n2] = (1 + [n1 = n2])
  // w[p(min, max), l] (nr x nr, symmetric), C_l[n2][m] = c[n2, l^2 + m]
  // (nr x (2l + 1)); 16 x 16 output tiles, 4 n2 per MFMA. Lane L supplies
  // A[L % 16][L / 16] and B[L / 16][L % 16]; it holds output rows
  // tile_row(L, i), column L % 16. The l loop is innermost (one accumulator
  // per l) so the pair index of (n1, n2) is formed once for all l.
  int L1 = a.lmax + 1, nr = a.nrad;
  T *urow = Ug + (size_t)i * nc;
  int r16 = lane & 15, k4 = lane >> 4;
  for (int ct = 0; ct < 2 * a.lmax + 1; ct += 16) {
    for (int rt = 16 * sub; rt < nr; rt += 32) {
      constexpr int KL = LC >= 0 ? LC : kMaxL;
      typename MfmaAcc<T>::type acc[KL + 1];
#pragma unroll
      for (int l = 0; l <= KL; l++) acc[l] = {};
      int n1 = rt + r16, m = ct + r16;
      for (int k0 = 0; k0 < nr; k0 += 4) {
        int n2 = k0 + k4;
        bool ok = n1 < nr && n2 < nr;
        int lo = n1 < n2 ? n1 : n2, hi = n1 < n2 ? n2 : n1;
        int p0 = ok ? pair_index(lo, hi, nr) * L1 : 0;
        T coef = n1 == n2 ? T(2) : T(1);
        const T *cr = c + (n2 < nr ? n2 : 0) * a.H;
        // Loads clamped in bounds and zeroed by selects (no divergent branches).
#pragma unroll
        for (int l = 0; l <= KL; l++) {
          int nm = 2 * l + 1;
          if (l <= a.lmax && ct < nm) {
            int mc = m < nm ? m : nm - 1;
            T g = wl[p0 + l], x = cr[l * l + mc];
            T av = ok ? coef * g : T(0);
            T bv = (n2 < nr && m < nm) ? x : T(0);
            acc[l] = mfma_4(av, bv, acc[l]);
          }
        }
      }
#pragma unroll
      for (int l = 0; l <= KL; l++) {
        int nm = 2 * l + 1;
        if (l <= a.lmax && ct < nm) {
#pragma unroll
          for (int q = 0; q < 4; q++) {
            int row = rt + mfma_row<T>(lane, q);
            if (row < nr && m < nm) urow[row * a.H + l * l + m] = acc[l][q];
          }
        }
      }
    }
  }
}

__device__ __forceinline__ void load4_lds(const float *p, float (&o)[4]) {
  float4 v = *reinterpret_cast<const float4 *>(p);
  o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w;
}
__device__ __forceinline__ void load4_lds(const double *p, double (&o)[4]) {
  double2 v0 = *reinterpret_cast<const double2 *>(p);
  double2 v1 = *reinterpret_cast<const double2 *>(p + 2);
  o[0] = v0.x, o[1] = v0.y, o[2] = v1.x, o[3] = v1.y;
}

// Adjoint, second half: per bond f = dE_i/dr_ij = sum_{n,lm} u[s_j n, lm]
// grad(g_n Y_lm) (pair_flare.cpp:141-150, the adjoint of b2.cpp:219-228).
// One wavefront per atom, lanes over its bonds; each lane sweeps every
// (l, m) of its bond (ylm_sweep with gradients) and contracts with u staged
// in LDS:
//   f = a u_hat + (P - (u_hat . P) u_hat) / r,
//   a = sum_lm (sum_n u dg_n) Y_lm,  P = sum_lm (sum_n u g_n) grad_u Y_lm,
// with grad_u the gradient of the degree-0 extension in u. Writes f for every
// bond and the atom's virial -sum_b f (x) r_b (xx, xy, xz, yy, yz, zz).
template <typename T, int LC>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_bond_forces(B2Args a,
                                                                       const double *norm,
                                                                       const double *nnv,
                                                                       AdjArgs aa, const T *Ug,
                                                                       T *fbond, double *vir,
                                                                       const int *atom_list,
                                                                       int n_list) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int k = blockIdx.x * (blockDim.x >> 6) + w;
  if (k >= n_list) return;
  int i = atom_list ? atom_list[k] : k;  // all atoms, or those of a row tile
  int nc = a.nrad * a.H, Np = pad4(a.N);
  // u in LDS as [species][lm][n] with n padded to 4, so a lane reads the
  // radial coefficients of one (species, lm) with 16-byte loads.
  T *u = reinterpret_cast<T *>(smem) + (size_t)w * a.nos * a.H * Np;
  int row = a.row_of_atom[i];
  int b0 = a.bond_off[i], b1 = a.bond_off[i + 1];
  double nrm = norm[row];
  if (nrm < aa.empty_norm || nnv[row] < aa.empty_nn) {
    // Empty environment: no force contribution (normalized_dot_product.cpp:171,
    // pair_flare.cpp:134-136).
    for (int b = b0 + lane; b < b1; b += 64) {
      fbond[3 * (size_t)b] = 0;
      fbond[3 * (size_t)b + 1] = 0;
      fbond[3 * (size_t)b + 2] = 0;
    }
    if (lane < 6) vir[6 * (size_t)i + lane] = 0;
    return;
  }
  for (int k = lane; k < nc; k += 64) {
    int n1 = k / a.H, lm = k - n1 * a.H, sj = n1 / a.N, n = n1 - sj * a.N;
    u[((size_t)sj * a.H + lm) * Np + n] = Ug[(size_t)i * nc + k];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int si = a.species[i];
  int L = a.lmax, H = a.H, N = a.N;
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int bb = b0; bb < b1; bb += 64) {
    int b = bb + lane;
    if (b < b1) {
      int q = a.bond_q[b];
      int sj = a.nspec[q];
      const double *rp = a.rel + 4 * (size_t)q;
      T r = (T)rp[0], x = (T)rp[1], y = (T)rp[2], z = (T)rp[3];
      T rc = (T)a.cutoffs[si * a.nos + sj];
      T g[kMaxN], dg[kMaxN];
      radial_basis<T>(r, rc, (T)a.r1, N, a.cutkind, g, dg);
      T rinv = T(1) / r;
      T ux = x * rinv, uy = y * rinv, uz = z * rinv;
      const T *ub = u + (size_t)sj * H * Np;
      T acc_a = 0, Px = 0, Py = 0, Pz = 0;
      ylm_sweep<T, true, LC>(L, ux, uy, uz, [&](int lm, T Y, T px, T py, T pz) {
        T s0 = 0, s1 = 0;
        const T *ul = ub + (size_t)lm * Np;
#pragma unroll
        for (int n4 = 0; n4 < kMaxN; n4 += 4) {
          if (n4 < N) {
            T uu[4];
            load4_lds(ul + n4, uu);
#pragma unroll
            for (int e = 0; e < 4; e++)
              if (n4 + e < N) {
                s0 += uu[e] * g[n4 + e];
                s1 += uu[e] * dg[n4 + e];
              }
          }
        }
        acc_a += s1 * Y;
        Px += s0 * px;
        Py += s0 * py;
        Pz += s0 * pz;
      });
      T up = ux * Px + uy * Py + uz * Pz;
      T fx = acc_a * ux + (Px - up * ux) * rinv;
      T fy = acc_a * uy + (Py - up * uy) * rinv;
      T fz = acc_a * uz + (Pz - up * uz) * rinv;
      fbond[3 * (size_t)b] = fx;
      fbond[3 * (size_t)b + 1] = fy;
      fbond[3 * (size_t)b + 2] = fz;
      double f3[3] = {(double)fx, (double)fy, (double)fz};
      double r3[3] = {rp[1], rp[2], rp[3]};
      int sc = 0;
      for (int c1 = 0; c1 < 3; c1++)
        for (int c2 = c1; c2 < 3; c2++) v[sc++] -= f3[c1] * r3[c2];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
    for (int k = 0; k < 6; k++) vir[6 * (size_t)i + k] = v[k];
}

// Atoms (wavefronts) per block: of 1, 2 and 4, the one that keeps the most
// atoms resident on a CU within its 160 KB of LDS (larger on ties).
static int atoms_per_block(size_t lds_per_atom) {
  const size_t cap = 160 * 1024;
  FM_CHECK(lds_per_atom <= cap, FM_ERR_CAPACITY, "per-atom descriptor workspace exceeds LDS");
  int best = 1;
  size_t best_res = 0;
  for (int apb = 1; apb <= kAtomsPerBlock; apb <<= 1) {
    size_t blocks = cap / (apb * lds_per_atom);
    size_t res = std::min<size_t>(blocks, 8) * apb;  // at most 8 blocks of ours per CU here
    if (blocks > 0 && res >= best_res) {
      best = apb;
      best_res = res;
    }
  }
  return best;
}

// Atoms resident per CU with lds_per_atom bytes each (the residency
// atoms_per_block maximises).
static size_t resident_atoms(size_t lds_per_atom) {
  const size_t cap = 160 * 1024;
  if (lds_per_atom > cap) return 0;
  int apb = atoms_per_block(lds_per_atom);
  return std::min<size_t>(cap / (apb * lds_per_atom), 8) * apb;
}

// Calls f(std::integral_constant<int, L>) for the l_max values with an
// unrolled harmonic sweep (those of the common B2 / B3 settings), else with
// L = -1 (the sweep loops over the run-time l_max).
template <typename F>
static void with_lmax(int lmax, F &&f) {
  switch (lmax) {
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    default: f(std::integral_constant<int, -1>{}); break;
  }
}

static B2Args make_args(fm_structure s) {
  B2Args a;
  a.n = s->n;
  a.nos = s->nos;
  a.N = s->N;
  a.lmax = s->lmax;
  a.H = (s->lmax + 1) * (s->lmax + 1);
  a.nrad = s->nos * s->N;
  a.nd = s->nd;
  a.cutkind = s->cutkind;
  a.r1 = s->r1;
  a.species = s->species.as<int>();
  a.bond_off = s->bond_off.as<int>();
  a.bond_q = s->bond_q.as<int>();
  a.nspec = s->nspecies.as<int>();
  a.row_of_atom = s->row_of_atom.as<int>();
  a.rel = s->rel.as<double>();
  a.cutoffs = s->d_cutoffs.as<double>();
  a.pairs = s->pair_table.as<int2>();
  return a;
}

void knorm_values(int lmax, double *out) {
  for (int l = 0; l <= lmax; l++)
    for (int m = -l; m <= l; m++) {
      int mm = m < 0 ? -m : m;
      double ratio = 1;
      for (int k = l - mm + 1; k <= l + mm; k++) ratio /= k;
      double K = std::sqrt((2 * l + 1) / (4 * 3.14159265358979323846) * ratio);
      if (mm > 0) K *= std::sqrt(2.0);
      out[l * l + m + l] = K;
    }
}

// The whole table (l <= 10) is written every time: entries do not depend on
// the descriptor's lmax, so uploads from several streams write equal values.
static void upload_knorm(hipStream_t st, int) {
  static thread_local double host[kMaxH];
  knorm_values(kMaxL, host);
  FM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_knorm), host, sizeof(host), 0, hipMemcpyHostToDevice,
                                st));
}

static void upload_rcp(hipStream_t st) {
  static thread_local double host[kMaxL + 2];
  host[0] = 0;
  for (int k = 1; k < kMaxL + 2; k++) host[k] = 1.0 / k;
  FM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_rcp), host, sizeof(host), 0, hipMemcpyHostToDevice,
                                st));
}

void launch_b2(fm_structure s, bool want_b2) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  FM_CHECK(s->N <= kMaxN, FM_ERR_INVALID, "n_max above 16 is not supported");
  FM_CHECK(s->lmax <= kMaxL, FM_ERR_INVALID, "l_max above 10 is not supported");
  upload_knorm(st, s->lmax);
  upload_rcp(st);
  B2Args a = make_args(s);
  // The chosen form's B2 rows are computed in FP64 (mixed: FP32 row and c
  // for the adjoint, int8 slices of the FP64 row for the kernel
  // contraction); c alone (the B3 bond pass) stays FP32.
  bool mixed = want_b2 && s->form != FM_FORM_FP64;
  size_t esz = s->form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
  size_t csz = s->form == FM_FORM_FP64 || mixed ? sizeof(double) : sizeof(float);
  // FP32 and mixed mode 1 stage the B row after c (over the bond tables);
  // mixed mode 2 stages only the row's int8 slices (kI8Slices x Kq bytes).
  auto work_of = [&](int ch, int mode) {
    size_t w = b2_work(a.nrad * a.H, a.H, a.N, ch);
    if (want_b2 && s->form != FM_FORM_FP64 && mode != 2)
      w = std::max(w, (size_t)a.nrad * a.H + (size_t)a.nd);
    if (mode == 2)
      w = std::max(w, (size_t)a.nrad * a.H +
                          ((size_t)kI8Slices * (size_t)i8_ld(a.nd) + csz - 1) / csz);
    return w;
  };
  // 64-bond chunks take half the chunk passes of 32-bond ones; they are used
  // when they cost at most a third more LDS and keep as many atoms resident
  // per CU (a staged row makes the two sizes close in LDS, where only the
  // residency tells them apart).
  auto chunk_of = [&](int mode) {
    bool wide = 3 * work_of(64, mode) <= 4 * work_of(32, mode) &&
                resident_atoms(work_of(64, mode) * csz) >= resident_atoms(work_of(32, mode) * csz);
    return wide ? 64 : 32;
  };
  // Mixed mode 2 (slices formed with the values, FP32 row stored directly)
  // spends more ALU per element than mode 1 (FP64 row staged, sliced after)
  // and pays off only where its smaller LDS footprint keeps more atoms
  // resident per CU.
  int mode = 0;
  if (mixed) {
    size_t w1 = work_of(chunk_of(1), 1), w2 = work_of(chunk_of(2), 2);
    mode = resident_atoms(w2 * csz) > resident_atoms(w1 * csz) ? 2 : 1;
  }
  int chunk = chunk_of(mode);
  bool wide = chunk == 64;
  size_t work = work_of(chunk, mode);
  size_t per = work * csz;
  int apb = atoms_per_block(per);
  size_t lds = apb * per;
  int nblk = (s->n + apb - 1) / apb;
  s->C.alloc(&ctx->pool, std::max<size_t>(1, (size_t)s->n * a.nrad * a.H * esz));
  MixedOut mo{};
  if (mixed) {
    // Each species' rows start a row tile of the slices.
    s->ldq = i8_ld(a.nd);
    s->q_tile_row.assign(s->nos, 0);
    std::vector<int> &shift = s->q_shift_host;  // kept with the structure for the async copy
    shift.assign(s->nos, 0);
    long long qrows = 0;
    for (int t = 0; t < s->nos; t++) {
      s->q_tile_row[t] = qrows;
      shift[t] = (int)(qrows - s->type_off[t]);
      qrows += (s->n_by_type[t] + kI8TileA - 1) / kI8TileA * kI8TileA;
    }
    FM_CHECK(qrows < (1LL << 31), FM_ERR_CAPACITY, "too many descriptor rows for the slices");
    s->Bq.alloc(&ctx->pool, std::max<size_t>(1, i8_tiled_bytes(qrows, kI8TileA, s->ldq)));
    s->Bq_scale.alloc(&ctx->pool, std::max<size_t>(1, (size_t)s->n * sizeof(double)));
    s->q_shift.alloc(&ctx->pool, (size_t)s->nos * sizeof(int));
    FM_HIP(hipMemcpyAsync(s->q_shift.as<int>(), shift.data(), shift.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    mo = MixedOut{s->B.as<float>(), s->C.as<float>(), s->Bq.as<int8_t>(),
                  s->Bq_scale.as<double>(), s->ldq / kI8Pad, s->q_shift.as<int>()};
  } else {
    s->Bq.release();
    s->Bq_scale.release();
    s->q_shift.release();
  }
  s->has_q = mixed;
  if (nblk == 0) return;
  // No BF16 copies of the rows: the chosen form's kernel contraction slices
  // the FP32 rows itself.
  SplitOut none{{nullptr, nullptr}};
  // The unrolled sweep (compile-time l_max) is faster for the FP64 kernel;
  // for FP32 its extra registers cost more occupancy than it saves at
  // l_max 6.
  auto go = [&](auto ch) {
    constexpr int CH = decltype(ch)::value;
    if (s->form == FM_FORM_FP64)
      with_lmax(a.lmax, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        hipLaunchKernelGGL((k_b2<double, LC, CH>), dim3(nblk), dim3(64 * apb), lds, st, a,
                           s->B.as<double>(), s->norm.as<double>(), s->nn.as<double>(), none, 0,
                           s->C.as<double>(), (int)want_b2, (int)work, mo);
      });
    else if (mode == 1)
      with_lmax(a.lmax, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        hipLaunchKernelGGL((k_b2<double, LC, CH, 1>), dim3(nblk), dim3(64 * apb), lds, st,
                           a, (double *)nullptr, s->norm.as<double>(), s->nn.as<double>(), none,
                           0, (double *)nullptr, (int)want_b2, (int)work, mo);
      });
    else if (mode == 2)
      with_lmax(a.lmax, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        hipLaunchKernelGGL((k_b2<double, LC, CH, 2>), dim3(nblk), dim3(64 * apb), lds, st,
                           a, (double *)nullptr, s->norm.as<double>(), s->nn.as<double>(), none,
                           0, (double *)nullptr, (int)want_b2, (int)work, mo);
      });
    else
      with_lmax(a.lmax <= 4 ? a.lmax : -1, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        hipLaunchKernelGGL((k_b2<float, LC, CH>), dim3(nblk), dim3(64 * apb), lds, st, a,
                           s->B.as<float>(), s->norm.as<double>(), s->nn.as<double>(), none, 0,
                           s->C.as<float>(), (int)want_b2, (int)work, mo);
      });
  };
  {
    KScope timer(ctx, KT_DESC);
    if (wide)
      go(std::integral_constant<int, 64>{});
    else
      go(std::integral_constant<int, 32>{});
  }
  FM_HIP(hipGetLastError());
}

void launch_adjoint(fm_structure s, const void *G, int64_t g_ld, void *fbond,
                    double *virial_atom, int project, double empty_norm, double empty_nn,
                    const double *energy, double power, void *u_buf) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  B2Args a = make_args(s);
  FM_CHECK(a.lmax + 1 < kMaxL + 2, FM_ERR_INVALID, "l_max above 10 is not supported");
  upload_knorm(st, s->lmax);
  upload_rcp(st);
  bool f64 = s->form == FM_FORM_FP64;
  size_t esz = f64 ? sizeof(double) : sizeof(float);
  size_t nc = (size_t)a.nrad * a.H;
  DevBuf U;
  if (!u_buf) {
    U.alloc(&ctx->pool, adjoint_u_bytes(s));
    u_buf = U.as<void>();
  }
  AdjArgs aa{project, empty_norm, empty_nn, energy, power};
  // dE/dc (B2): c and w rows in LDS per atom (each padded to 4 values).
  size_t per1 = ((size_t)pad4((int)nc) + (size_t)pad4(a.nd)) * esz;
  int apb1 = s->desc_kind == 2 ? atoms_per_block(per1) : 1;
  int nblk1 = (s->n + apb1 - 1) / apb1;
  if (s->n == 0) return;
  if (s->desc_kind == 3)
    launch_dEdc_b3(s, G, g_ld, project, empty_norm, empty_nn, u_buf);
  // FP32 projection from the rows in one pass over G and B where LDS
  // already holds the CU to two blocks (the one-pass registers cost no
  // occupancy there; with three or more blocks they would).
  bool one_pass = !f64 && project && !energy && a.nd <= kOnePassRow &&
                  3 * (size_t)apb1 * per1 > (size_t)160 * 1024;
  if (s->desc_kind == 2) {
    KScope timer(ctx, KT_DEDC);
    with_lmax(a.lmax, [&](auto lc) {
      constexpr int LC = decltype(lc)::value;
      if (f64)
        hipLaunchKernelGGL((k_dEdc<double, LC>), dim3(nblk1), dim3(128 * apb1), apb1 * per1, st,
                           a, s->B.as<double>(), s->norm.as<double>(), s->nn.as<double>(),
                           (const double *)G, (long long)g_ld, aa, s->C.as<double>(),
                           (double *)u_buf);
      else if (one_pass)
        hipLaunchKernelGGL((k_dEdc<float, LC, true>), dim3(nblk1), dim3(128 * apb1),
                           apb1 * per1, st, a, s->B.as<float>(), s->norm.as<double>(),
                           s->nn.as<double>(), (const float *)G, (long long)g_ld, aa,
                           s->C.as<float>(), (float *)u_buf);
      else
        hipLaunchKernelGGL((k_dEdc<float, LC>), dim3(nblk1), dim3(128 * apb1), apb1 * per1, st,
                           a, s->B.as<float>(), s->norm.as<double>(), s->nn.as<double>(),
                           (const float *)G, (long long)g_ld, aa, s->C.as<float>(),
                           (float *)u_buf);
    });
  }
  FM_HIP(hipGetLastError());
  launch_bond_forces(s, u_buf, fbond, virial_atom, nullptr, s->n, empty_norm, empty_nn);
}

size_t adjoint_u_bytes(fm_structure s) {
  B2Args a = make_args(s);
  size_t esz = s->form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
  return std::max<size_t>(1, (size_t)s->n * a.nrad * a.H * esz);
}

void launch_bond_forces(fm_structure s, const void *U, void *fbond, double *virial_atom,
                        const int *atom_list, int n_list, double empty_norm, double empty_nn) {
  if (n_list == 0) return;
  hipStream_t st = s->ctx->stream;
  B2Args a = make_args(s);
  upload_knorm(st, s->lmax);
  upload_rcp(st);
  bool f64 = s->form == FM_FORM_FP64;
  size_t esz = f64 ? sizeof(double) : sizeof(float);
  AdjArgs aa{1, empty_norm, empty_nn, nullptr, 0};
  // u in LDS per atom (radial index padded to 4).
  size_t per2 = (size_t)a.nos * a.H * pad4(a.N) * esz;
  int apb2 = atoms_per_block(per2);
  int nblk2 = (n_list + apb2 - 1) / apb2;
  KScope timer(s->ctx, KT_BOND);
  if (f64)
    hipLaunchKernelGGL((k_bond_forces<double, -1>), dim3(nblk2), dim3(64 * apb2), apb2 * per2,
                       st, a, s->norm.as<double>(), s->nn.as<double>(), aa, (const double *)U,
                       (double *)fbond, virial_atom, atom_list, n_list);
  else
    hipLaunchKernelGGL((k_bond_forces<float, -1>), dim3(nblk2), dim3(64 * apb2), apb2 * per2,
                       st, a, s->norm.as<double>(), s->nn.as<double>(), aa, (const float *)U,
                       (float *)fbond, virial_atom, atom_list, n_list);
  FM_HIP(hipGetLastError());
}

}  // namespace fm
