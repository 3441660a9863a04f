// Device-side helpers shared by the gfx950 kernels: wavefront reductions and
// the bond basis (Chebyshev radial x cutoff, real spherical harmonics).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fm {

constexpr int kWave = 64;
constexpr double kEmptyThresh = 1e-8;  // normalized_dot_product.cpp:45,131

// Largest radial count per species and lmax the bond kernels take.
constexpr int kMaxN = 16;
constexpr int kMaxL = 10;
constexpr int kMaxH = (kMaxL + 1) * (kMaxL + 1);

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Chebyshev values T_n(x) and dT_n/dr on x = (r - r1) / (rc - r1)
// (radial.cpp:138-169), times the cutoff f_c(r) (cutoffs.cpp:44-73).
// g[n] = T_n f_c, dg[n] = T_n' f_c + T_n f_c' (so grad g_n = dg[n] r_hat).
// Outside [r1, rc] both are zero, as the reference's zero-initialised basis.
template <typename T>
__device__ __forceinline__ void radial_basis(T r, T rc, T r1, int N, int cutkind,
                                             T (&g)[kMaxN], T (&dg)[kMaxN]) {
  T fc, dfc;
  if (r > rc) {
    fc = 0;
    dfc = 0;
  } else if (cutkind == 0) {
    T rd = r - rc;
    fc = rd * rd;
    dfc = 2 * rd;
  } else {
    const T pi = T(3.14159265358979323846);
    fc = T(0.5) * (cos(pi * r / rc) + 1);
    dfc = -pi * sin(pi * r / rc) / (2 * rc);
  }
  bool inside = !((r < r1) || (r > rc));
  T c = T(1) / (rc - r1);
  T x = (r - r1) * c;
  T tm2 = 0, tm1 = 0, dm2 = 0, dm1 = 0;
#pragma unroll
  for (int n = 0; n < kMaxN; n++) {
    T t, d;
    if (n == 0) {
      t = 1;
      d = 0;
    } else if (n == 1) {
      t = x;
      d = c;
    } else {
      t = 2 * x * tm1 - tm2;
      d = 2 * tm1 * c + 2 * x * dm1 - dm2;
    }
    tm2 = tm1;
    tm1 = t;
    dm2 = dm1;
    dm1 = d;
    if (n < N) {
      g[n] = inside ? t * fc : T(0);
      dg[n] = inside ? d * fc + t * dfc : T(0);
    } else {
      g[n] = 0;
      dg[n] = 0;
    }
  }
}

// One real spherical harmonic Y_lm (no Condon-Shortley phase: l = 1 gives
// y, z, x as y_grad.cpp:718-736) and optionally its Cartesian gradient, for a
// lane that owns (l, m). Y_lm = K_lm Q_l^|m|(u_z) {Re, Im}(u_x + i u_y)^|m|
// with Q_l^m = d^m P_l/dt^m; the gradient of the degree-0 extension is
// (grad_u P - (u . grad_u P) u) / r.
template <typename T, bool kGrad>
__device__ __forceinline__ void sph_harm_lane(int l, int m, T knorm, T ux, T uy,
                                              T uz, T rinv, T &Y, T &Yx, T &Yy,
                                              T &Yz) {
  int mm = m < 0 ? -m : m;
  // (ux + i uy)^k for k = mm - 1 and mm.
  T cm1 = 1, sm1 = 0, cm = 1, sm = 0;
  for (int k = 1; k <= mm; k++) {
    cm1 = cm;
    sm1 = sm;
    cm = cm1 * ux - sm1 * uy;
    sm = sm1 * ux + cm1 * uy;
  }
  // Q_l^mm and Q_l^(mm+1) by the upward recursion in l.
  T q = 1, dq = 0;
  {
    T qmm = 1;
    for (int k = 1; k <= 2 * mm - 1; k += 2) qmm *= k;
    T qa = qmm, qb = (2 * mm + 1) * uz * qmm;  // Q_mm^mm, Q_{mm+1}^mm
    if (l == mm) q = qa;
    else if (l == mm + 1) q = qb;
    else {
      for (int ll = mm + 2; ll <= l; ll++) {
        T qn = ((2 * ll - 1) * uz * qb - (ll + mm - 1) * qa) / (ll - mm);
        qa = qb;
        qb = qn;
      }
      q = qb;
    }
    if (kGrad) {
      int m1 = mm + 1;
      if (l < m1) dq = 0;
      else {
        T pmm = qmm * (2 * m1 - 1);  // (2 m1 - 1)!!
        T pa = pmm, pb = (2 * m1 + 1) * uz * pmm;
        if (l == m1) dq = pa;
        else if (l == m1 + 1) dq = pb;
        else {
          for (int ll = m1 + 2; ll <= l; ll++) {
            T pn = ((2 * ll - 1) * uz * pb - (ll + m1 - 1) * pa) / (ll - m1);
            pa = pb;
            pb = pn;
          }
          dq = pb;
        }
      }
    }
  }
  T trig, dtx, dty;
  if (m > 0) {
    trig = cm;
    dtx = mm * cm1;
    dty = -mm * sm1;
  } else if (m < 0) {
    trig = sm;
    dtx = mm * sm1;
    dty = mm * cm1;
  } else {
    trig = 1;
    dtx = 0;
    dty = 0;
  }
  Y = knorm * q * trig;
  if (kGrad) {
    T px = knorm * q * dtx, py = knorm * q * dty, pz = knorm * dq * trig;
    T ud = ux * px + uy * py + uz * pz;
    Yx = (px - ud * ux) * rinv;
    Yy = (py - ud * uy) * rinv;
    Yz = (pz - ud * uz) * rinv;
  }
}

}  // namespace fm
