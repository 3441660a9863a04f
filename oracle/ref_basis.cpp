// Oracle: bond basis — Chebyshev radial functions, cutoff functions and real
// spherical harmonics with Cartesian gradients (test infrastructure only).
#include <cmath>

#include "ref_cpu.hpp"

namespace ref {

static const double kPi = 3.14159265358979323846;

// radial.cpp:138-169. Outside [r1, r2] the values are left untouched; the
// caller (radial.cpp:357-360) hands in zero-initialised vectors.
void chebyshev(double *vals, double *derivs, double r, int N, double r1, double r2) {
  if ((r < r1) || (r > r2)) return;
  double c = 1 / (r2 - r1);
  double x = (r - r1) * c;
  for (int n = 0; n < N; n++) {
    if (n == 0) {
      vals[n] = 1;
      derivs[n] = 0;
    } else if (n == 1) {
      vals[n] = x;
      derivs[n] = c;
    } else {
      vals[n] = 2 * x * vals[n - 1] - vals[n - 2];
      derivs[n] = 2 * vals[n - 1] * c + 2 * x * derivs[n - 1] - derivs[n - 2];
    }
  }
}

// cutoffs.cpp:44-56 (quadratic) and :58-73 (cosine).
void cutoff_fn(int kind, double r, double rc, double &fc, double &dfc) {
  if (r > rc) {
    fc = 0;
    dfc = 0;
    return;
  }
  if (kind == CUT_QUADRATIC) {
    double rd = r - rc;
    fc = rd * rd;
    dfc = 2 * rd;
  } else {
    fc = 0.5 * (std::cos(kPi * r / rc) + 1);
    dfc = -kPi * std::sin(kPi * r / rc) / (2 * rc);
  }
}

// radial.cpp:344-379: g_n = T_n(r) f_c(r), grad g_n = (T_n' f_c + T_n f_c') r_hat.
void radial_basis(double *g, double *gx, double *gy, double *gz, double x,
                  double y, double z, double r, double rc, int N, double r1,
                  int cutoff_kind) {
  double fc, dfc;
  cutoff_fn(cutoff_kind, r, rc, fc, dfc);
  double bv[64], bd[64];
  for (int n = 0; n < N; n++) bv[n] = bd[n] = 0;
  chebyshev(bv, bd, r, N, r1, rc);
  double xr = x / r, yr = y / r, zr = z / r;
  for (int n = 0; n < N; n++) {
    g[n] = bv[n] * fc;
    gx[n] = bd[n] * xr * fc + bv[n] * xr * dfc;
    gy[n] = bd[n] * yr * fc + bv[n] * yr * dfc;
    gz[n] = bd[n] * zr * fc + bv[n] * zr * dfc;
  }
}

// Real spherical harmonics without the Condon-Shortley phase, the convention
// of y_grad.cpp:667-2754 (l = 1 gives y, z, x). Written from the textbook
// form Y_lm = K_lm Q_l^|m|(u_z) {Re, Im}(u_x + i u_y)^|m|, where
// Q_l^m = d^m P_l / dt^m, and the gradient
// grad Y = (grad_u P - (u . grad_u P) u) / r of the degree-0 extension.
void real_sph_harm(double *Y, double *Yx, double *Yy, double *Yz, double x,
                   double y, double z, int lmax) {
  double r = std::sqrt(x * x + y * y + z * z);
  double ux = x / r, uy = y / r, uz = z / r;
  double C[24], S[24];
  C[0] = 1;
  S[0] = 0;
  for (int m = 1; m <= lmax + 1; m++) {
    C[m] = C[m - 1] * ux - S[m - 1] * uy;
    S[m] = S[m - 1] * ux + C[m - 1] * uy;
  }
  // Q[l][m] for m <= l + 1.
  double Q[24][24];
  for (int l = 0; l <= lmax; l++)
    for (int m = 0; m <= lmax + 1; m++) Q[l][m] = 0;
  for (int m = 0; m <= lmax; m++) {
    double qmm = 1;
    for (int k = 1; k <= 2 * m - 1; k += 2) qmm *= k;
    Q[m][m] = qmm;
    if (m + 1 <= lmax) Q[m + 1][m] = (2 * m + 1) * uz * qmm;
    for (int l = m + 2; l <= lmax; l++)
      Q[l][m] = ((2 * l - 1) * uz * Q[l - 1][m] - (l + m - 1) * Q[l - 2][m]) / (l - m);
  }
  for (int l = 0; l <= lmax; l++) {
    for (int m = -l; m <= l; m++) {
      int mm = m < 0 ? -m : m;
      double ratio = 1;
      for (int k = l - mm + 1; k <= l + mm; k++) ratio /= k;
      double K = std::sqrt((2 * l + 1) / (4 * kPi) * ratio);
      if (mm > 0) K *= std::sqrt(2.0);
      double trig, dtx, dty;
      if (m > 0) {
        trig = C[mm];
        dtx = mm * C[mm - 1];
        dty = -mm * S[mm - 1];
      } else if (m < 0) {
        trig = S[mm];
        dtx = mm * S[mm - 1];
        dty = mm * C[mm - 1];
      } else {
        trig = 1;
        dtx = 0;
        dty = 0;
      }
      double q = Q[l][mm], dq = (mm + 1 <= l) ? Q[l][mm + 1] : 0.0;
      double P = K * q * trig;
      double px = K * q * dtx, py = K * q * dty, pz = K * dq * trig;
      double ud = ux * px + uy * py + uz * pz;
      int idx = l * l + m + l;
      Y[idx] = P;
      Yx[idx] = (px - ud * ux) / r;
      Yy[idx] = (py - ud * uy) / r;
      Yz[idx] = (pz - ud * uz) / r;
    }
  }
}

}  // namespace ref
