// Oracle helper (test infrastructure only): C entry points over the
// reference's own radial.cpp / cutoffs.cpp, compiled where they lie in the
// reference tree into oracle/_ref/. Used to pin the oracle's Chebyshev and
// cutoff restatement.
#include <cstring>
#include <vector>

#include "cutoffs.h"
#include "radial.h"

extern "C" {
// out: vals[N], derivs[N]
void refsrc_chebyshev(double r, int N, double r1, double r2, double *out) {
  std::vector<double> v(N, 0), d(N, 0);
  chebyshev(v, d, r, N, std::vector<double>{r1, r2});
  std::memcpy(out, v.data(), N * 8);
  std::memcpy(out + N, d.data(), N * 8);
}
// kind 0 quadratic, 1 cosine; out: value, derivative
void refsrc_cutoff(int kind, double r, double rc, double *out) {
  std::vector<double> v(2, 0);
  if (kind == 0) quadratic_cutoff(v, r, rc, {});
  else cos_cutoff(v, r, rc, {});
  out[0] = v[0];
  out[1] = v[1];
}
// calculate_radial with chebyshev + the chosen cutoff; out: g, gx, gy, gz (N each)
void refsrc_radial(double x, double y, double z, double r, double rc, int N,
                   double r1, int kind, double *out) {
  std::vector<double> g(N, 0), gx(N, 0), gy(N, 0), gz(N, 0);
  std::function<void(std::vector<double> &, double, double, std::vector<double>)> cf;
  if (kind == 0) cf = quadratic_cutoff;
  else cf = cos_cutoff;
  calculate_radial(g, gx, gy, gz, chebyshev, cf, x, y, z, r, rc, N,
                   std::vector<double>{r1, rc}, {});
  for (int n = 0; n < N; n++) {
    out[n] = g[n];
    out[N + n] = gx[n];
    out[2 * N + n] = gy[n];
    out[3 * N + n] = gz[n];
  }
}
}
