// vb_lut.cpp -- host-built FP64 tables: ETD coefficient records indexed by
// the integer m = kx^2 + ky'^2 + kz'^2, and FFT twiddles.
//
// Built with -ffp-contract=off: in VB_LUT_REFERENCE mode every coefficient is
// the reference's own expression in the reference's operation order, so the
// table reproduces what the reference kernels compute per mode (given the same
// exp).  The ETDRK4 brackets (Calc.cu:558-567) cancel catastrophically for
// small |z| = |h L^|; VB_LUT_ACCURATE evaluates the same phi-functions from
// their Taylor series there instead.
#include <cmath>
#include <vector>

#include "vb_host.hpp"

namespace vb {

int mmax_of(int nx, int ny, int nz) {
  const int a = nx / 2, b = ny / 2, c = nz / 2;
  return a * a + b * b + c * c;
}

// Calc::init_fEIF (Calc.cu:599-600) with c = -(kx^2 + ky^2 + kz^2)(2 pi/L)^2;
// for L = 2 pi the factor is exactly 1 and c is the same double the
// reference forms from its three integer terms.
double lhat(double m, double kfac2, double eps, double kappa, double D) {
  const double c = -(m * kfac2);
  return -eps * c * c + kappa * c / eps + D * (c - kappa / (eps * eps)) / eps;
}

// (k-2)^2/k!, (k-2)/k!, (4-k)/k! series of the ETDRK4 phi-combinations
// f1 = [-4 - z + e^z (4 - 3z + z^2)] / z^3, f2 = [2 + z + e^z (z - 2)] / z^3,
// f3 = [-4 - 3z - z^2 + e^z (4 - z)] / z^3, and phi2 = (e^z - 1 - z) / z^2.
static void phi_series(long double z, long double* f1, long double* f2, long double* f3, long double* p2) {
  long double s1 = 0, s2 = 0, s3 = 0, sp = 0;
  // term_k = z^(k-3)/k!, k >= 3
  long double t = 1.0L / 6.0L;
  for (int k = 3; k < 40; ++k) {
    s1 += (long double)((k - 2) * (k - 2)) * t;
    s2 += (long double)(k - 2) * t;
    s3 += (long double)(4 - k) * t;
    t *= z / (long double)(k + 1);
  }
  long double u = 0.5L;   // z^(k-2)/k!, k >= 2
  for (int k = 2; k < 40; ++k) {
    sp += u;
    u *= z / (long double)(k + 1);
  }
  *f1 = s1; *f2 = s2; *f3 = s3; *p2 = sp;
}

static void phi_closed(long double z, long double* f1, long double* f2, long double* f3, long double* p2) {
  const long double ez = expl(z), z3 = z * z * z;
  *f1 = (-4.0L - z + ez * (4.0L - 3.0L * z + z * z)) / z3;
  *f2 = (2.0L + z + ez * (z - 2.0L)) / z3;
  *f3 = (-4.0L - 3.0L * z - z * z + ez * (4.0L - z)) / z3;
  *p2 = (ez - 1.0L - z) / (z * z);
}

void build_lut(std::vector<Coef>& out, int mmax, int scheme, int mode, double h, double kfac2, double eps,
               double kappa, double D) {
  out.assign((size_t)mmax + 1, Coef{});
  for (int m = 0; m <= mmax; ++m) {
    Coef& r = out[(size_t)m];
    const double H = lhat((double)m, kfac2, eps, kappa, D);
    r.c[7] = H;
    if (mode == VB_LUT_REFERENCE) {
      if (scheme == VB_ETDRK2) {
        // etd1_kernel / etd2_kernel (Calc.cu:453-475)
        const double c = exp(H * h);
        r.c[0] = c;
        r.c[1] = (c - 1.0) / H;
        double c2;
        if (h * H == 0) c2 = h / 2;
        else c2 = (c - 1.0 - h * H) / (h * H * H);
        r.c[2] = c2;
      } else {
        // etd41..etd44 (Calc.cu:495, 556-567)
        const double e1 = exp(H * h / 2);
        const double e2 = e1 * e1;
        const double s = 1.0 / (h * h * H * H * H);
        const double A = -4.0 - h * H + e2 * (4.0 - 3.0 * h * H + h * h * H * H);
        const double B = 2.0 + h * H + e2 * (-2 + h * H);
        const double C = -4 - 3 * h * H - h * h * H * H + e2 * (4 - h * H);
        r.c[0] = e1;
        r.c[1] = (e1 - 1.0) / H;
        r.c[2] = e2;
        r.c[3] = s * A;
        r.c[4] = s * (2 * B);
        r.c[5] = s * C;
      }
    } else {
      const long double z = (long double)h * (long double)H;
      long double f1, f2, f3, p2;
      if (fabsl(z) < 1.0L) phi_series(z, &f1, &f2, &f3, &p2);
      else phi_closed(z, &f1, &f2, &f3, &p2);
      if (scheme == VB_ETDRK2) {
        r.c[0] = (double)expl(z);
        r.c[1] = (double)(expm1l(z) / (long double)H);
        r.c[2] = (double)((long double)h * p2);
      } else {
        r.c[0] = (double)expl(z / 2);
        r.c[1] = (double)(expm1l(z / 2) / (long double)H);
        r.c[2] = (double)expl(z);
        r.c[3] = (double)((long double)h * f1);
        r.c[4] = (double)(2.0L * (long double)h * f2);
        r.c[5] = (double)((long double)h * f3);
      }
    }
  }
}

void build_twiddles(std::vector<double>& out, int n) {
  out.resize(2 * (size_t)n);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int m = 0; m < n; ++m) {
    const long double a = -2.0L * pi * (long double)m / (long double)n;
    out[2 * (size_t)m] = (double)cosl(a);
    out[2 * (size_t)m + 1] = (double)sinl(a);
  }
}

Geo make_geo(int nx, int nz, double L) {
  Geo g;
  g.nx = nx;
  g.ny = nx;
  g.nz = nz;
  g.nh = nx / 2 + 1;
  g.nr = (long long)nx * nx * nz;
  g.nc = (long long)g.nh * nx * nz;
  const double PI = 3.1415926535897932384626433832795;  // tool.h:7
  g.kfac = 2.0 * PI / L;
  g.kfac2 = g.kfac * g.kfac;
  g.L3 = L * L * L;
  g.pitch = g.nh;
  g.aligned = 0;
  g.pp = (long long)g.ny * g.nh;
  g.ncp = g.nc;
  g.nq = -1;
  return g;
}

}  // namespace vb
