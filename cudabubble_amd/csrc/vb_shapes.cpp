// vb_shapes.cpp -- the reference's deterministic initial shapes selected by
// the ini field ninit, and the raw field files of its results folder.
//
// Grid x_i = i L/n - L/2 (fluidbase.cpp:152-161); every shape is a tanh
// profile of a (pseudo) signed distance, written here from the formulas of
// fluidbase.cpp so that an ini file drives the same initial field.
#include <cmath>
#include <cstdio>
#include <vector>

#include "vb_host.hpp"

namespace {
constexpr double kPI = 3.1415926535897932384626433832795;   // tool.h:7
}

extern "C" {

int vb_shape_ninit(int ninit, int nx, int ny, int nz, double len, double eps, double* phi) {
  if (!phi || nx <= 0 || ny <= 0 || nz <= 0 || !(len > 0) || !(eps > 0)) return VB_ERR_PARAM;
  switch (ninit) {
    case 0: case 1: case 2: case 4: case 5: case 6: case 8: case 11: case 50: case 51: break;
    default: return VB_ERR_PARAM;
  }
  const double L = len * 2.0 * kPI;   // xlen = ylen = zlen (Calc.cu:48-49)
  const double s2 = sqrt(2.0);
  std::vector<double> xx(nx), yy(ny), zz(nz);
  for (int i = 0; i < nx; ++i) xx[i] = i * L / nx - L / 2.0;
  for (int i = 0; i < ny; ++i) yy[i] = i * L / ny - L / 2.0;
  for (int i = 0; i < nz; ++i) zz[i] = i * L / nz - L / 2.0;
  for (int i = 0; i < nz; ++i) {
    const double z = zz[i];
    for (int j = 0; j < ny; ++j) {
      const double y = yy[j];
      for (int k = 0; k < nx; ++k) {
        const double x = xx[k];
        double val;
        if (ninit == 0) {
          // one sphere of radius L/4 (fluidbase.cpp:1080-1099)
          const double r = sqrt(x * x + y * y + z * z);
          val = tanh((0.5 * L / 2.0 - r) / (s2 * eps));
        } else if (ninit == 1) {
          // spheres of radii 0.3 L/2 and 0.2 L/2 on the z axis (fluidbase.cpp:1197-1215)
          const double za = z - 0.4 * L / 2.0, zb = z + 0.5 * L / 2.0;
          const double r1 = sqrt(x * x + y * y + za * za), r2 = sqrt(x * x + y * y + zb * zb);
          val = tanh((0.3 * L / 2.0 - r1) / (s2 * eps)) + tanh((0.2 * L / 2.0 - r2) / (s2 * eps)) + 1.0;
        } else if (ninit == 2) {
          // spheroid with z semi-axis 0.7 of the others, half-width profile (fluidbase.cpp:1862-1877)
          const double r = sqrt((x * x + y * y) + z * z / 0.49);
          val = tanh((0.7 * L / 2.0 - r) / (0.5 * s2 * eps));
        } else if (ninit == 5 || ninit == 6 || ninit == 11) {
          // oblate spheroids: x, y semi-axis a, z semi-axis a / q with q = 2, 5, 3
          // and a = 0.83, 0.8, 0.8 L/2, half-width profile (init_ellipse2,
          // init_ellipse3, init_ellipse4, fluidbase.cpp:1880-1932)
          const double q = ninit == 5 ? 2.0 : ninit == 6 ? 5.0 : 3.0;
          const double a = ninit == 5 ? 0.83 : 0.8;
          const double r = sqrt((x * x + y * y) / (q * q) + z * z) * q;
          val = tanh((a * L / 2.0 - r) / (0.5 * s2 * eps));
        } else if (ninit == 8) {
          // torus about the y axis: centre circle of radius 0.5 L/2 in the x-z
          // plane, tube radius 0.25 L/2 (init_vtorus, fluidbase.cpp:2121-2143)
          const double r = sqrt(x * x + z * z);
          if (r == 0) {
            val = tanh(-0.5 * L / 2.0 / (s2 * eps));
          } else {
            double th = acos(x / r);
            if (z < 0) th = 2 * kPI - th;
            const double cx = x - 0.5 * L / 2.0 * cos(th), cz = z - 0.5 * L / 2.0 * sin(th);
            const double d = 0.25 * L / 2.0 - sqrt(cx * cx + cz * cz + y * y);
            val = tanh(d / (s2 * eps));
          }
        } else if (ninit == 50) {
          // "tian": a 3 x 3 x 3 frame of axis-parallel rods of radius r = L/12
          // through the planes 0 and +-4r, cut at |x|, |y|, |z| = 5r; a sharp
          // +-1 field (init_tian, fluidbase.cpp:1825-1858)
          const double r = L / 12, rr = r * r;
          const double X[3] = {x, x - 4 * r, x + 4 * r}, Y[3] = {y, y - 4 * r, y + 4 * r},
                       Z[3] = {z, z - 4 * r, z + 4 * r};
          bool in = false;
          // the reference lists, for each pair of axes, the rods at all nine offset
          // pairs in {0, 4r, -4r}^2
          for (int a = 0; a < 3 && !in; ++a)
            for (int b = 0; b < 3 && !in; ++b)
              in = X[a] * X[a] + Y[b] * Y[b] < rr || X[a] * X[a] + Z[b] * Z[b] < rr || Y[a] * Y[a] + Z[b] * Z[b] < rr;
          if (fabs(x) > 5 * r || fabs(y) > 5 * r || fabs(z) > 5 * r) in = false;
          val = in ? 1.0 : -1.0;
        } else if (ninit == 51) {
          // the gyroid cos x sin y + cos y sin z + cos z sin x > 0, a sharp +-1
          // field (init_Gyroid, fluidbase.cpp:2255-2270)
          const double g = cos(x) * sin(y) + cos(y) * sin(z) + cos(z) * sin(x);
          val = g > 0 ? 1.0 : -1.0;
        } else {
          // torus whose tube radius varies with the angle (fluidbase.cpp:2068-2090)
          const double r = sqrt(x * x + y * y);
          if (r == 0) {
            val = tanh(-0.5 * L / 2.0 / (s2 * eps));
          } else {
            double th = acos(x / r);
            if (y < 0) th = 2 * kPI - th;
            const double cx = x - 0.5 * L / 2.0 * cos(th), cy = y - 0.5 * L / 2.0 * sin(th);
            const double d = 0.1 * L / 2.0 * (0.75 * cos(th) + 2.25) - sqrt(cx * cx + cy * cy + z * z);
            val = tanh(d / (s2 * eps));
          }
        }
        phi[((size_t)i * ny + j) * nx + k] = val;
      }
    }
  }
  return VB_OK;
}

// Raw doubles [nz][ny][nx] as write_array stores them (tool.cpp:28-42).  A
// file holding (nx+1) x (ny+1) x nz values (the 65x65x64 fields of the
// results folder repeat the first x and y planes at the end) is trimmed to
// nx x ny x nz.
int vb_read_field(const char* path, int nx, int ny, int nz, double* out) {
  if (!path || !out || nx <= 0 || ny <= 0 || nz <= 0) return VB_ERR_PARAM;
  FILE* f = fopen(path, "rb");
  if (!f) return VB_ERR_IO;
  fseek(f, 0, SEEK_END);
  const long bytes = ftell(f);
  fseek(f, 0, SEEK_SET);
  const size_t exact = (size_t)nx * ny * nz, padded = (size_t)(nx + 1) * (ny + 1) * nz;
  int rc = VB_OK;
  if ((size_t)bytes == exact * sizeof(double)) {
    if (fread(out, sizeof(double), exact, f) != exact) rc = VB_ERR_IO;
  } else if ((size_t)bytes == padded * sizeof(double)) {
    std::vector<double> row((size_t)nx + 1);
    for (int i = 0; i < nz && rc == VB_OK; ++i) {
      for (int j = 0; j <= ny; ++j) {
        if (fread(row.data(), sizeof(double), row.size(), f) != row.size()) {
          rc = VB_ERR_IO;
          break;
        }
        if (j < ny)
          for (int k = 0; k < nx; ++k) out[((size_t)i * ny + j) * nx + k] = row[(size_t)k];
      }
    }
  } else {
    rc = VB_ERR_PARAM;
  }
  fclose(f);
  return rc;
}

}  // extern "C"
