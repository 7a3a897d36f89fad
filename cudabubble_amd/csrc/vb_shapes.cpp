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
  if (ninit != 0 && ninit != 1 && ninit != 2 && ninit != 4) return VB_ERR_PARAM;
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
