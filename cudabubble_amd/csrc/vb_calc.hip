// vb_calc.hip -- the Calc level of the ABI: the reference's individual
// spectral and pointwise operations (class Calc, Calc.h:8-119) on caller
// device memory.  The fused solver does not use these entry points; they
// exist so every reference operation can be called and checked on its own.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>

#include "vb_host.hpp"
#include "vb_kernels.hpp"
#include "vb_ops.hpp"

namespace vb {

constexpr int PW = 256;

static unsigned pgrid(long long n) { return unsigned((n + PW - 1) / PW); }

// laplace_kernel (Calc.cu:107-124)
template <typename T>
__global__ void k_laplace(cpx<T>* u, Geo g) {
  const long long tid = (long long)blockIdx.x * PW + threadIdx.x;
  if (tid >= g.nc) return;
  int z = int(tid / ((long long)g.ny * g.nh));
  int y = int((tid % ((long long)g.ny * g.nh)) / g.nh);
  int x = int((tid % ((long long)g.ny * g.nh)) % g.nh);
  z = fold(z, g.nz);
  y = fold(y, g.ny);
  x = fold(x, g.nx);
  const double c = g.kfac2;
  const double tmp = -(x * x * c + y * y * c + z * z * c);
  u[tid].x = T(double(u[tid].x) * tmp);
  u[tid].y = T(double(u[tid].y) * tmp);
}

// gradient_kernel (Calc.cu:300-341)
__global__ void k_gradient(const cpx<double>* u, cpx<double>* u1, cpx<double>* u2, cpx<double>* u3, Geo g) {
  const long long tid = (long long)blockIdx.x * PW + threadIdx.x;
  if (tid >= g.nc) return;
  const double fac = g.kfac;
  const int x = int(tid % g.nh);
  if (x == g.nx / 2) u1[tid] = mk(0.0, 0.0);
  else u1[tid] = mk(-(u[tid].y * fac) * x, (u[tid].x * fac) * x);
  int y = int((tid % ((long long)g.ny * g.nh)) / g.nh);
  if (y == g.ny / 2) u2[tid] = mk(0.0, 0.0);
  else {
    y = sgnk(y, g.ny);
    u2[tid] = mk(-(u[tid].y * fac) * y, (u[tid].x * fac) * y);
  }
  int z = int(tid / ((long long)g.ny * g.nh));
  if (z == g.nz / 2) u3[tid] = mk(0.0, 0.0);
  else {
    z = sgnk(z, g.nz);
    u3[tid] = mk(-(u[tid].y * fac) * z, (u[tid].x * fac) * z);
  }
}

enum { PW_F, PW_G0, PW_NLD1K, PW_NLD2K, PW_STERM, PW_MULTI, PW_PLUS, PW_NLD1, PW_NLD2 };

struct PwArgs {
  double* out;
  const double* a;
  const double* b;
  const double* c;
  const double* d;
  double s0, s1, s2, s3;
  long long n;
};

// pointwise kernels of Calc.cu:153-437
template <int OP>
__global__ void k_pw(PwArgs p) {
  const long long i = (long long)blockIdx.x * PW + threadIdx.x;
  if (i >= p.n) return;
  if constexpr (OP == PW_F) {            // f_kernel: a = lap, b = d, s0 = eps
    p.out[i] = -p.a[i] * p.s0 + (p.b[i] * p.b[i] - 1.0) * p.b[i] / p.s0;
  } else if constexpr (OP == PW_G0) {    // g0_kernel: out = g0 (in/out), a = d, b = f
    p.out[i] = -p.out[i] + (3.0 * p.a[i] * p.a[i] - 1.0) * p.b[i] / (p.s0 * p.s0);
  } else if constexpr (OP == PW_NLD1K) { // nld_kernel1: a = d, s0 = kappa
    p.out[i] = p.a[i] * (p.a[i] * p.a[i] - 1.0 - p.s0);
  } else if constexpr (OP == PW_NLD2K) { // nld_kernel2: out = nld (in/out), a = tmp1, b = d, c = f
    const double e = 1 / p.s0, e2 = e * e, e3 = e * e2, D = p.s1;
    p.out[i] = -p.out[i] * e + D * p.a[i] * e3 + (3 * p.b[i] * p.b[i] - D - 1) * e2 * p.c[i];
  } else if constexpr (OP == PW_STERM) { // s_term_kernel: a = d, b,c,d = dx,dy,dz
    p.out[i] = 0.5 * p.s0 * (p.b[i] * p.b[i] + p.c[i] * p.c[i] + p.d[i] * p.d[i]) +
               0.25 / p.s0 * (p.a[i] * p.a[i] - 1) * (p.a[i] * p.a[i] - 1);
  } else if constexpr (OP == PW_MULTI) { // multi_kernel
    p.out[i] = p.a[i] * p.b[i];
  } else if constexpr (OP == PW_PLUS) {  // plus_kernel: out += s0 * a
    p.out[i] += p.s0 * p.a[i];
  } else if constexpr (OP == PW_NLD1) {  // nld1_kernel: a = f, b = g
    p.out[i] = p.s0 * 0.5 + p.s1 * p.a[i] + p.b[i] * p.s2;
  } else {                               // nld2_kernel
    p.out[i] = -p.s0 * 0.5 - p.s1 * p.a[i] * 3.0 / (2.0 * sqrt(2.0)) - p.b[i] * p.s2;
  }
}

// Block partial sums of intf (intf_kernel, Calc.cu:224-252).  Workgroup b
// adds the elements [b kSumChunk, (b + 1) kSumChunk) in a fixed order, so the
// result does not depend on the number of CUs (the reference's grid is twice
// the SM count, Calc.cu:259).
constexpr long long kSumChunk = 256 * 256;
inline long long sum_parts(long long n) { return (n + kSumChunk - 1) / kSumChunk; }

__global__ void __launch_bounds__(256) k_sum(const double* __restrict__ u, long long n, double* __restrict__ part) {
  __shared__ double red[4];
  double acc[1] = {0.0};
  const long long b0 = (long long)blockIdx.x * kSumChunk;
  const long long e = b0 + kSumChunk < n ? b0 + kSumChunk : n;
  for (long long i = b0 + threadIdx.x; i < e; i += 256) acc[0] += u[i];
  block_sum_store<1, 256>(acc, red, part + blockIdx.x);
}

// (sum of the nb partials / N + addition) * mL3 into out[0], mL3 = multiply L^3
// (sumArray_kernel, Calc.cu:191-221, 255-263)
__global__ void __launch_bounds__(256) k_intf_final(const double* __restrict__ part, int nb, double N,
                                                    double addition, double mL3, double* __restrict__ out) {
  __shared__ double red[4];
  __shared__ double tot[1];
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < nb; i += 256) acc[0] += part[i];
  block_sum_store<1, 256>(acc, red, tot);
  if (threadIdx.x == 0) out[0] = (tot[0] / N + addition) * mL3;
}

// the two launches of intf: the partials of n points into part (sum_parts(n)
// doubles), the value into val[0]
static void launch_intf(const double* u, long long n, double L, double addition, double multiply, double* part,
                        double* val, hipStream_t st) {
  const long long nb = sum_parts(n);
  hipLaunchKernelGGL(k_sum, dim3(unsigned(nb)), dim3(256), 0, st, u, n, part);
  hipLaunchKernelGGL(k_intf_final, dim3(1), dim3(256), 0, st, part, int(nb), double(n), addition,
                     multiply * L * L * L, val);
}

enum { ETD_1, ETD_2, ETD_41, ETD_42, ETD_43, ETD_44 };

struct EtdArgs {
  cpx<double>* out;
  const cpx<double>* fd;
  const cpx<double>* n1;
  const cpx<double>* n2;
  const cpx<double>* n3;
  const Coef* lut;
  Geo g;
};

// ETD combines (Calc.cu:446-569) with the per-mode coefficients of the table
template <int OP>
__global__ void k_etd(EtdArgs a) {
  const long long tid = (long long)blockIdx.x * PW + threadIdx.x;
  if (tid >= a.g.nc) return;
  const Geo& g = a.g;
  const int z = int(tid / ((long long)g.ny * g.nh));
  const int y = int((tid % ((long long)g.ny * g.nh)) / g.nh);
  const int x = int(tid % g.nh);
  const Coef& cf = a.lut[mindex(g, x, y, z)];
  const double H = cf.c[7];
  if constexpr (OP == ETD_1) {
    const double c = cf.c[0];
    a.out[tid] = mk(c * a.fd[tid].x + (c - 1.0) * a.n1[tid].x / H, c * a.fd[tid].y + (c - 1.0) * a.n1[tid].y / H);
  } else if constexpr (OP == ETD_2) {   // n1 = nnld, n2 = nld
    const double c = cf.c[2];
    a.out[tid] = mk(a.fd[tid].x + c * (a.n1[tid].x - a.n2[tid].x), a.fd[tid].y + c * (a.n1[tid].y - a.n2[tid].y));
  } else if constexpr (OP == ETD_41 || OP == ETD_42) {
    const double e1 = cf.c[0];
    a.out[tid] = mk(e1 * a.fd[tid].x + (e1 - 1.0) * a.n1[tid].x / H, e1 * a.fd[tid].y + (e1 - 1.0) * a.n1[tid].y / H);
  } else if constexpr (OP == ETD_43) {  // out = c (in/out), n1 = nld, n2 = nldb
    const double e1 = cf.c[0];
    const cpx<double> c = a.out[tid];
    a.out[tid] = mk(c.x * e1 + (e1 - 1.0) * (2 * a.n2[tid].x - a.n1[tid].x) / H,
                    c.y * e1 + (e1 - 1.0) * (2 * a.n2[tid].y - a.n1[tid].y) / H);
  } else {                              // ETD_44: n1 = nlda, n2 = nldb, n3 = nldc
    const double e2 = cf.c[2], fa = cf.c[3], fb = cf.c[4], fc = cf.c[5];
    const cpx<double> na = a.n1[tid], nb = a.n2[tid], nc = a.n3[tid];
    a.out[tid] = mk(e2 * a.fd[tid].x + (na.x * fa + (na.x + nb.x) * fb + nc.x * fc),
                    e2 * a.fd[tid].y + (na.y * fa + (na.y + nb.y) * fb + nc.y * fc));
  }
}

struct CalcImpl {
  int nx = 0, nz = 0, prec = VB_FP64;
  double L = 0;
  Geo g{};
  hipStream_t st = nullptr;
  const SizeOps<double>* oxd = nullptr;
  const SizeOps<double>* ozd = nullptr;
  const SizeOps<float>* oxf = nullptr;
  const SizeOps<float>* ozf = nullptr;
  DBuf<cpx<double>> c1, c2, c3, c4;
  DBuf<cpx<float>> fc1;
  DBuf<cpx<double>> twx, twz;
  DBuf<cpx<float>> twxf, twzf;
  DBuf<double> part, val, rtmp;
  std::map<std::tuple<int, double, double>, std::unique_ptr<DBuf<Coef>>> luts;
  long long npairs = 0;

  CalcImpl(int nx_, int nz_, double lenx, int precision) {
    nx = nx_;
    nz = nz_;
    prec = precision;
    L = lenx;
    g = make_geo(nx, nz, L);
    VB_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    oxd = get_ops<double>(nx);
    ozd = get_ops<double>(nz);
    oxf = get_ops<float>(nx);
    ozf = get_ops<float>(nz);
    const size_t nc = (size_t)g.nc;
    c1.alloc(nc);
    c2.alloc(nc);
    c3.alloc(nc);
    c4.alloc(nc);
    if (prec == VB_FP32_MIXED) fc1.alloc(nc);
    part.alloc((size_t)sum_parts(g.nr));
    val.alloc(1);
    rtmp.alloc((size_t)g.nr);
    npairs = (long long)nz * nx / 2;
    std::vector<double> t;
    build_twiddles(t, nx);
    upload(twx, twxf, t, nx);
    build_twiddles(t, nz);
    upload(twz, twzf, t, nz);
  }
  ~CalcImpl() {
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
  static void upload(DBuf<cpx<double>>& d, DBuf<cpx<float>>& f, const std::vector<double>& t, int n) {
    std::vector<cpx<double>> hd((size_t)n);
    std::vector<cpx<float>> hf((size_t)n);
    for (int m = 0; m < n; ++m) {
      hd[(size_t)m].x = t[2 * (size_t)m];
      hd[(size_t)m].y = t[2 * (size_t)m + 1];
      hf[(size_t)m].x = float(t[2 * (size_t)m]);
      hf[(size_t)m].y = float(t[2 * (size_t)m + 1]);
    }
    d.alloc((size_t)n);
    f.alloc((size_t)n);
    VB_HIP(hipMemcpy(d.p, hd.data(), sizeof(cpx<double>) * n, hipMemcpyHostToDevice));
    VB_HIP(hipMemcpy(f.p, hf.data(), sizeof(cpx<float>) * n, hipMemcpyHostToDevice));
  }

  void done() {
    VB_HIP(hipGetLastError());
    VB_HIP(hipStreamSynchronize(st));
  }

  // forward normalised R2C (r2c_3d_f + normalize_kernel, Calc.cu:88-96)
  void r2c(const void* u, void* s) {
    if (prec == VB_FP32_MIXED) {
      cpx<float>* o = (cpx<float>*)s;
      oxf->xr2c((const float*)u, o, npairs, twxf.p, 1.0 / double(g.nr), g, st);
      oxf->ypass(o, nullptr, 1, -1, g, twxf.p, nullptr, st);
      ozf->zpass(o, -1, g, twzf.p, st);
    } else {
      cpx<double>* o = (cpx<double>*)s;
      oxd->xr2c_d((const double*)u, o, npairs, twx.p, 1.0 / double(g.nr), g, st);
      oxd->ypass(o, nullptr, 1, -1, g, twx.p, nullptr, st);
      ozd->zpass(o, -1, g, twz.p, st);
    }
  }
  // backward C2R of a buffer that may be overwritten (c2r_3d_b, Calc.cu:99-104)
  void c2r_destroy(cpx<double>* s, double* u) {
    ozd->zpass(s, +1, g, twz.p, st);
    oxd->ypass(s, nullptr, 1, +1, g, twx.p, nullptr, st);
    oxd->xc2r_d(s, u, npairs, twx.p, g, st);
  }
  void c2r(const void* s, void* u) {
    if (prec == VB_FP32_MIXED) {
      VB_HIP(hipMemcpyAsync(fc1.p, s, sizeof(cpx<float>) * (size_t)g.nc, hipMemcpyDeviceToDevice, st));
      ozf->zpass(fc1.p, +1, g, twzf.p, st);
      oxf->ypass(fc1.p, nullptr, 1, +1, g, twxf.p, nullptr, st);
      oxf->xc2r(fc1.p, (float*)u, npairs, twxf.p, g, st);
    } else {
      VB_HIP(hipMemcpyAsync(c4.p, s, sizeof(cpx<double>) * (size_t)g.nc, hipMemcpyDeviceToDevice, st));
      c2r_destroy(c4.p, (double*)u);
    }
  }
  void laplace(void* s) {
    if (prec == VB_FP32_MIXED)
      hipLaunchKernelGGL(k_laplace<float>, dim3(pgrid(g.nc)), dim3(PW), 0, st, (cpx<float>*)s, g);
    else
      hipLaunchKernelGGL(k_laplace<double>, dim3(pgrid(g.nc)), dim3(PW), 0, st, (cpx<double>*)s, g);
  }
  void lap_spec(const void* fd, void* out) {   // Calc.cu:139-144
    VB_HIP(hipMemcpyAsync(c1.p, fd, sizeof(cpx<double>) * (size_t)g.nc, hipMemcpyDeviceToDevice, st));
    laplace(c1.p);
    c2r_destroy(c1.p, (double*)out);
  }
  void lap_real(const void* d, void* out) {    // Calc.cu:146-151
    r2c(d, c1.p);
    laplace(c1.p);
    c2r_destroy(c1.p, (double*)out);
  }
  void fftgradient(const void* fd, void* dx, void* dy, void* dz) {   // Calc.cu:344-356
    hipLaunchKernelGGL(k_gradient, dim3(pgrid(g.nc)), dim3(PW), 0, st, (const cpx<double>*)fd, c1.p, c2.p, c3.p, g);
    c2r_destroy(c1.p, (double*)dx);
    c2r_destroy(c2.p, (double*)dy);
    c2r_destroy(c3.p, (double*)dz);
  }
  template <int OP>
  void pw(double* out, const void* a, const void* b, const void* c, const void* d, double s0, double s1 = 0,
          double s2 = 0) {
    PwArgs p{out, (const double*)a, (const double*)b, (const double*)c, (const double*)d, s0, s1, s2, 0, g.nr};
    hipLaunchKernelGGL(k_pw<OP>, dim3(pgrid(g.nr)), dim3(PW), 0, st, p);
  }
  // Calc::intf (Calc.cu:255-263): (sum(u)/N + addition) * multiply * L^3
  double intf(const void* u, double addition, double multiply) {
    launch_intf((const double*)u, g.nr, L, addition, multiply, part.p, val.p, st);
    double s = 0;
    VB_HIP(hipMemcpyAsync(&s, val.p, sizeof(double), hipMemcpyDeviceToHost, st));
    done();
    return s;
  }
  const Coef* lut(int sch, double eps, double h) {
    auto key = std::make_tuple(sch, eps, h);
    auto it = luts.find(key);
    if (it != luts.end()) return it->second->p;
    std::vector<Coef> tab;
    build_lut(tab, mmax_of(nx, nx, nz), sch, VB_LUT_REFERENCE, h, g.kfac2, eps, 2.0, 1.5);
    auto b = std::make_unique<DBuf<Coef>>();
    b->alloc(tab.size());
    VB_HIP(hipMemcpy(b->p, tab.data(), sizeof(Coef) * tab.size(), hipMemcpyHostToDevice));
    const Coef* r = b->p;
    luts.emplace(key, std::move(b));
    return r;
  }
  template <int OP>
  void etd(void* out, const void* fd, const void* n1, const void* n2, const void* n3, double eps, double h) {
    EtdArgs a{(cpx<double>*)out, (const cpx<double>*)fd, (const cpx<double>*)n1, (const cpx<double>*)n2,
              (const cpx<double>*)n3, lut(OP == ETD_1 || OP == ETD_2 ? 2 : 4, eps, h), g};
    hipLaunchKernelGGL(k_etd<OP>, dim3(pgrid(g.nc)), dim3(PW), 0, st, a);
  }
  // Calc::init_fEIF (Calc.cu:579-606): host table, then H2D
  void init_fEIF(void* Het, double eps) {
    const int nh = nx / 2 + 1, ny = nx;
    std::vector<double> t((size_t)g.nc);
    for (int i = 0; i < nz; i++) {
      const int i2 = i < nz - i ? i : nz - i;
      for (int j = 0; j < ny; j++) {
        const int j2 = j < ny - j ? j : ny - j;
        for (int k = 0; k < nh; k++) {
          const int k2 = k < nx - k ? k : nx - k;
          t[((size_t)i * ny + j) * nh + k] = lhat((double)(i2 * i2 + j2 * j2 + k2 * k2), g.kfac2, eps, 2.0, 1.5);
        }
      }
    }
    VB_HIP(hipMemcpyAsync(Het, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, st));
    done();
  }
};

}  // namespace vb

using vb::CalcImpl;

struct vb_calc {
  std::unique_ptr<CalcImpl> impl;
};

namespace {
template <class F>
int guard(F&& f) {
  try {
    f();
    return VB_OK;
  } catch (const vb::HipError& e) {
    vb::set_last_error(std::string(hipGetErrorString(e.err)) + " in " + e.where);
    return e.err == hipErrorOutOfMemory ? VB_ERR_OOM : VB_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return VB_ERR_OOM;
  }
}
}  // namespace

extern "C" {

int vb_calc_create(int nx, int nz, double lenx, int precision, vb_calc** out) {
  if (!out) return VB_ERR_PARAM;
  *out = nullptr;
  if (nx <= 0 || nz <= 0) return VB_ERR_PARAM;
  if (nx % 2) return VB_ERR_ODD_NX;
  if (!vb::size_supported(nx) || !vb::size_supported(nz) || lenx <= 0) return VB_ERR_PARAM;
  if (precision != VB_FP64 && precision != VB_FP32_MIXED) return VB_ERR_PARAM;
  int ndev = 0;
  const hipError_t de = hipGetDeviceCount(&ndev);
  if (de != hipSuccess || ndev <= 0) {
    vb::set_last_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(de) + ", devices " +
                       std::to_string(ndev));
    return VB_ERR_NO_DEVICE;
  }
  auto c = std::make_unique<vb_calc>();
  int rc = guard([&] { c->impl = std::make_unique<CalcImpl>(nx, nz, lenx, precision); });
  if (rc == VB_OK) *out = c.release();
  return rc;
}

void vb_calc_destroy(vb_calc* c) { delete c; }

#define CALC_BODY(expr)                      \
  if (!c || !c->impl) return VB_ERR_PARAM;   \
  CalcImpl& k = *c->impl;                    \
  return guard([&] { expr; k.done(); })

int vb_calc_r2c(vb_calc* c, const void* u, void* s) { CALC_BODY(k.r2c(u, s)); }
int vb_calc_c2r(vb_calc* c, const void* s, void* u) { CALC_BODY(k.c2r(s, u)); }
int vb_calc_laplace(vb_calc* c, void* s) { CALC_BODY(k.laplace(s)); }
int vb_calc_lap_spec(vb_calc* c, const void* fd, void* out) { CALC_BODY(k.lap_spec(fd, out)); }
int vb_calc_lap_real(vb_calc* c, const void* d, void* out) { CALC_BODY(k.lap_real(d, out)); }
int vb_calc_fftgradient(vb_calc* c, const void* fd, void* dx, void* dy, void* dz) {
  CALC_BODY(k.fftgradient(fd, dx, dy, dz));
}
int vb_calc_f(vb_calc* c, void* f, const void* lap, const void* d, double eps) {
  CALC_BODY(k.pw<vb::PW_F>((double*)f, lap, d, nullptr, nullptr, eps));
}
int vb_calc_g0(vb_calc* c, void* g0, const void* d, const void* f, double eps) {
  CALC_BODY(k.pw<vb::PW_G0>((double*)g0, d, f, nullptr, nullptr, eps));
}
int vb_calc_intf(vb_calc* c, const void* u, double addition, double multiply, double* host_val) {
  if (!host_val) return VB_ERR_PARAM;
  CALC_BODY(*host_val = k.intf(u, addition, multiply));
}
int vb_calc_nld(vb_calc* c, void* nldp, const void* d, const void* f, void* tmp1, double eps, double kappa,
                double D) {
  CALC_BODY({
    k.pw<vb::PW_NLD1K>((double*)tmp1, d, nullptr, nullptr, nullptr, kappa);
    k.lap_real(tmp1, nldp);
    k.pw<vb::PW_NLD2K>((double*)nldp, tmp1, d, f, nullptr, eps, D);
  });
}
int vb_calc_s_term(vb_calc* c, void* s, const void* d, const void* dx, const void* dy, const void* dz, double eps) {
  CALC_BODY(k.pw<vb::PW_STERM>((double*)s, d, dx, dy, dz, eps));
}
int vb_calc_multi(vb_calc* c, const void* a, const void* b, void* out) {
  CALC_BODY(k.pw<vb::PW_MULTI>((double*)out, a, b, nullptr, nullptr, 0.0));
}
int vb_calc_plus(vb_calc* c, void* u, const void* nld1, double t) {
  CALC_BODY(k.pw<vb::PW_PLUS>((double*)u, nld1, nullptr, nullptr, nullptr, t));
}
int vb_calc_nld1(vb_calc* c, void* nld, const void* f, const void* g, double l1, double l2, double eta) {
  CALC_BODY(k.pw<vb::PW_NLD1>((double*)nld, f, g, nullptr, nullptr, l1, l2, eta));
}
int vb_calc_nld2(vb_calc* c, void* nld, const void* f, const void* g, double l1, double l2, double eta) {
  CALC_BODY(k.pw<vb::PW_NLD2>((double*)nld, f, g, nullptr, nullptr, l1, l2, eta));
}
int vb_calc_init_fEIF(vb_calc* c, void* Het, double eps) { CALC_BODY(k.init_fEIF(Het, eps)); }
int vb_calc_etd1(vb_calc* c, void* a, const void* fd, const void* nld, double eps, double h) {
  CALC_BODY(k.etd<vb::ETD_1>(a, fd, nld, nullptr, nullptr, eps, h));
}
int vb_calc_etd2(vb_calc* c, void* a, const void* fd, const void* nnld, const void* nld, double eps, double h) {
  CALC_BODY(k.etd<vb::ETD_2>(a, fd, nnld, nld, nullptr, eps, h));
}
int vb_calc_etd41(vb_calc* c, void* d_real, const void* fd, const void* nld, double eps, double h) {
  CALC_BODY({
    k.etd<vb::ETD_41>(k.c1.p, fd, nld, nullptr, nullptr, eps, h);
    k.c2r_destroy(k.c1.p, (double*)d_real);
  });
}
int vb_calc_etd42(vb_calc* c, void* b, const void* fd, const void* nlda, double eps, double h) {
  CALC_BODY(k.etd<vb::ETD_42>(b, fd, nlda, nullptr, nullptr, eps, h));
}
int vb_calc_etd43(vb_calc* c, void* cc, const void* nld, const void* nldb, double eps, double h) {
  CALC_BODY(k.etd<vb::ETD_43>(cc, nullptr, nld, nldb, nullptr, eps, h));
}
int vb_calc_etd44(vb_calc* c, void* d_real, const void* fd, const void* nlda, const void* nldb, const void* nldc,
                  double eps, double h) {
  CALC_BODY({
    k.etd<vb::ETD_44>(k.c1.p, fd, nlda, nldb, nldc, eps, h);
    k.c2r_destroy(k.c1.p, (double*)d_real);
  });
}
int vb_calc_sync(vb_calc* c) { CALC_BODY((void)0); }

}  // extern "C"

// ---------------------------------------------------------------------------
// The allocation-free integral (vb_intf_static).  Its block partial sums live
// in `scratch`, a file-scope device array: every GPU the library runs on holds
// one copy (128 KiB) until the process ends.  Host code never passes the
// array's name to a kernel (in host code it is a stand-in, not a device
// address): the address of each device's copy is looked up once per device,
// with that device current, and kept.
// ---------------------------------------------------------------------------
static constexpr long long kStaticParts = 16384;   // partials: inputs of up to 16384 * 65536 = 2^30 points
static __device__ __align__(256) double scratch[kStaticParts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch` (the calling thread's current device)
int scratch_of_current(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return VB_ERR_NO_DEVICE;
  if (dev >= kMaxDevices) return VB_ERR_PARAM;
  std::lock_guard<std::mutex> l(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      vb::set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return VB_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return VB_OK;
}
}  // namespace

extern "C" int vb_intf_static(const double* u, long long n, double lenx, double addition, double multiply,
                              double* val, void* stream) {
  if (!u || !val || n <= 0 || !(lenx > 0)) return VB_ERR_PARAM;
  if (vb::sum_parts(n) > kStaticParts) {
    vb::set_last_error("vb_intf_static: " + std::to_string(n) + " points exceed its maximum of " +
                       std::to_string(kStaticParts * vb::kSumChunk));
    return VB_ERR_PARAM;
  }
  double* part = nullptr;
  const int rc = scratch_of_current(&part);
  if (rc != VB_OK) return rc;
  vb::launch_intf(u, n, lenx, addition, multiply, part, val, static_cast<hipStream_t>(stream));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    vb::set_last_error(std::string("vb_intf_static launch: ") + hipGetErrorString(e));
    return VB_ERR_HIP;
  }
  return VB_OK;
}
