// CPU driver for the reliable-update decisions of the PCG engine
// (lowlevelfem.jl_amd/csrc/engine_policy.h, the same header the engine's
// update step calls).  It runs the engine's process on the host: CG in FP32
// on K rounded to FP32 with the residual scaled to unit norm, Jacobi
// preconditioner, FP64 dot products; whenever the recursive residual has
// dropped by 0.1 since the last update, the FP32 correction is folded into a
// trial FP64 solution whose true residual decides adoption or rollback
// (adopt_trial), and update_status decides whether to stop.  On a rollback
// the recursion restarts from the kept solution (p = z), as on the device.
//
// Usage: engine_policy_check DIR reltol maxiter [k ...]
//   DIR holds colptr.bin (int64, n+1), rowval.bin (int64), nzval.bin (f64),
//   b.bin (f64, n): the free-DoF block K_ff and its right-hand side.
//   Each k plants an overflow in the FP32 correction before the k-th update
//   (x32[0] = +inf), the failure that the rollback guards against.
// Prints "status S iters I updates U rollbacks R relres X" and writes x.bin.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "engine_policy.h"

template <typename T>
static std::vector<T> load(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) {
    std::fprintf(stderr, "missing %s\n", path.c_str());
    std::exit(2);
  }
  const size_t bytes = (size_t)f.tellg();
  std::vector<T> v(bytes / sizeof(T));
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)bytes);
  return v;
}

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::string d = argv[1];
  const double reltol = std::atof(argv[2]);
  const int64_t maxiter = std::atoll(argv[3]);
  std::set<int64_t> plant;
  for (int a = 4; a < argc; ++a) plant.insert(std::atoll(argv[a]));
  const auto cp = load<int64_t>(d + "/colptr.bin");
  const auto rv = load<int64_t>(d + "/rowval.bin");
  const auto nz = load<double>(d + "/nzval.bin");
  const auto b = load<double>(d + "/b.bin");
  const int64_t n = (int64_t)b.size();
  if ((int64_t)cp.size() != n + 1) return 2;
  std::vector<float> nz32(nz.begin(), nz.end());
  std::vector<float> dinv(n, 0.f);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = cp[j]; k < cp[j + 1]; ++k)
      if (rv[k] == j && nz[k] != 0.0) dinv[j] = (float)(1.0 / nz[k]);
  // K is symmetric: y = K x column by column
  auto spmv64 = [&](const std::vector<double>& x, std::vector<double>& y) {
    std::fill(y.begin(), y.end(), 0.0);
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = cp[j]; k < cp[j + 1]; ++k) y[rv[k]] += nz[k] * x[j];
  };
  auto spmv32 = [&](const std::vector<float>& x, std::vector<float>& y) {
    std::fill(y.begin(), y.end(), 0.f);
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = cp[j]; k < cp[j + 1]; ++k) y[rv[k]] += nz32[k] * x[j];
  };
  auto dot = [&](const auto& u, const auto& v) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (double)u[i] * (double)v[i];
    return s;
  };
  std::vector<double> xh(n, 0.0), xt(n), y(n), rh(n);
  std::vector<float> x(n, 0.f), r(n), z(n), p(n), q(n);
  llfem::UpdateBook bk;
  const double bnorm = std::sqrt(dot(b, b));
  double scale = 1.0 / bnorm, rho = 0.0, maxrr = 0.0, tol2 = 0.0;
  int64_t iters = 0;
  int status = llfem::kRunning;
  // restart of the FP32 recursion from the true residual rh (fresh: p = z)
  auto restart = [&](double rn, bool fresh) {
    const double s_new = 1.0 / rn;
    const float pscale = (float)(s_new / scale);
    scale = s_new;
    for (int64_t i = 0; i < n; ++i) {
      r[i] = (float)(rh[i] * s_new);
      z[i] = dinv[i] * r[i];
      p[i] = fresh ? z[i] : pscale * p[i];
    }
    rho = dot(r, z);
    maxrr = dot(r, r);
    tol2 = (reltol * bnorm * s_new) * (reltol * bnorm * s_new);
  };
  for (int64_t i = 0; i < n; ++i) rh[i] = b[i];
  status = llfem::update_status(bnorm, bnorm, reltol, 0, maxiter, bk);
  if (status == llfem::kRunning) restart(bnorm, true);
  while (status == llfem::kRunning) {
    // FP32 iterations until the recursive residual drops by 0.1 or reaches the tolerance
    for (;;) {
      spmv32(p, q);
      const double pq = dot(p, q);
      const double alpha = pq > 0.0 ? rho / pq : 0.0;
      const float a32 = (float)alpha;
      for (int64_t i = 0; i < n; ++i) {
        x[i] += a32 * p[i];
        r[i] -= a32 * q[i];
        z[i] = dinv[i] * r[i];
      }
      const double rz = dot(r, z), rr = dot(r, r);
      const float beta = (float)(rho != 0.0 ? rz / rho : 0.0);
      rho = rz;
      ++iters;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      if (rr <= tol2 || rr <= 0.01 * maxrr || iters >= maxiter || !(rz > 0.0)) break;
    }
    // FP64 update: trial, adoption or rollback, status
    if (plant.count(bk.updates + 1)) x[0] = INFINITY;
    for (int64_t i = 0; i < n; ++i) {
      xt[i] = xh[i] + (double)x[i] / scale;
      x[i] = 0.f;
    }
    spmv64(xt, y);
    for (int64_t i = 0; i < n; ++i) rh[i] = b[i] - y[i];
    double r2 = dot(rh, rh);
    bool fresh = false;
    if (llfem::adopt_trial(r2, &bk)) {
      xh.swap(xt);
    } else {
      fresh = true;
      spmv64(xh, y);
      for (int64_t i = 0; i < n; ++i) rh[i] = b[i] - y[i];
      r2 = dot(rh, rh);
    }
    ++bk.updates;
    const double rn = std::sqrt(r2);
    status = llfem::update_status(rn, bnorm, reltol, iters, maxiter, bk);
    if (status == llfem::kRunning) restart(rn, fresh);
  }
  spmv64(xh, y);
  double rr = 0.0;
  for (int64_t i = 0; i < n; ++i) rr += (b[i] - y[i]) * (b[i] - y[i]);
  std::printf("status %d iters %lld updates %lld rollbacks %lld relres %.6e\n", status, (long long)iters,
              (long long)bk.updates, (long long)bk.rollbacks, std::sqrt(rr) / bnorm);
  std::ofstream(d + "/x.bin", std::ios::binary)
      .write(reinterpret_cast<const char*>(xh.data()), (std::streamsize)(sizeof(double) * n));
  return 0;
}
