// Dirichlet-eliminated Jacobi-PCG on the device, FP64 or FP32-inner with
// FP64 iterative refinement.
//
// Reference: solveDisplacement(K, f; support, iterative=true, reltol, maxiter,
// preconditioner) (src/linear.jl:3711-3743): u_fixed on the constrained DoFs,
// rhs = f_free - K[free,fixed] u_fixed, u_free = cg(K_ff, rhs; Pl, reltol,
// maxiter) with IterativeSolvers.cg semantics (x0 = 0, converged when
// ||r|| <= reltol ||r0||).  K_ff is applied as the masked operator
// P_f K P_f (no submatrix is extracted); fixed entries stay 0 in every
// Krylov vector.
//
// SpMV runs on the node-panel layout of the assembled K (block CSR of K^T,
// equal to K by symmetry): one int32 column index per 3x3 block, so the
// index stream is 1/9 of scalar CSR.  Dot products reduce per block and then
// in one fixed-order pass, so every run gives bit-identical iterates.
// Scalars (alpha, beta, rho, ||r||, convergence) live on the device; the host
// only polls the convergence flag every few iterations.
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.h"

namespace llfem {
namespace {

constexpr int kBS = 256;
constexpr int kG = 8;   // lanes per node row in the SpMV

enum { S_RHO = 0, S_ALPHA, S_BETA, S_RR, S_PQ, S_TOL2, S_DONE, S_ITER, S_NS };

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// y = mask(K x); optional partial of dot(x, y).  Node row j uses kG lanes.
template <typename MT, typename VT>
__global__ __launch_bounds__(kBS) void k_spmv(int64_t N, const int64_t* __restrict__ adj_ptr,
                                             const int32_t* __restrict__ adj, const MT* __restrict__ val,
                                             const VT* __restrict__ x, const uint8_t* __restrict__ fixed,
                                             VT* __restrict__ y, double* __restrict__ part,
                                             const double* __restrict__ scal) {
  __shared__ double sh[kBS];
  if (scal && scal[S_DONE] != 0.0) return;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = gt / kG;
  const int l = (int)(gt % kG);
  VT a0 = 0, a1 = 0, a2 = 0;
  if (j < N) {
    const int64_t p0 = adj_ptr[j];
    const int R = (int)(adj_ptr[j + 1] - p0);
    const MT* P = val + 9 * p0;
    for (int s = l; s < R; s += kG) {
      const int64_t i = adj[p0 + s];
      const VT x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
      a0 += (VT)P[3 * s] * x0 + (VT)P[3 * s + 1] * x1 + (VT)P[3 * s + 2] * x2;
      a1 += (VT)P[3 * R + 3 * s] * x0 + (VT)P[3 * R + 3 * s + 1] * x1 + (VT)P[3 * R + 3 * s + 2] * x2;
      a2 += (VT)P[6 * R + 3 * s] * x0 + (VT)P[6 * R + 3 * s + 1] * x1 + (VT)P[6 * R + 3 * s + 2] * x2;
    }
  }
#pragma unroll
  for (int o = kG / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kG);
    a1 += __shfl_xor(a1, o, kG);
    a2 += __shfl_xor(a2, o, kG);
  }
  double d = 0.0;
  if (j < N && l == 0) {
    const VT r[3] = {a0, a1, a2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const VT v = (fixed && fixed[3 * j + c]) ? (VT)0 : r[c];
      y[3 * j + c] = v;
      d += (double)x[3 * j + c] * (double)v;
    }
  }
  if (part) {
    const double s = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

// Fixed-order final reduction of `n` partial rows of width w into out[0..w).
__global__ void k_reduce(int64_t n, int w, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[kBS];
  for (int c = 0; c < w; ++c) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += part[i * w + c];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = s;
  }
}

// alpha = rho / (p.q)
__global__ void k_alpha(double* scal, const double* red) {
  if (scal[S_DONE] != 0.0) return;
  scal[S_PQ] = red[0];
  scal[S_ALPHA] = scal[S_RHO] / red[0];
}

// x += alpha p; r -= alpha q; z = dinv r; partials (r.r, r.z)
template <typename VT>
__global__ __launch_bounds__(kBS) void k_update(int64_t n, const double* __restrict__ scal,
                                               VT* __restrict__ x, VT* __restrict__ r,
                                               const VT* __restrict__ p, const VT* __restrict__ q,
                                               const VT* __restrict__ dinv, VT* __restrict__ z,
                                               double* __restrict__ part) {
  __shared__ double sh[kBS];
  if (scal[S_DONE] != 0.0) return;
  const VT alpha = (VT)scal[S_ALPHA];
  double rr = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const VT ri = r[i] - alpha * q[i];
    r[i] = ri;
    const VT zi = dinv[i] * ri;
    z[i] = zi;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  rr = block_sum(rr, sh);
  rz = block_sum(rz, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rr;
    part[2 * blockIdx.x + 1] = rz;
  }
}

// beta = rz / rho; rho = rz; convergence test on ||r||
__global__ void k_beta(double* scal, const double* red, int64_t maxiter) {
  if (scal[S_DONE] != 0.0) return;
  scal[S_RR] = red[0];
  scal[S_BETA] = red[1] / scal[S_RHO];
  scal[S_RHO] = red[1];
  scal[S_ITER] += 1.0;
  if (red[0] <= scal[S_TOL2] || scal[S_ITER] >= (double)maxiter) scal[S_DONE] = 1.0;
}

template <typename VT>
__global__ void k_pupd(int64_t n, const double* __restrict__ scal, const VT* __restrict__ z,
                       VT* __restrict__ p) {
  if (scal[S_DONE] != 0.0) return;
  const VT beta = (VT)scal[S_BETA];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// init: x = 0; z = dinv r; p = z; partials (r.r, r.z)
template <typename VT>
__global__ __launch_bounds__(kBS) void k_init(int64_t n, VT* __restrict__ x, const VT* __restrict__ r,
                                             const VT* __restrict__ dinv, VT* __restrict__ z,
                                             VT* __restrict__ p, double* __restrict__ part) {
  __shared__ double sh[kBS];
  double rr = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = 0;
    const VT zi = dinv[i] * r[i];
    z[i] = zi;
    p[i] = zi;
    rr += (double)r[i] * (double)r[i];
    rz += (double)r[i] * (double)zi;
  }
  rr = block_sum(rr, sh);
  rz = block_sum(rz, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rr;
    part[2 * blockIdx.x + 1] = rz;
  }
}

__global__ void k_init_scal(double* scal, const double* red, double rtol) {
  scal[S_RR] = red[0];
  scal[S_RHO] = red[1];
  scal[S_TOL2] = rtol * rtol * red[0];
  scal[S_ITER] = 0.0;
  scal[S_DONE] = (red[0] <= scal[S_TOL2] || red[0] == 0.0) ? 1.0 : 0.0;
}

// diag / Jacobi inverse from the panels; 0 on fixed DoFs
template <typename VT>
__global__ void k_dinv(int64_t N, const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ dslot,
                       const double* __restrict__ val, const uint8_t* __restrict__ fixed,
                       VT* __restrict__ dinv) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int64_t p0 = adj_ptr[j];
  const int R = (int)(adj_ptr[j + 1] - p0);
  const int s = dslot[j];
  for (int c = 0; c < 3; ++c) {
    const double dv = s >= 0 ? val[9 * p0 + (int64_t)c * 3 * R + 3 * s + c] : 0.0;
    dinv[3 * j + c] = (fixed[3 * j + c] || dv == 0.0) ? (VT)0 : (VT)(1.0 / dv);
  }
}

// r = mask(f - y); for the FP64 residual of the refinement loop
__global__ void k_resid(int64_t n, const double* __restrict__ f, const double* __restrict__ y,
                        const uint8_t* __restrict__ fixed, double* __restrict__ r,
                        double* __restrict__ part) {
  __shared__ double sh[kBS];
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = fixed[i] ? 0.0 : f[i] - y[i];
    r[i] = v;
    rr += v * v;
  }
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, double scale, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)(a[i] * scale);
}

__global__ void k_axpy_f32(int64_t n, double s, const float* __restrict__ d, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] += s * (double)d[i];
}

__global__ void k_masked_set(int64_t n, const uint8_t* __restrict__ fixed, const double* __restrict__ ufix,
                             double* __restrict__ u) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    u[i] = fixed[i] ? (ufix ? ufix[i] : 0.0) : 0.0;
}

__global__ void k_cast_val(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_axpy(int64_t n, double s, const double* __restrict__ d, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] += s * d[i];
}

// Scalar CSC operator (caller-supplied symmetric matrix, used as CSR of K^T).
template <typename MT, typename VT>
__global__ __launch_bounds__(kBS) void k_spmv_csc(int64_t n, const int64_t* __restrict__ colptr,
                                                 const int32_t* __restrict__ rowval,
                                                 const MT* __restrict__ val, const VT* __restrict__ x,
                                                 const uint8_t* __restrict__ fixed, VT* __restrict__ y,
                                                 double* __restrict__ part,
                                                 const double* __restrict__ scal) {
  __shared__ double sh[kBS];
  if (scal && scal[S_DONE] != 0.0) return;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = gt / kG;
  const int l = (int)(gt % kG);
  VT acc = 0;
  if (j < n)
    for (int64_t p = colptr[j] + l; p < colptr[j + 1]; p += kG) acc += (VT)val[p] * x[rowval[p]];
#pragma unroll
  for (int o = kG / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kG);
  double d = 0.0;
  if (j < n && l == 0) {
    const VT v = (fixed && fixed[j]) ? (VT)0 : acc;
    y[j] = v;
    d = (double)x[j] * (double)v;
  }
  if (part) {
    const double s = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

template <typename VT>
__global__ void k_dinv_csc(int64_t n, const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                           const double* __restrict__ val, const uint8_t* __restrict__ fixed,
                           VT* __restrict__ dinv) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double dv = 0.0;
  for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p)
    if (rowval[p] == j) dv = val[p];
  dinv[j] = (fixed[j] || dv == 0.0) ? (VT)0 : (VT)(1.0 / dv);
}

// Operator views.
struct PanelOp {
  int64_t N;
  const int64_t* adj_ptr;
  const int32_t* adj;
  const int32_t* dslot;
  int64_t ndof() const { return 3 * N; }
  unsigned grid() const { return (unsigned)((N * kG + kBS - 1) / kBS); }
  template <typename MT, typename VT>
  void spmv(hipStream_t st, const MT* val, const VT* x, const uint8_t* fixed, VT* y, double* part,
            const double* scal) const {
    k_spmv<MT, VT><<<grid(), kBS, 0, st>>>(N, adj_ptr, adj, val, x, fixed, y, part, scal);
  }
  template <typename VT>
  void dinv(hipStream_t st, const double* val, const uint8_t* fixed, VT* d) const {
    k_dinv<VT><<<(unsigned)((N + 255) / 256), 256, 0, st>>>(N, adj_ptr, dslot, val, fixed, d);
  }
};

struct CscOp {
  int64_t n;
  const int64_t* colptr;
  const int32_t* rowval;
  int64_t ndof() const { return n; }
  unsigned grid() const { return (unsigned)((n * kG + kBS - 1) / kBS); }
  template <typename MT, typename VT>
  void spmv(hipStream_t st, const MT* val, const VT* x, const uint8_t* fixed, VT* y, double* part,
            const double* scal) const {
    k_spmv_csc<MT, VT><<<grid(), kBS, 0, st>>>(n, colptr, rowval, val, x, fixed, y, part, scal);
  }
  template <typename VT>
  void dinv(hipStream_t st, const double* val, const uint8_t* fixed, VT* d) const {
    k_dinv_csc<VT><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, colptr, rowval, val, fixed, d);
  }
};

inline unsigned vgrid(int64_t n, llfem_ctx* ctx) {
  int64_t g = (n + kBS - 1) / kBS;
  int64_t cap = (int64_t)std::max(ctx->num_cus, 1) * 8;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

struct Scratch {   // partial sums + scalars
  double* part;
  double* red;
  double* scal;
};

// Jacobi-PCG (IterativeSolvers.cg semantics, x0 = 0) on vectors of type VT.
template <typename Op, typename MT, typename VT>
int pcg(llfem_ctx* ctx, const Op& K, const MT* val, const uint8_t* fixed, const VT* dinv, const VT* b,
        VT* x, VT* r, VT* z, VT* p, VT* q, const Scratch& S, double rtol, int64_t maxiter,
        int64_t* iters, double* res_rel) {
  hipStream_t st = ctx->stream;
  const int64_t n = K.ndof();
  const unsigned gv = vgrid(n, ctx);
  const unsigned gs = K.grid();
  LLFEM_HIP(hipMemcpyAsync(r, b, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
  k_init<VT><<<gv, kBS, 0, st>>>(n, x, r, dinv, z, p, S.part);
  k_reduce<<<1, kBS, 0, st>>>(gv, 2, S.part, S.red);
  k_init_scal<<<1, 1, 0, st>>>(S.scal, S.red, rtol);
  double h[S_NS];
  const int poll = 16;
  for (int64_t it = 0; it < maxiter;) {
    const int64_t batch = std::min<int64_t>(poll, maxiter - it);
    for (int64_t k = 0; k < batch; ++k) {
      K.template spmv<MT, VT>(st, val, p, fixed, q, S.part, S.scal);
      k_reduce<<<1, kBS, 0, st>>>(gs, 1, S.part, S.red);
      k_alpha<<<1, 1, 0, st>>>(S.scal, S.red);
      k_update<VT><<<gv, kBS, 0, st>>>(n, S.scal, x, r, p, q, dinv, z, S.part);
      k_reduce<<<1, kBS, 0, st>>>(gv, 2, S.part, S.red);
      k_beta<<<1, 1, 0, st>>>(S.scal, S.red, maxiter);
      k_pupd<VT><<<gv, kBS, 0, st>>>(n, S.scal, z, p);
    }
    it += batch;
    LLFEM_HIP(hipMemcpyAsync(h, S.scal, sizeof(h), hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipStreamSynchronize(st));
    if (h[S_DONE] != 0.0) break;
  }
  LLFEM_HIP(hipMemcpyAsync(h, S.scal, sizeof(h), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  LLFEM_HIP(hipGetLastError());
  *iters = (int64_t)h[S_ITER];
  const double b2 = rtol > 0 ? h[S_TOL2] / (rtol * rtol) : 0.0;
  *res_rel = b2 > 0 ? std::sqrt(h[S_RR] / b2) : 0.0;
  return LLFEM_OK;
}

__global__ void k_zero(int64_t n, double* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

// Device workspace of solve_device_ws: 8 FP64 vectors of K.ndof(), the
// reduction partials and scalars, and 7 FP32 vectors for the mixed method.
template <typename Op>
size_t solve_ws_bytes(llfem_ctx* ctx, const Op& K, int method) {
  const int64_t n = K.ndof();
  const int64_t nparts = std::max<int64_t>(K.grid(), 2 * (int64_t)vgrid(n, ctx)) + 2;
  const size_t nf = method == LLFEM_PCG_MIXED_REFINE ? 7 : 0;
  return sizeof(double) * (8 * (size_t)n + (size_t)nparts + 8 + S_NS + 8) + sizeof(float) * (nf * (size_t)n + 8);
}

// Full solve on device data: f, ufix (nullable) and u are device vectors of
// K.ndof(); fixed is a device 0/1 mask.  val32 is needed for the mixed method.
// `base` holds solve_ws_bytes(ctx, K, method) bytes; every part of it is
// written before it is read, so it may hold anything on entry.
template <typename Op>
int solve_device_ws(llfem_ctx* ctx, const Op& K, const double* val, const float* val32,
                    const uint8_t* fixed, const double* f, const double* ufix, int method, double reltol,
                    int64_t maxiter, double* u, llfem_stats* stats, void* base) {
  hipStream_t st = ctx->stream;
  const int64_t n = K.ndof();
  const unsigned gv = vgrid(n, ctx);
  const int64_t nparts = std::max<int64_t>(K.grid(), 2 * (int64_t)gv) + 2;
  const bool mixed = (method == LLFEM_PCG_MIXED_REFINE);
  double* b = (double*)base;
  double* x = b + n;
  double* r = x + n;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  double* dinv = q + n;
  double* y = dinv + n;
  Scratch S{y + n, y + n + nparts, y + n + nparts + 8};
  float* f32 = (float*)(S.scal + S_NS + 8);
  float *b32 = f32, *x32 = b32 + n, *r32 = x32 + n, *z32 = r32 + n, *p32 = z32 + n, *q32 = p32 + n,
        *dinv32 = q32 + n;

  K.template dinv<double>(st, val, fixed, dinv);
  if (mixed) K.template dinv<float>(st, val, fixed, dinv32);
  // u = u_fixed on constrained DoFs; b = mask(f - K u)  (src/linear.jl:3730-3733)
  k_masked_set<<<gv, kBS, 0, st>>>(n, fixed, ufix, u);
  K.template spmv<double, double>(st, val, u, (const uint8_t*)nullptr, y, nullptr, nullptr);
  k_resid<<<gv, kBS, 0, st>>>(n, f, y, fixed, b, S.part);
  k_reduce<<<1, kBS, 0, st>>>(gv, 1, S.part, S.red);
  double bn2 = 0.0;
  LLFEM_HIP(hipMemcpyAsync(&bn2, S.red, sizeof(double), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  const double bnorm = std::sqrt(bn2);
  int64_t total_it = 0, refinements = 0;
  bool converged = false;
  if (!mixed) {
    int64_t it = 0;
    double rr = 0.0;
    LLFEM_TRY((pcg<Op, double, double>(ctx, K, val, fixed, dinv, b, x, r, z, p, q, S, reltol, maxiter,
                                       &it, &rr)));
    total_it = it;
    converged = rr <= reltol;
    k_axpy<<<gv, kBS, 0, st>>>(n, 1.0, x, u);
  } else {
    // x accumulates the free-DoF correction in FP64; r is the FP64 residual
    k_zero<<<gv, kBS, 0, st>>>(n, x);
    LLFEM_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    double rnorm = bnorm;
    const double inner_floor = 1e-5;
    const int max_ref = 40;
    while (refinements < max_ref && total_it < maxiter) {
      if (rnorm <= reltol * bnorm || bnorm == 0.0) { converged = true; break; }
      k_to_f32<<<gv, kBS, 0, st>>>(n, r, 1.0 / rnorm, b32);
      const double need = 0.5 * reltol * bnorm / rnorm;
      const double rin = std::max(inner_floor, std::min(1e-2, need));
      int64_t it = 0;
      double rr = 0.0;
      LLFEM_TRY((pcg<Op, float, float>(ctx, K, val32, fixed, dinv32, b32, x32, r32, z32, p32, q32, S,
                                       rin, maxiter - total_it, &it, &rr)));
      total_it += it;
      k_axpy_f32<<<gv, kBS, 0, st>>>(n, rnorm, x32, x);
      // FP64 residual r = mask(b - K x)
      K.template spmv<double, double>(st, val, x, fixed, y, nullptr, nullptr);
      k_resid<<<gv, kBS, 0, st>>>(n, b, y, fixed, r, S.part);
      k_reduce<<<1, kBS, 0, st>>>(gv, 1, S.part, S.red);
      double r2 = 0.0;
      LLFEM_HIP(hipMemcpyAsync(&r2, S.red, sizeof(double), hipMemcpyDeviceToHost, st));
      LLFEM_HIP(hipStreamSynchronize(st));
      rnorm = std::sqrt(r2);
      ++refinements;
      if (it == 0) break;
    }
    if (!converged && rnorm <= reltol * bnorm) converged = true;
    k_axpy<<<gv, kBS, 0, st>>>(n, 1.0, x, u);
  }
  // reported residual: ||(K u - f)_free|| / ||f_free|| (test/test_dsl.jl:106)
  K.template spmv<double, double>(st, val, u, (const uint8_t*)nullptr, y, nullptr, nullptr);
  k_resid<<<gv, kBS, 0, st>>>(n, f, y, fixed, r, S.part);
  k_reduce<<<1, kBS, 0, st>>>(gv, 1, S.part, S.red);
  double rr2 = 0.0;
  LLFEM_HIP(hipMemcpyAsync(&rr2, S.red, sizeof(double), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  // ||f_free||
  k_zero<<<gv, kBS, 0, st>>>(n, y);
  k_resid<<<gv, kBS, 0, st>>>(n, f, y, fixed, z, S.part);
  k_reduce<<<1, kBS, 0, st>>>(gv, 1, S.part, S.red);
  double ff2 = 0.0;
  LLFEM_HIP(hipMemcpyAsync(&ff2, S.red, sizeof(double), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  LLFEM_HIP(hipGetLastError());
  const double relres = ff2 > 0 ? std::sqrt(rr2 / ff2) : std::sqrt(rr2);
  if (stats) {
    stats->iterations = total_it;
    stats->rollbacks = 0;
    stats->restarts = 0;
    stats->refinements = refinements;
    stats->relres = relres;
  }
  if (!converged)
    return fail(LLFEM_ENOCONV, "cg: maxiter reached (relative residual " + std::to_string(relres) + ")");
  return LLFEM_OK;
}

// solve_device_ws on a workspace allocated for this call.
template <typename Op>
int solve_device(llfem_ctx* ctx, const Op& K, const double* val, const float* val32,
                 const uint8_t* fixed, const double* f, const double* ufix, int method, double reltol,
                 int64_t maxiter, double* u, llfem_stats* stats) {
  const size_t bytes = solve_ws_bytes(ctx, K, method);
  void* base = nullptr;
  LLFEM_TRY(dmalloc(ctx, &base, bytes));
  struct Free {
    llfem_ctx* c; void* p; size_t b;
    ~Free() { (void)hipStreamSynchronize(c->stream); dfree(c, p, b); }
  } guard{ctx, base, bytes};
  return solve_device_ws(ctx, K, val, val32, fixed, f, ufix, method, reltol, maxiter, u, stats, base);
}

__global__ void k_set_fixed(int64_t nfix, const int64_t* __restrict__ idx, int64_t base, int64_t n,
                            const double* __restrict__ vals, uint8_t* __restrict__ mask,
                            double* __restrict__ ufix, int* __restrict__ err) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfix) return;
  const int64_t i = idx[k] - base;
  if (i < 0 || i >= n) { atomicOr(err, 1); return; }
  mask[i] = 1;
  ufix[i] = vals ? vals[k] : 0.0;
}

// Upload f, the fixed set and u_fixed; run solve_device; download u.
template <typename Op>
int solve_host(llfem_ctx* ctx, const Op& K, const double* val, const float* val32, int64_t n,
               const double* f, const int64_t* fixed, int64_t nfixed, int32_t one_based,
               const double* u_fixed, int32_t method, double reltol, int64_t maxiter, double* u_out,
               llfem_stats* stats) {
  hipStream_t st = ctx->stream;
  const size_t bytes = sizeof(double) * 3 * (size_t)n + (size_t)n + sizeof(int64_t) * (size_t)(nfixed + 1) +
                       sizeof(double) * (size_t)(nfixed + 1) + 64;
  char* base = nullptr;
  LLFEM_TRY(dmalloc(ctx, (void**)&base, bytes));
  struct Free {
    llfem_ctx* c; void* p; size_t b;
    ~Free() { (void)hipStreamSynchronize(c->stream); dfree(c, p, b); }
  } guard{ctx, base, bytes};
  double* d_f = (double*)base;
  double* d_u = d_f + n;
  double* d_ufix = d_u + n;
  int64_t* d_idx = (int64_t*)(d_ufix + n);
  double* d_vals = (double*)(d_idx + nfixed + 1);
  int* d_err = (int*)(d_vals + nfixed + 1);
  uint8_t* d_mask = (uint8_t*)(d_err + 2);
  LLFEM_HIP(hipMemcpyAsync(d_f, f, sizeof(double) * n, hipMemcpyHostToDevice, st));
  LLFEM_HIP(hipMemsetAsync(d_mask, 0, (size_t)n, st));
  LLFEM_HIP(hipMemsetAsync(d_ufix, 0, sizeof(double) * n, st));
  LLFEM_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
  if (nfixed > 0) {
    LLFEM_HIP(hipMemcpyAsync(d_idx, fixed, sizeof(int64_t) * nfixed, hipMemcpyHostToDevice, st));
    if (u_fixed)
      LLFEM_HIP(hipMemcpyAsync(d_vals, u_fixed, sizeof(double) * nfixed, hipMemcpyHostToDevice, st));
    k_set_fixed<<<(unsigned)((nfixed + 255) / 256), 256, 0, st>>>(nfixed, d_idx, one_based ? 1 : 0, n,
                                                                   u_fixed ? d_vals : nullptr, d_mask,
                                                                   d_ufix, d_err);
  }
  int herr = 0;
  LLFEM_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  if (herr) return fail(LLFEM_EINVAL, "solveDisplacement: constrained DoF out of range");
  auto t0 = std::chrono::steady_clock::now();
  int s = solve_device(ctx, K, val, val32, d_mask, d_f, d_ufix, method, reltol, maxiter, d_u, stats);
  if (s != LLFEM_OK && s != LLFEM_ENOCONV) return s;
  LLFEM_HIP(hipMemcpyAsync(u_out, d_u, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  if (stats)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return s;
}

__global__ void k_i64_to_i32(int64_t n, const int64_t* __restrict__ a, int64_t base, int32_t* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (int32_t)(a[i] - base);
}

__global__ void k_i64_sub(int64_t n, int64_t* __restrict__ a, int64_t base) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] -= base;
}

}  // namespace

// Upload f and the fixed set, run one system through the batched engine,
// download u.
int solve_host_reliable(llfem_ctx* ctx, SysSpec sp, bool f64_vectors, const double* f, const int64_t* fixed,
                        int64_t nfixed, int32_t one_based, const double* u_fixed, double* u_out,
                        llfem_stats* stats, llfem_mesh* amg_mesh) {
  hipStream_t st = ctx->stream;
  const int64_t n = 3 * sp.N;
  const size_t bytes = sizeof(double) * 3 * (size_t)n + (size_t)n + sizeof(int64_t) * (size_t)(nfixed + 1) +
                       sizeof(double) * (size_t)(nfixed + 1) + 64;
  char* base = nullptr;
  LLFEM_TRY(dmalloc(ctx, (void**)&base, bytes));
  struct Free {
    llfem_ctx* c; void* p; size_t b;
    ~Free() { (void)hipStreamSynchronize(c->stream); dfree(c, p, b); }
  } guard{ctx, base, bytes};
  double* d_f = (double*)base;
  double* d_u = d_f + n;
  double* d_ufix = d_u + n;
  int64_t* d_idx = (int64_t*)(d_ufix + n);
  double* d_vals = (double*)(d_idx + nfixed + 1);
  int* d_err = (int*)(d_vals + nfixed + 1);
  uint8_t* d_mask = (uint8_t*)(d_err + 2);
  LLFEM_HIP(hipMemcpyAsync(d_f, f, sizeof(double) * n, hipMemcpyHostToDevice, st));
  LLFEM_HIP(hipMemsetAsync(d_mask, 0, (size_t)n, st));
  LLFEM_HIP(hipMemsetAsync(d_ufix, 0, sizeof(double) * n, st));
  LLFEM_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
  if (nfixed > 0) {
    LLFEM_HIP(hipMemcpyAsync(d_idx, fixed, sizeof(int64_t) * nfixed, hipMemcpyHostToDevice, st));
    if (u_fixed)
      LLFEM_HIP(hipMemcpyAsync(d_vals, u_fixed, sizeof(double) * nfixed, hipMemcpyHostToDevice, st));
    k_set_fixed<<<(unsigned)((nfixed + 255) / 256), 256, 0, st>>>(nfixed, d_idx, one_based ? 1 : 0, n,
                                                                   u_fixed ? d_vals : nullptr, d_mask,
                                                                   d_ufix, d_err);
  }
  int herr = 0;
  LLFEM_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  if (herr) return fail(LLFEM_EINVAL, "solveDisplacement: constrained DoF out of range");
  auto t0 = std::chrono::steady_clock::now();
  sp.fixed = d_mask;
  sp.f = d_f;
  sp.ufix = d_ufix;
  if (amg_mesh) {
    // the hierarchy of this support, built on first use and kept on the mesh
    uint64_t key = hash_bytes(&one_based, sizeof(one_based));
    key = hash_bytes(&nfixed, sizeof(nfixed), key);
    if (nfixed > 0) key = hash_bytes(fixed, sizeof(int64_t) * (size_t)nfixed, key);
    auto it = amg_mesh->amg.find(key);
    if (it == amg_mesh->amg.end()) {
      uint8_t* keep = nullptr;
      LLFEM_TRY(dalloc(ctx, &keep, (size_t)n));
      LLFEM_HIP(hipMemcpyAsync(keep, d_mask, (size_t)n, hipMemcpyDeviceToDevice, st));
      sp.fixed = keep;
      AmgHier* h = nullptr;
      const int rc = amg_build(ctx, sp, amg_mesh->d_xyz, &h);
      if (rc != LLFEM_OK) {
        (void)hipStreamSynchronize(st);
        dfree(ctx, keep, (size_t)n);
        return rc;
      }
      it = amg_mesh->amg.emplace(key, std::make_pair(h, keep)).first;
    }
    sp.fixed = it->second.second;
    sp.amg = it->second.first;
  }
  BatchSolver eng(ctx, f64_vectors);
  SolveSys* s = nullptr;
  LLFEM_TRY(eng.add(sp, &s));
  const double setup_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // AMG systems reach the reliable-update trigger every ~10 iterations: short
  // chunks keep a paused system from idling through the rest of a chunk
  const int chunk = sp.amg ? 8 : 32;
  while (sys_status(s) == 0) {
    std::vector<SolveSys*> done;
    LLFEM_TRY(eng.run(chunk, &done));
  }
  LLFEM_TRY(eng.finish(s, d_u));
  LLFEM_HIP(hipMemcpyAsync(u_out, d_u, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  const int status = sys_status(s);
  if (stats) {
    stats->iterations = sys_iterations(s);
    stats->rollbacks = sys_rollbacks(s);
    stats->restarts = sys_refreshes(s);
    stats->refinements = sys_updates(s);
    stats->relres = sys_relres(s);
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->setup_ms = setup_ms;
  }
  const double rr = sys_relres(s);
  eng.remove(s);
  if (status == -2)
    return fail(LLFEM_ENOCONV, "cg: the FP32 correction was not finite in three successive updates");
  if (status != 1)
    return fail(LLFEM_ENOCONV, "cg: maxiter reached (relative residual " + std::to_string(rr) + ")");
  return LLFEM_OK;
}

// Make sure the mesh's FP32 copy of K exists (mixed-precision solve).
int ensure_k32(llfem_mesh* m) {
  if (m->d_valK32 && m->solver_ready) return LLFEM_OK;
  llfem_ctx* ctx = m->ctx;
  const int64_t nnz = 9 * m->nadj;
  if (!m->d_valK32) LLFEM_TRY(dalloc(ctx, &m->d_valK32, (size_t)nnz));
  k_cast_val<<<vgrid(nnz, ctx), kBS, 0, ctx->stream>>>(nnz, m->d_valK, m->d_valK32);
  m->solver_ready = true;
  return LLFEM_OK;
}

namespace {
__global__ void k_check_csc(int64_t n, const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                            int64_t nnz, int* __restrict__ err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    if (colptr[j + 1] < colptr[j]) atomicOr(err, 1);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += stride)
    if (rowval[k] < 0 || (int64_t)rowval[k] >= n) atomicOr(err, 2);
}

__global__ void k_zero_int(int* a) { *a = 0; }
}  // namespace

}  // namespace llfem

using namespace llfem;

extern "C" {

int llfem_solve(llfem_mesh* m, const double* f, const int64_t* fixed, int64_t nfixed, int32_t one_based,
                const double* u_fixed, int32_t method, double reltol, int64_t maxiter, double* u_out,
                llfem_stats* stats) {
  if (!m || !f || !u_out || (nfixed > 0 && !fixed)) return fail(LLFEM_EINVAL, "llfem_solve: NULL argument");
  if (!m->have_K) return fail(LLFEM_EINVAL, "llfem_solve: no stiffness assembled on this mesh");
  if (method != LLFEM_PCG_JACOBI_FP64 && method != LLFEM_PCG_MIXED_REFINE &&
      method != LLFEM_PCG_MIXED_RELIABLE && method != LLFEM_PCG_F32OP_RELIABLE &&
      method != LLFEM_PCG_AMG_RELIABLE)
    return fail(LLFEM_EUNSUP, "llfem_solve: unknown method");
  if (!(reltol > 0)) return fail(LLFEM_EINVAL, "llfem_solve: reltol must be positive");
  LLFEM_HIP(hipSetDevice(m->ctx->device));
  if (method == LLFEM_PCG_MIXED_RELIABLE || method == LLFEM_PCG_F32OP_RELIABLE ||
      method == LLFEM_PCG_AMG_RELIABLE) {
    SysSpec sp;
    sp.N = m->N;
    sp.nadj = m->nadj;
    sp.adj_ptr = m->d_adj_ptr;
    sp.adj = m->d_adj;
    sp.dslot = m->d_diag_slot;
    sp.val64 = m->d_valK;
    sp.reltol = reltol;
    sp.maxiter = maxiter;
    return solve_host_reliable(m->ctx, sp, method == LLFEM_PCG_F32OP_RELIABLE, f, fixed, nfixed, one_based,
                               u_fixed, u_out, stats, method == LLFEM_PCG_AMG_RELIABLE ? m : nullptr);
  }
  if (method == LLFEM_PCG_MIXED_REFINE) LLFEM_TRY(ensure_k32(m));
  PanelOp K{m->N, m->d_adj_ptr, m->d_adj, m->d_diag_slot};
  return solve_host(m->ctx, K, m->d_valK, m->d_valK32, 3 * m->N, f, fixed, nfixed, one_based, u_fixed,
                    method, reltol, maxiter, u_out, stats);
}

int llfem_solve_csc(llfem_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval,
                    const double* nzval, int32_t one_based, const double* f, const int64_t* fixed,
                    int64_t nfixed, const double* u_fixed, int32_t method, double reltol, int64_t maxiter,
                    double* u_out, llfem_stats* stats) {
  if (!ctx || !colptr || !rowval || !nzval || !f || !u_out || (nfixed > 0 && !fixed))
    return fail(LLFEM_EINVAL, "llfem_solve_csc: NULL argument");
  if (n <= 0 || n >= ((int64_t)1 << 31)) return fail(LLFEM_EINVAL, "llfem_solve_csc: bad size");
  if (method != LLFEM_PCG_JACOBI_FP64 && method != LLFEM_PCG_MIXED_REFINE)
    return fail(LLFEM_EUNSUP, "llfem_solve_csc: unknown method");
  if (!(reltol > 0)) return fail(LLFEM_EINVAL, "llfem_solve_csc: reltol must be positive");
  const int64_t base = one_based ? 1 : 0;
  const int64_t nnz = colptr[n] - base;
  if (nnz < 0 || colptr[0] != base) return fail(LLFEM_EINVAL, "llfem_solve_csc: bad colptr");
  // the device SpMV trusts these indices: check them here (O(n + nnz))
  for (int64_t j = 0; j < n; ++j)
    if (colptr[j + 1] < colptr[j])
      return fail(LLFEM_EINVAL, "llfem_solve_csc: colptr is not non-decreasing at column " +
                                    std::to_string(j + base));
  for (int64_t k = 0; k < nnz; ++k)
    if (rowval[k] < base || rowval[k] >= n + base)
      return fail(LLFEM_EINVAL, "llfem_solve_csc: row index " + std::to_string(rowval[k]) + " out of range");
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t bytes = sizeof(int64_t) * (size_t)(n + 1) + sizeof(int64_t) * (size_t)nnz +
                       sizeof(int32_t) * (size_t)nnz + sizeof(double) * (size_t)nnz +
                       (method == LLFEM_PCG_MIXED_REFINE ? sizeof(float) * (size_t)nnz : 0) + 64;
  char* basep = nullptr;
  LLFEM_TRY(dmalloc(ctx, (void**)&basep, bytes));
  struct Free {
    llfem_ctx* c; void* p; size_t b;
    ~Free() { (void)hipStreamSynchronize(c->stream); dfree(c, p, b); }
  } guard{ctx, basep, bytes};
  int64_t* d_cp = (int64_t*)basep;
  int64_t* d_rv64 = d_cp + n + 1;
  double* d_val = (double*)(d_rv64 + nnz);
  int32_t* d_rv = (int32_t*)(d_val + nnz);
  float* d_val32 = (float*)(d_rv + nnz + (nnz & 1));
  LLFEM_HIP(hipMemcpyAsync(d_cp, colptr, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
  LLFEM_HIP(hipMemcpyAsync(d_rv64, rowval, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, st));
  LLFEM_HIP(hipMemcpyAsync(d_val, nzval, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
  k_i64_sub<<<vgrid(n + 1, ctx), kBS, 0, st>>>(n + 1, d_cp, base);
  k_i64_to_i32<<<vgrid(nnz, ctx), kBS, 0, st>>>(nnz, d_rv64, base, d_rv);
  if (method == LLFEM_PCG_MIXED_REFINE) k_cast_val<<<vgrid(nnz, ctx), kBS, 0, st>>>(nnz, d_val, d_val32);
  CscOp K{n, d_cp, d_rv};
  return solve_host(ctx, K, d_val, d_val32, n, f, fixed, nfixed, one_based, u_fixed, method, reltol,
                    maxiter, u_out, stats);
}

int llfem_solve_csc_noalloc(llfem_ctx* ctx, int64_t n, const int64_t* d_colptr, const int32_t* d_rowval,
                            const double* d_nzval, const double* d_f, const uint8_t* d_fixed,
                            const double* d_u_fixed, double reltol, int64_t maxiter, double* d_u_out,
                            llfem_stats* stats) {
  if (!ctx || !d_colptr || !d_rowval || !d_nzval || !d_f || !d_fixed || !d_u_out)
    return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: NULL argument");
  if (n <= 0 || n >= ((int64_t)1 << 31)) return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: bad size");
  if (!(reltol > 0)) return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: reltol must be positive");
  CscOp K{n, d_colptr, d_rowval};
  // the flag word of the index check sits after the solver's workspace
  const size_t ws = (solve_ws_bytes(ctx, K, LLFEM_PCG_JACOBI_FP64) + 255) & ~(size_t)255;
  if (ws + 256 > kNoAllocScratchBytes)
    return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: ndof " + std::to_string(n) + " needs " +
                                  std::to_string(ws + 256) + " bytes of workspace; the maximum is " +
                                  std::to_string(kNoAllocScratchBytes) + " (about " +
                                  std::to_string(kNoAllocScratchBytes / 65) + " DoFs)");
  const int dev = ctx->device;
  void* ws_ptr = nullptr;
  std::mutex* ws_lock = nullptr;
  if (!noalloc_scratch(dev, &ws_ptr, &ws_lock))
    return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: no workspace on device " + std::to_string(dev));
  LLFEM_HIP(hipSetDevice(dev));
  std::lock_guard<std::mutex> lk(*ws_lock);
  unsigned char* base = (unsigned char*)ws_ptr;
  int* d_err = (int*)(base + ws);
  hipStream_t st = ctx->stream;
  auto t0 = std::chrono::steady_clock::now();
  // the same index checks as llfem_solve_csc, on the device
  int64_t ends[2] = {0, 0};
  LLFEM_HIP(hipMemcpyAsync(&ends[0], d_colptr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipMemcpyAsync(&ends[1], d_colptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  if (ends[0] != 0 || ends[1] < 0) return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: bad colptr");
  k_zero_int<<<1, 1, 0, st>>>(d_err);
  k_check_csc<<<vgrid(std::max<int64_t>(n, ends[1]), ctx), kBS, 0, st>>>(n, d_colptr, d_rowval, ends[1], d_err);
  int herr = 0;
  LLFEM_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  if (herr & 1) return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: colptr is not non-decreasing");
  if (herr & 2) return fail(LLFEM_EINVAL, "llfem_solve_csc_noalloc: row index out of range");
  const int s = solve_device_ws(ctx, K, d_nzval, nullptr, d_fixed, d_f, d_u_fixed, LLFEM_PCG_JACOBI_FP64,
                                reltol, maxiter, d_u_out, stats, base);
  if (s != LLFEM_OK && s != LLFEM_ENOCONV) return s;
  LLFEM_HIP(hipStreamSynchronize(st));
  if (stats)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return s;
}

}  // extern "C"
