// Batched mixed-precision PCG engine: FP32 Krylov iterations with FP64
// reliable updates, 3x3 block-Jacobi preconditioner, many systems per launch.
//
// Reference: solveDisplacement(K, f; support, iterative=true, reltol, ...)
// (src/linear.jl:3711-3743): u = u_fixed on the constrained DoFs and
// K_ff u_f = f_f - K_fc u_c by preconditioned CG, converged at
// ||r|| <= reltol ||r0||.  Here the convergence test is made on the TRUE
// FP64 residual ||b - K x|| <= reltol ||b|| (b = the eliminated right-hand
// side), so the mixed-precision result meets the same relative residual as
// the FP64 solver.
//
// Algorithm (reliable updates, Sleijpen & van der Vorst 1996, as used for
// mixed-precision CG in lattice QCD solvers):
//   * CG runs in FP32 on K32 (the assembled panels rounded to FP32) with the
//     residual scaled to unit norm; dot products accumulate in FP64.
//   * The FP32 correction x32 is folded into the FP64 solution and the true
//     residual r = b - K x is recomputed in FP64 whenever the recursive
//     residual has dropped by delta (0.1) since the last update, or reaches
//     the tolerance.  The search direction p is kept (rescaled), so the
//     Krylov process is not restarted and the iteration count stays close to
//     the FP64 solver's.
//
// Batching: every launch covers all RUNNING systems.  Blocks map to systems
// through per-launch tables; each system's rows are split into blocks that
// depend on that system only and its dot products reduce in a fixed order,
// so a system's iterates are bit-identical whether it runs alone or among
// others.  A system that needs a reliable update pauses (its blocks exit at
// once) until the host performs the update between chunks of iterations.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"
#include "engine_policy.h"

namespace llfem {

namespace {

constexpr int kSB = 256;                 // threads per block
constexpr int kSG = 8;                   // lanes per node row in the SpMV
constexpr int kSpmvNodes = kSB / kSG;    // nodes per SpMV block
constexpr int kVecNodes = kSB;           // nodes per vector block pass
// Blocks per system are capped (the blocks stride over the system's nodes),
// so the per-system reductions sum at most a few thousand partials.
constexpr int64_t kMaxSpmvBlocks = 2048;
constexpr int64_t kMaxVecBlocks = 1024;

// scal words of a system; SC_LAST: iteration count at the last FP64 update,
// SC_MAXRUN: force an update after this many iterations (0: never)
enum { SC_RHO = 0, SC_ALPHA, SC_BETA, SC_RR, SC_MAXRR, SC_TOL2, SC_DELTA2, SC_STATE, SC_ITER, SC_PQ, SC_LAST,
       SC_MAXRUN, SC_N = 16 };
enum { ST_RUN = 0, ST_PENDING = 1, ST_PAUSED = 2, ST_IDLE = 3 };

struct SysDev {
  int64_t N;               // nodes (3 DoFs each)
  const int64_t* adj_ptr;
  const int32_t* adj;
  const float* val;        // FP32 node blocks, block-contiguous: [adj slot][c][d]
  const uint8_t* fixed;    // [3N] 1 = constrained
  const float* binv;       // [N][9] inverse diagonal blocks (free part)
  void *x, *r, *z, *p, *q; // Krylov vectors, FP32 or FP64 (the kernels' VT)
  double* scal;            // [SC_N]
  double* part;            // [nsblk] SpMV partials, then [2 * nvblk]
  int32_t sblk0, nsblk, vblk0, nvblk;
  int32_t amg;             // 1: z and the r.z partials come from the AMG V-cycle
};

__device__ __forceinline__ double bsum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// q = mask(K32 p) for RUNNING systems; partial p.q per block.
template <typename VT>
__global__ __launch_bounds__(kSB) void k_bs_spmv(const SysDev* __restrict__ sys, const int32_t* __restrict__ bsys) {
  __shared__ double sh[kSB];
  const SysDev& S = sys[bsys[blockIdx.x]];
  if (S.scal[SC_STATE] != ST_RUN) return;   // block-uniform
  const VT* __restrict__ pv = (const VT*)gp(static_cast<const VT*>(S.p));
  VT* __restrict__ qv = (VT*)gp(static_cast<VT*>(S.q));
  const int64_t* __restrict__ adj_ptr = (const int64_t*)gp(S.adj_ptr);
  const int32_t* __restrict__ adj = (const int32_t*)gp(S.adj);
  const float* __restrict__ val = (const float*)gp(S.val);
  const uint8_t* __restrict__ fixed = (const uint8_t*)gp(S.fixed);
  const int lb = blockIdx.x - S.sblk0;
  const int l = threadIdx.x % kSG;
  double d = 0.0;
  for (int64_t j = (int64_t)lb * kSpmvNodes + threadIdx.x / kSG; j - threadIdx.x / kSG < S.N;
       j += (int64_t)S.nsblk * kSpmvNodes) {
  VT a0 = 0, a1 = 0, a2 = 0;
  if (j < S.N) {
    const int64_t p0 = adj_ptr[j];
    const int R = (int)(adj_ptr[j + 1] - p0);
    const float* P = val + 9 * p0;
#pragma unroll 2
    for (int t = l; t < R; t += kSG) {
      const int64_t i = adj[p0 + t];
      const VT x0 = pv[3 * i], x1 = pv[3 * i + 1], x2 = pv[3 * i + 2];
      const float* B = P + 9 * t;   // block t: [c][d], 36 contiguous bytes
      a0 += (VT)B[0] * x0 + (VT)B[1] * x1 + (VT)B[2] * x2;
      a1 += (VT)B[3] * x0 + (VT)B[4] * x1 + (VT)B[5] * x2;
      a2 += (VT)B[6] * x0 + (VT)B[7] * x1 + (VT)B[8] * x2;
    }
  }
#pragma unroll
  for (int o = kSG / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kSG);
    a1 += __shfl_xor(a1, o, kSG);
    a2 += __shfl_xor(a2, o, kSG);
  }
  if (j < S.N && l == 0) {
    const VT a[3] = {a0, a1, a2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const VT v = fixed[3 * j + c] ? (VT)0 : a[c];
      qv[3 * j + c] = v;
      d += (double)pv[3 * j + c] * (double)v;
    }
  }
  }
  d = bsum(d, sh);
  if (threadIdx.x == 0) S.part[lb] = d;
}

// alpha = rho / p.q (one block per system).  A system flagged PENDING at the
// end of the previous iteration becomes PAUSED here.
__global__ __launch_bounds__(kSB) void k_bs_alpha(const SysDev* __restrict__ sys) {
  __shared__ double sh[kSB];
  const SysDev& S = sys[blockIdx.x];
  const double st = S.scal[SC_STATE];
  if (st == ST_PENDING) {
    if (threadIdx.x == 0) S.scal[SC_STATE] = ST_PAUSED;
    return;
  }
  if (st != ST_RUN) return;
  double v = 0.0;
  for (int i = threadIdx.x; i < S.nsblk; i += kSB) v += S.part[i];
  v = bsum(v, sh);
  if (threadIdx.x == 0) {
    S.scal[SC_PQ] = v;
    if (v > 0.0) {
      S.scal[SC_ALPHA] = S.scal[SC_RHO] / v;
    } else {   // breakdown (p = 0 or K not SPD on the free set): check the true residual
      S.scal[SC_ALPHA] = 0.0;
      S.scal[SC_STATE] = ST_PAUSED;
    }
  }
}

// x += alpha p; r -= alpha q; z = Binv r; partials (r.r, r.z).  Thread per node.
template <typename VT>
__global__ __launch_bounds__(kSB) void k_bs_update(const SysDev* __restrict__ sys, const int32_t* __restrict__ bsys) {
  __shared__ double sh[kSB];
  const SysDev& S = sys[bsys[blockIdx.x]];
  if (S.scal[SC_STATE] != ST_RUN) return;
  VT* __restrict__ xv = (VT*)gp(static_cast<VT*>(S.x));
  VT* __restrict__ rv = (VT*)gp(static_cast<VT*>(S.r));
  VT* __restrict__ zv = (VT*)gp(static_cast<VT*>(S.z));
  const VT* __restrict__ pv = (const VT*)gp(static_cast<const VT*>(S.p));
  const VT* __restrict__ qv = (const VT*)gp(static_cast<const VT*>(S.q));
  const float* __restrict__ binv = (const float*)gp(S.binv);
  const int lb = blockIdx.x - S.vblk0;
  double rr = 0.0, rz = 0.0;
  const VT alpha = (VT)S.scal[SC_ALPHA];
  for (int64_t j = (int64_t)lb * kVecNodes + threadIdx.x; j < S.N; j += (int64_t)S.nvblk * kVecNodes) {
    VT r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = 3 * j + c;
      xv[i] += alpha * pv[i];
      r[c] = rv[i] - alpha * qv[i];
      rv[i] = r[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) rr += (double)r[c] * (double)r[c];
    if (S.amg) continue;
    const float* B = binv + 9 * j;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const VT z = (VT)B[3 * c] * r[0] + (VT)B[3 * c + 1] * r[1] + (VT)B[3 * c + 2] * r[2];
      zv[3 * j + c] = z;
      rz += (double)r[c] * (double)z;
    }
  }
  rr = bsum(rr, sh);
  rz = bsum(rz, sh);
  if (threadIdx.x == 0) {
    S.part[S.nsblk + 2 * lb] = rr;
    if (!S.amg) S.part[S.nsblk + 2 * lb + 1] = rz;
  }
}

// beta = rz / rho; reliable-update / convergence triggers (one block per system).
__global__ __launch_bounds__(kSB) void k_bs_beta(const SysDev* __restrict__ sys) {
  __shared__ double sh[kSB];
  const SysDev& S = sys[blockIdx.x];
  if (S.scal[SC_STATE] != ST_RUN) return;
  double rr = 0.0, rz = 0.0;
  for (int i = threadIdx.x; i < S.nvblk; i += kSB) {
    rr += S.part[S.nsblk + 2 * i];
    rz += S.part[S.nsblk + 2 * i + 1];
  }
  rr = bsum(rr, sh);
  rz = bsum(rz, sh);
  if (threadIdx.x == 0) {
    const double rho = S.scal[SC_RHO];
    S.scal[SC_BETA] = rho != 0.0 ? rz / rho : 0.0;
    S.scal[SC_RHO] = rz;
    S.scal[SC_RR] = rr;
    S.scal[SC_ITER] += 1.0;
    const double maxrun = S.scal[SC_MAXRUN];
    if (rr <= S.scal[SC_TOL2] || rr < S.scal[SC_DELTA2] * S.scal[SC_MAXRR] || !(rz > 0.0) ||
        (maxrun > 0.0 && S.scal[SC_ITER] - S.scal[SC_LAST] >= maxrun))
      S.scal[SC_STATE] = ST_PENDING;
  }
}

// p = z + beta p (RUNNING, and PENDING systems finishing their iteration).
template <typename VT>
__global__ __launch_bounds__(kSB) void k_bs_pupd(const SysDev* __restrict__ sys, const int32_t* __restrict__ bsys) {
  const SysDev& S = sys[bsys[blockIdx.x]];
  const double st = S.scal[SC_STATE];
  if (st != ST_RUN && st != ST_PENDING) return;
  VT* __restrict__ pv = (VT*)gp(static_cast<VT*>(S.p));
  const VT* __restrict__ zv = (const VT*)gp(static_cast<const VT*>(S.z));
  const int lb = blockIdx.x - S.vblk0;
  const VT beta = (VT)S.scal[SC_BETA];
  for (int64_t j = (int64_t)lb * kVecNodes + threadIdx.x; j < S.N; j += (int64_t)S.nvblk * kVecNodes)
#pragma unroll
    for (int c = 0; c < 3; ++c) pv[3 * j + c] = zv[3 * j + c] + beta * pv[3 * j + c];
}

__global__ void k_bs_states(const SysDev* __restrict__ sys, int nsys, double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  out[2 * s] = sys[s].scal[SC_STATE];
  out[2 * s + 1] = sys[s].scal[SC_ITER];
}

// ---- single-system kernels of the reliable update ------------------------------

// trial solution xt = xh + x / s (xh is kept until xt is known to be finite); x = 0
template <typename VT>
__global__ void k_ru_trial(int64_t n, VT* __restrict__ x, double inv_s, const double* __restrict__ xh,
                           double* __restrict__ xt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xt[i] = xh[i] + (double)x[i] * inv_s;
  x[i] = (VT)0;
}

// y = mask(K64 x) (mask may be NULL)
__global__ __launch_bounds__(kSB) void k_spmv64(int64_t N, const int64_t* __restrict__ adj_ptr,
                                                const int32_t* __restrict__ adj, const double* __restrict__ val,
                                                const double* __restrict__ x, const uint8_t* __restrict__ fixed,
                                                double* __restrict__ y) {
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = gt / kSG;
  const int l = (int)(gt % kSG);
  double a0 = 0, a1 = 0, a2 = 0;
  if (j < N) {
    const int64_t p0 = adj_ptr[j];
    const int R = (int)(adj_ptr[j + 1] - p0);
    const double* P = val + 9 * p0;
    for (int t = l; t < R; t += kSG) {
      const int64_t i = adj[p0 + t];
      const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
      a0 += P[3 * t] * x0 + P[3 * t + 1] * x1 + P[3 * t + 2] * x2;
      a1 += P[3 * R + 3 * t] * x0 + P[3 * R + 3 * t + 1] * x1 + P[3 * R + 3 * t + 2] * x2;
      a2 += P[6 * R + 3 * t] * x0 + P[6 * R + 3 * t + 1] * x1 + P[6 * R + 3 * t + 2] * x2;
    }
  }
#pragma unroll
  for (int o = kSG / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kSG);
    a1 += __shfl_xor(a1, o, kSG);
    a2 += __shfl_xor(a2, o, kSG);
  }
  if (j < N && l == 0) {
    const double a[3] = {a0, a1, a2};
#pragma unroll
    for (int c = 0; c < 3; ++c) y[3 * j + c] = (fixed && fixed[3 * j + c]) ? 0.0 : a[c];
  }
}

// out = mask(b - y) (y may be NULL: out = mask(b)); partial sum of out^2 per block
__global__ __launch_bounds__(kSB) void k_ru_resid(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                                                  const uint8_t* __restrict__ fixed, double* __restrict__ out,
                                                  double* __restrict__ part) {
  __shared__ double sh[kSB];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    v = fixed[i] ? 0.0 : b[i] - (y ? y[i] : 0.0);
    if (out) out[i] = v;
  }
  const double s = bsum(v * v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// scal[SC_STATE] = v (one thread; an eager write that needs no host copy)
__global__ void k_set_state(double* __restrict__ scal, double v) { scal[SC_STATE] = v; }

__global__ __launch_bounds__(kSB) void k_sum(int64_t n, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[kSB];
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kSB) v += part[i];
  v = bsum(v, sh);
  if (threadIdx.x == 0) out[0] = v;
}

// r = rh * s; p = pscale * p (= z when init); z = Binv r; partials (r.r, r.z)
template <typename VT>
__global__ __launch_bounds__(kSB) void k_ru_restart(int64_t N, const double* __restrict__ rh, double s, double pscale,
                                                    int init, const float* __restrict__ binv, VT* __restrict__ r32,
                                                    VT* __restrict__ z32, VT* __restrict__ p32,
                                                    double* __restrict__ part) {
  __shared__ double sh[kSB];
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rz = 0.0;
  if (j < N) {
    VT r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[c] = (VT)(rh[3 * j + c] * s);
      r32[3 * j + c] = r[c];
    }
    const float* B = binv + 9 * j;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const VT z = (VT)B[3 * c] * r[0] + (VT)B[3 * c + 1] * r[1] + (VT)B[3 * c + 2] * r[2];
      z32[3 * j + c] = z;
      p32[3 * j + c] = init ? z : (VT)pscale * p32[3 * j + c];
      rr += (double)r[c] * (double)r[c];
      rz += (double)r[c] * (double)z;
    }
  }
  rr = bsum(rr, sh);
  rz = bsum(rz, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rr;
    part[2 * blockIdx.x + 1] = rz;
  }
}

__global__ __launch_bounds__(kSB) void k_ru_scal(int64_t nb, const double* __restrict__ part, double* __restrict__ scal,
                                                 double tol2, double delta2) {
  __shared__ double sh[kSB];
  double rr = 0.0, rz = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += kSB) {
    rr += part[2 * i];
    rz += part[2 * i + 1];
  }
  rr = bsum(rr, sh);
  rz = bsum(rz, sh);
  if (threadIdx.x == 0) {
    scal[SC_RHO] = rz;
    scal[SC_RR] = rr;
    scal[SC_MAXRR] = rr;
    scal[SC_TOL2] = tol2;
    scal[SC_DELTA2] = delta2;
    scal[SC_BETA] = 0.0;
    scal[SC_STATE] = ST_RUN;
  }
}

// AMG restart, first half: r = rh * s; p = pscale * p (not on the first
// start); partial r.r per engine vector block (the V-cycle adds r.z).
__global__ __launch_bounds__(kSB) void k_ru_restart_amg(int64_t N, int32_t nvblk, const double* __restrict__ rh,
                                                        double s, double pscale, int init, float* __restrict__ r32,
                                                        float* __restrict__ p32, double* __restrict__ part) {
  __shared__ double sh[kSB];
  double rr = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kVecNodes + threadIdx.x; j < N; j += (int64_t)nvblk * kVecNodes)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float r = (float)(rh[3 * j + c] * s);
      r32[3 * j + c] = r;
      if (!init) p32[3 * j + c] = (float)pscale * p32[3 * j + c];
      rr += (double)r * (double)r;
    }
  rr = bsum(rr, sh);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = rr;
}

// AMG restart without a V-cycle: r = rh * s, and p, z rescaled by pscale.
// z = M r carries over with the same scale only while the true residual is
// close to the rescaled recursive one; the pass also sums the drift
// ||r_true - pscale r_rec||^2 (into part[2 nvblk + block]) so the host can
// tell when z must be recomputed by a V-cycle instead.
__global__ __launch_bounds__(kSB) void k_ru_rescale_amg(int64_t N, int32_t nvblk, const double* __restrict__ rh,
                                                        double s, float pscale, float* __restrict__ r32,
                                                        float* __restrict__ z32, float* __restrict__ p32,
                                                        double* __restrict__ part) {
  __shared__ double sh[kSB];
  double rr = 0.0, rz = 0.0, dd = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kVecNodes + threadIdx.x; j < N; j += (int64_t)nvblk * kVecNodes)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float r = (float)(rh[3 * j + c] * s);
      const float z = pscale * z32[3 * j + c];
      const double dr = (double)r - (double)pscale * (double)r32[3 * j + c];
      r32[3 * j + c] = r;
      z32[3 * j + c] = z;
      p32[3 * j + c] = pscale * p32[3 * j + c];
      rr += (double)r * (double)r;
      rz += (double)r * (double)z;
      dd += dr * dr;
    }
  rr = bsum(rr, sh);
  rz = bsum(rz, sh);
  dd = bsum(dd, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rr;
    part[2 * blockIdx.x + 1] = rz;
    part[2 * nvblk + blockIdx.x] = dd;
  }
}

// AMG restart, second half: scalars from the partials (one block).
__global__ __launch_bounds__(kSB) void k_ru_scal_amg(int32_t nvblk, const double* __restrict__ part,
                                                     double* __restrict__ scal, double tol2, double delta2,
                                                     double maxrun) {
  __shared__ double sh[kSB];
  double rr = 0.0, rz = 0.0;
  for (int i = threadIdx.x; i < nvblk; i += kSB) {
    rr += part[2 * i];
    rz += part[2 * i + 1];
  }
  rr = bsum(rr, sh);
  rz = bsum(rz, sh);
  if (threadIdx.x == 0) {
    scal[SC_RHO] = rz;
    scal[SC_RR] = rr;
    scal[SC_MAXRR] = rr;
    scal[SC_TOL2] = tol2;
    scal[SC_DELTA2] = delta2;
    scal[SC_BETA] = 0.0;
    scal[SC_LAST] = scal[SC_ITER];
    scal[SC_MAXRUN] = maxrun;
    scal[SC_STATE] = ST_RUN;
  }
}

// Jacobi preconditioner, stored as 3x3 blocks: the inverse diagonal of K on
// the free components, zero elsewhere.  (Scalar Jacobi converged faster than
// the inverse 3x3 diagonal blocks on the elasticity systems measured; the
// block storage lets the update kernel take either.)
__global__ void k_binv(int64_t N, const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ dslot,
                       const double* __restrict__ val, const uint8_t* __restrict__ fixed, float* __restrict__ binv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int64_t a0 = adj_ptr[j];
  const int64_t R = adj_ptr[j + 1] - a0;
  const int64_t s = dslot[j];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double kdd = val[9 * a0 + d * 3 * R + 3 * s + d];   // K[(j,d),(j,d)] = panel[d][s][d]
    const bool free = fixed[3 * j + d] == 0 && kdd != 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) binv[9 * j + d * 3 + c] = (c == d && free) ? (float)(1.0 / kdd) : 0.f;
  }
}

// FP32 operator in block-contiguous order: out[(a0+t)*9 + c*3 + d] = panel
// entry [c][t][d] of node j (thread per node block).
__global__ void k_to32_blocks(int64_t N, const int64_t* __restrict__ adj_ptr, const double* __restrict__ val,
                              float* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
  if (j >= N) return;
  const int64_t a0 = adj_ptr[j];
  const int64_t R = adj_ptr[j + 1] - a0;
  const double* P = val + 9 * a0;
  for (int64_t t = threadIdx.x; t < R; t += blockDim.x)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) out[9 * (a0 + t) + 3 * c + d] = (float)P[c * 3 * R + 3 * t + d];
}

__global__ void k_fill_u(int64_t n, const uint8_t* __restrict__ fixed, const double* __restrict__ ufix,
                         const double* __restrict__ x, double* __restrict__ u) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u[i] = fixed[i] ? (ufix ? ufix[i] : 0.0) : (x ? x[i] : 0.0);
}

inline unsigned blocks(int64_t n, int bs = kSB) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }
inline int64_t spmv_blocks(int64_t N) {
  return std::max<int64_t>(1, std::min<int64_t>((N + kSpmvNodes - 1) / kSpmvNodes, kMaxSpmvBlocks));
}
inline int64_t vec_blocks(int64_t N) {
  return std::max<int64_t>(1, std::min<int64_t>((N + kVecNodes - 1) / kVecNodes, kMaxVecBlocks));
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct SolveSys {
  SysSpec spec;
  int64_t n = 0, nadj = 0;
  void* mem = nullptr;
  size_t bytes = 0;
  double *b = nullptr, *xh = nullptr, *xt = nullptr, *rh = nullptr, *y = nullptr, *scal = nullptr, *part = nullptr;
  UpdateBook book;          // rollbacks, consecutive non-finite corrections, updates
  int64_t refreshes = 0;    // AMG updates whose z had to be recomputed (drift or r.z <= 0)
  void *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;   // Krylov vectors (VT)
  float *binv = nullptr, *val32 = nullptr;
  int64_t npart = 0;
  double bnorm = 0.0, scale = 1.0, relres = 0.0;
  int64_t iters = 0;
  int status = kRunning;
  int32_t sblk0 = 0, nsblk = 0, vblk0 = 0, nvblk = 0;
  AmgSys* amg = nullptr;
};

BatchSolver::BatchSolver(llfem_ctx* c, bool f64_vectors) : ctx_(c), f64v_(f64_vectors) {}

BatchSolver::~BatchSolver() {
  drop_graph();
  if (hslot_) (void)hipHostFree(hslot_);
  for (SolveSys* s : sys_) release(s);
  sys_.clear();
  if (tables_) dfree(ctx_, tables_, tables_bytes_);
  amg_batch_free(ctx_, amgb_);
}

void BatchSolver::release(SolveSys* s) {
  amg_sys_free(ctx_, s->amg);
  if (s->mem) dfree(ctx_, s->mem, s->bytes);
  delete s;
}

size_t BatchSolver::bytes_needed(int64_t N, int64_t nadj, bool f64_vectors) {
  const int64_t n = 3 * N;
  const int64_t nsblk = spmv_blocks(N), nvblk = vec_blocks(N);
  const int64_t npart = std::max<int64_t>(nsblk + 3 * nvblk, blocks(n)) + 64;
  return sizeof(double) * (5 * (size_t)n + SC_N + (size_t)npart) +
         (f64_vectors ? sizeof(double) : sizeof(float)) * 5 * (size_t)n +
         sizeof(float) * (9 * (size_t)N + 9 * (size_t)nadj) + 16 * 256;   // 14 buffers, 256-B aligned
}

// Pinned host slots for the update readbacks (grown on demand).
int BatchSolver::host_slots(size_t n) {
  if (hslot_n_ >= n) return LLFEM_OK;
  if (hslot_) (void)hipHostFree(hslot_);
  hslot_ = nullptr;
  hslot_n_ = 0;
  LLFEM_HIP(hipHostMalloc((void**)&hslot_, sizeof(double) * n, hipHostMallocDefault));
  hslot_n_ = n;
  return LLFEM_OK;
}

// Reliable updates of several systems: fold x32 into xh, recompute the FP64
// residual, test convergence, restart the FP32 recursion.  The host waits a
// fixed number of times per call, whatever the number of systems:
//   1. every system's trial solution and true residual, one readback;
//   2. the decisions (engine_policy.h); rolled-back systems get the residual
//      of their kept solution, one more readback if there are any;
//   3. stop, or restart the recursion; AMG systems that keep a rescaled z
//      read r.z and the drift back, one readback for all of them;
//   4. those whose rescaled z fails the test get a V-cycle.
// Each system's launches are the same as for an update of that system alone.
int BatchSolver::update_many(const std::vector<SolveSys*>& ss, bool init) {
  if (ss.empty()) return LLFEM_OK;
  hipStream_t st = ctx_->stream;
  const size_t m = ss.size();
  LLFEM_TRY(host_slots(4 * m));   // per system: r^2, iterations, r.z, drift^2
  double* h = hslot_;
  // r = b - K x in FP64 for x (NULL: x = 0); ||r||^2 into scal[SC_N - 1]
  auto residual = [&](SolveSys* s, const double* x) {
    const SysSpec& sp = s->spec;
    const int64_t n = s->n;
    if (x) {
      k_spmv64<<<blocks(sp.N * kSG), kSB, 0, st>>>(sp.N, sp.adj_ptr, sp.adj, sp.val64, x, sp.fixed, s->y);
      k_ru_resid<<<blocks(n), kSB, 0, st>>>(n, s->b, s->y, sp.fixed, s->rh, s->part);
    } else {
      k_ru_resid<<<blocks(n), kSB, 0, st>>>(n, s->b, nullptr, sp.fixed, s->rh, s->part);
    }
    k_sum<<<1, kSB, 0, st>>>(blocks(n), s->part, s->scal + SC_N - 1);
  };
  // 1. trial solutions xt = xh + x / scale and their residuals
  for (size_t i = 0; i < m; ++i) {
    SolveSys* s = ss[i];
    const int64_t n = s->n;
    if (!init) {
      if (f64v_) k_ru_trial<double><<<blocks(n), kSB, 0, st>>>(n, (double*)s->x, 1.0 / s->scale, s->xh, s->xt);
      else k_ru_trial<float><<<blocks(n), kSB, 0, st>>>(n, (float*)s->x, 1.0 / s->scale, s->xh, s->xt);
      residual(s, s->xt);
    } else {
      residual(s, nullptr);
    }
    LLFEM_HIP(hipMemcpyAsync(h + 4 * i, s->scal + SC_N - 1, sizeof(double), hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipMemcpyAsync(h + 4 * i + 1, s->scal + SC_ITER, sizeof(double), hipMemcpyDeviceToHost, st));
  }
  LLFEM_HIP(hipStreamSynchronize(st));
  // 2. adoption or rollback.  LLFEM_ENGINE_FORCE_ROLLBACK=k treats the k-th
  // update's trial as not finite (a host-side switch for testing the
  // rollback path).
  std::vector<uint8_t> fresh(m, init ? 1 : 0);   // restart the Krylov recursion from scratch (p = z)
  if (!init) {
    const char* force_env = std::getenv("LLFEM_ENGINE_FORCE_ROLLBACK");
    const int64_t force_rb = force_env ? (int64_t)std::atoll(force_env) : 0;
    bool any = false;
    for (size_t i = 0; i < m; ++i) {
      SolveSys* s = ss[i];
      double r2 = h[4 * i];
      if (force_rb > 0 && s->book.updates + 1 == force_rb && s->book.rollbacks == 0) r2 = NAN;
      if (adopt_trial(r2, &s->book)) {
        std::swap(s->xh, s->xt);
      } else {
        // rollback: keep xh, restart the recursion from it
        fresh[i] = 1;
        any = true;
        residual(s, s->xh);
        LLFEM_HIP(hipMemcpyAsync(h + 4 * i, s->scal + SC_N - 1, sizeof(double), hipMemcpyDeviceToHost, st));
      }
    }
    if (any) LLFEM_HIP(hipStreamSynchronize(st));
  }
  // 3. stop or restart
  std::vector<uint8_t> check(m, 0);   // AMG systems whose rescaled z is tested
  std::vector<double> tol2(m, 0.0), s_new(m, 0.0), pscale(m, 0.0);
  bool any_check = false;
  for (size_t i = 0; i < m; ++i) {
    SolveSys* s = ss[i];
    const SysSpec& sp = s->spec;
    const int64_t n = s->n;
    const double rn = std::sqrt(h[4 * i]);
    s->iters = (int64_t)h[4 * i + 1];
    if (init) s->bnorm = rn;
    else ++s->book.updates;
    const int status = update_status(rn, s->bnorm, sp.reltol, s->iters, sp.maxiter, s->book);
    if (status != kRunning) {
      s->status = status;
      if (status != kConverged) s->relres = s->bnorm > 0 && std::isfinite(rn) ? rn / s->bnorm : 0.0;
      k_set_state<<<1, 1, 0, st>>>(s->scal, (double)ST_IDLE);
      continue;
    }
    s_new[i] = 1.0 / rn;
    pscale[i] = s_new[i] / s->scale;
    s->scale = s_new[i];
    const double tol = sp.reltol * s->bnorm * s_new[i];
    tol2[i] = tol * tol;
    if (s->amg) {
      const int32_t nsblk = (int32_t)spmv_blocks(sp.N), nvblk = (int32_t)vec_blocks(sp.N);
      double* vpart = s->part + nsblk;
      if (fresh[i]) {   // z = M r by a V-cycle of this system alone, p = z
        k_ru_restart_amg<<<(unsigned)nvblk, kSB, 0, st>>>(sp.N, nvblk, s->rh, s_new[i], pscale[i], 1, (float*)s->r,
                                                          (float*)s->p, vpart);
        LLFEM_TRY(amg_apply_solo(ctx_, s->amg));
        LLFEM_HIP(hipMemcpyAsync(s->p, s->z, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
      } else {
        k_ru_rescale_amg<<<(unsigned)nvblk, kSB, 0, st>>>(sp.N, nvblk, s->rh, s_new[i], (float)pscale[i],
                                                          (float*)s->r, (float*)s->z, (float*)s->p, vpart);
      }
      k_ru_scal_amg<<<1, kSB, 0, st>>>(nvblk, vpart, s->scal, tol2[i], kDeltaAmg * kDeltaAmg, (double)kMaxRunAmg);
      if (!fresh[i]) {
        // the rescaled z is consistent only if r.z > 0 and the true residual
        // is close to the rescaled recursive one (amg_keep_rescaled)
        k_sum<<<1, kSB, 0, st>>>(nvblk, vpart + 2 * nvblk, s->scal + SC_N - 1);
        LLFEM_HIP(hipMemcpyAsync(h + 4 * i + 2, s->scal + SC_RHO, sizeof(double), hipMemcpyDeviceToHost, st));
        LLFEM_HIP(hipMemcpyAsync(h + 4 * i + 3, s->scal + SC_N - 1, sizeof(double), hipMemcpyDeviceToHost, st));
        check[i] = 1;
        any_check = true;
      }
      continue;
    }
    const int64_t nb = blocks(sp.N);
    if (f64v_)
      k_ru_restart<double><<<(unsigned)nb, kSB, 0, st>>>(sp.N, s->rh, s_new[i], pscale[i], fresh[i] ? 1 : 0, s->binv,
                                                         (double*)s->r, (double*)s->z, (double*)s->p, s->part);
    else
      k_ru_restart<float><<<(unsigned)nb, kSB, 0, st>>>(sp.N, s->rh, s_new[i], pscale[i], fresh[i] ? 1 : 0, s->binv,
                                                        (float*)s->r, (float*)s->z, (float*)s->p, s->part);
    k_ru_scal<<<1, kSB, 0, st>>>(nb, s->part, s->scal, tol2[i], kDelta * kDelta);
  }
  // 4. AMG systems whose rescaled z failed: z = M r, restart the direction from it
  if (any_check) {
    LLFEM_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < m; ++i) {
      if (!check[i] || amg_keep_rescaled(h[4 * i + 2], h[4 * i + 3], kDriftAmg)) continue;
      SolveSys* s = ss[i];
      const SysSpec& sp = s->spec;
      const int32_t nsblk = (int32_t)spmv_blocks(sp.N), nvblk = (int32_t)vec_blocks(sp.N);
      double* vpart = s->part + nsblk;
      ++s->refreshes;
      k_ru_restart_amg<<<(unsigned)nvblk, kSB, 0, st>>>(sp.N, nvblk, s->rh, s_new[i], pscale[i], 1, (float*)s->r,
                                                        (float*)s->p, vpart);
      LLFEM_TRY(amg_apply_solo(ctx_, s->amg));
      LLFEM_HIP(hipMemcpyAsync(s->p, s->z, sizeof(float) * s->n, hipMemcpyDeviceToDevice, st));
      k_ru_scal_amg<<<1, kSB, 0, st>>>(nvblk, vpart, s->scal, tol2[i], kDeltaAmg * kDeltaAmg, (double)kMaxRunAmg);
    }
  }
  LLFEM_HIP(hipGetLastError());
  return LLFEM_OK;
}

int BatchSolver::add(const SysSpec& spec, SolveSys** out) {
  hipStream_t st = ctx_->stream;
  SolveSys* s = new SolveSys();
  s->spec = spec;
  s->n = 3 * spec.N;
  s->nadj = spec.nadj;
  s->bytes = bytes_needed(spec.N, spec.nadj, f64v_);
  int rc = dmalloc(ctx_, &s->mem, s->bytes);
  if (rc != LLFEM_OK) {
    delete s;
    return rc;
  }
  const int64_t n = s->n, N = spec.N;
  const int64_t nsblk = spmv_blocks(N), nvblk = vec_blocks(N);
  s->npart = std::max<int64_t>(nsblk + 3 * nvblk, blocks(n)) + 64;   // SpMV, V-cycle r.z and drift partials
  char* m = static_cast<char*>(s->mem);
  auto take = [&](size_t bytes) -> char* {
    char* p = m;
    m += (bytes + 255) / 256 * 256;
    return p;
  };
  s->b = (double*)take(sizeof(double) * n);
  s->xh = (double*)take(sizeof(double) * n);
  s->xt = (double*)take(sizeof(double) * n);
  s->rh = (double*)take(sizeof(double) * n);
  s->y = (double*)take(sizeof(double) * n);
  s->scal = (double*)take(sizeof(double) * SC_N);
  s->part = (double*)take(sizeof(double) * s->npart);
  const size_t vsz = f64v_ ? sizeof(double) : sizeof(float);
  s->x = take(vsz * n);
  s->r = take(vsz * n);
  s->z = take(vsz * n);
  s->p = take(vsz * n);
  s->q = take(vsz * n);
  s->binv = (float*)take(sizeof(float) * 9 * N);
  s->val32 = (float*)take(sizeof(float) * 9 * spec.nadj);
  if ((size_t)(m - static_cast<char*>(s->mem)) > s->bytes) {
    release(s);
    return fail(LLFEM_EINVAL, "engine: workspace layout overflow");
  }
  // b = mask(f - K u_fixed); x = 0; FP32 operator and block-Jacobi inverse
  k_fill_u<<<blocks(n), kSB, 0, st>>>(n, spec.fixed, spec.ufix, nullptr, s->y);
  k_spmv64<<<blocks(N * kSG), kSB, 0, st>>>(N, spec.adj_ptr, spec.adj, spec.val64, s->y, nullptr, s->rh);
  k_ru_resid<<<blocks(n), kSB, 0, st>>>(n, spec.f, s->rh, spec.fixed, s->b, s->part);
  k_to32_blocks<<<blocks(N, 8), dim3(32, 8), 0, st>>>(N, spec.adj_ptr, spec.val64, s->val32);
  k_binv<<<blocks(N), kSB, 0, st>>>(N, spec.adj_ptr, spec.dslot, spec.val64, spec.fixed, s->binv);
  LLFEM_HIP(hipMemsetAsync(s->xh, 0, sizeof(double) * n, st));
  LLFEM_HIP(hipMemsetAsync(s->x, 0, vsz * n, st));
  LLFEM_HIP(hipMemsetAsync(s->scal, 0, sizeof(double) * SC_N, st));
  s->scale = 1.0;
  if (spec.amg) {
    if (f64v_) {
      release(s);
      return fail(LLFEM_EUNSUP, "engine: the AMG preconditioner runs with FP32 Krylov vectors");
    }
    rc = amg_sys_create(ctx_, spec.amg, s->val32, (float*)s->r, (float*)s->z, (float*)s->q, s->part + nsblk, &s->amg);
    if (rc != LLFEM_OK) {
      release(s);
      return rc;
    }
  }
  rc = update_many({s}, true);
  if (rc != LLFEM_OK) {
    release(s);
    return rc;
  }
  sys_.push_back(s);
  dirty_ = true;
  *out = s;
  return LLFEM_OK;
}

int BatchSolver::rebuild_tables() {
  std::vector<SysDev> dev;
  std::vector<int32_t> sb, vb;
  live_.clear();
  for (SolveSys* s : sys_) {
    if (s->status != 0) continue;
    const SysSpec& sp = s->spec;
    SysDev d{};
    d.N = sp.N;
    d.adj_ptr = sp.adj_ptr;
    d.adj = sp.adj;
    d.val = s->val32;
    d.fixed = sp.fixed;
    d.binv = s->binv;
    d.x = s->x; d.r = s->r; d.z = s->z; d.p = s->p; d.q = s->q;
    d.scal = s->scal;
    d.part = s->part;
    d.nsblk = (int32_t)spmv_blocks(sp.N);
    d.nvblk = (int32_t)vec_blocks(sp.N);
    d.amg = s->amg ? 1 : 0;
    d.sblk0 = (int32_t)sb.size();
    d.vblk0 = (int32_t)vb.size();
    const int32_t id = (int32_t)dev.size();
    sb.insert(sb.end(), (size_t)d.nsblk, id);
    vb.insert(vb.end(), (size_t)d.nvblk, id);
    dev.push_back(d);
    live_.push_back(s);
  }
  {
    std::vector<AmgSys*> as;
    std::vector<const double*> st_words;
    for (SolveSys* s : live_)
      if (s->amg) {
        as.push_back(s->amg);
        st_words.push_back(s->scal + SC_STATE);
      }
    LLFEM_TRY(amg_batch_build(ctx_, as, st_words, &amgb_));
  }
  nsys_ = (int)dev.size();
  nsb_ = (int64_t)sb.size();
  nvb_ = (int64_t)vb.size();
  const size_t need = sizeof(SysDev) * dev.size() + sizeof(int32_t) * (sb.size() + vb.size()) +
                      sizeof(double) * 2 * dev.size() + 1024;
  if (tables_bytes_ < need) {
    if (tables_) dfree(ctx_, tables_, tables_bytes_);
    tables_ = nullptr;
    tables_bytes_ = 0;
    LLFEM_TRY(dmalloc(ctx_, &tables_, need * 2));
    tables_bytes_ = need * 2;
  }
  char* m = static_cast<char*>(tables_);
  d_sys_ = reinterpret_cast<SysDev*>(m);
  m += (sizeof(SysDev) * dev.size() + 255) / 256 * 256;
  d_sb_ = reinterpret_cast<int32_t*>(m);
  m += (sizeof(int32_t) * sb.size() + 255) / 256 * 256;
  d_vb_ = reinterpret_cast<int32_t*>(m);
  m += (sizeof(int32_t) * vb.size() + 255) / 256 * 256;
  d_states_ = reinterpret_cast<double*>(m);
  hipStream_t st = ctx_->stream;
  if (!dev.empty()) {
    LLFEM_HIP(hipMemcpyAsync(d_sys_, dev.data(), sizeof(SysDev) * dev.size(), hipMemcpyHostToDevice, st));
    LLFEM_HIP(hipMemcpyAsync(d_sb_, sb.data(), sizeof(int32_t) * sb.size(), hipMemcpyHostToDevice, st));
    LLFEM_HIP(hipMemcpyAsync(d_vb_, vb.data(), sizeof(int32_t) * vb.size(), hipMemcpyHostToDevice, st));
    LLFEM_HIP(hipStreamSynchronize(st));   // host vectors go out of scope
  }
  dirty_ = false;
  return LLFEM_OK;
}

void BatchSolver::drop_graph() {
  if (chunk_exec_) (void)hipGraphExecDestroy(chunk_exec_);
  chunk_exec_ = nullptr;
  chunk_iters_ = 0;
}

// Whether chunks run as a captured graph: on unless LLFEM_ENGINE_GRAPH=0.
static bool engine_graph_enabled() {
  const char* e = std::getenv("LLFEM_ENGINE_GRAPH");
  return !(e && e[0] == '0');
}

int BatchSolver::run(int iters, std::vector<SolveSys*>* finished) {
  hipStream_t st = ctx_->stream;
  if (dirty_) {
    drop_graph();
    LLFEM_TRY(rebuild_tables());
  }
  if (nsys_ == 0) return LLFEM_OK;
  const SysDev* sys = static_cast<const SysDev*>(d_sys_);
  if (engine_graph_enabled()) {
    // The chunk's launches (SpMV, scalars, updates, V-cycles) depend only on
    // the live set's tables; the per-system state they read and write lives
    // in device memory.  Capture them once per live set and replay.
    if (!chunk_exec_ || chunk_iters_ != iters) {
      drop_graph();
      hipGraph_t g = nullptr;
      LLFEM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = launch_chunk(iters);
      const hipError_t ec = hipStreamEndCapture(st, &g);
      LLFEM_TRY(rc);
      if (ec != hipSuccess) return fail(LLFEM_EHIP, std::string("engine graph capture: ") + hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&chunk_exec_, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        chunk_exec_ = nullptr;
        return fail(LLFEM_EHIP, std::string("engine graph instantiate: ") + hipGetErrorString(ei));
      }
      chunk_iters_ = iters;
    }
    LLFEM_HIP(hipGraphLaunch(chunk_exec_, st));
  } else {
    LLFEM_TRY(launch_chunk(iters));
  }
  k_bs_states<<<blocks(nsys_), kSB, 0, st>>>(sys, nsys_, d_states_);
  std::vector<double> h(2 * (size_t)nsys_);
  LLFEM_HIP(hipMemcpyAsync(h.data(), d_states_, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  LLFEM_HIP(hipGetLastError());
  std::vector<SolveSys*> paused;
  for (int i = 0; i < nsys_; ++i) {
    SolveSys* s = live_[i];
    s->iters = (int64_t)h[2 * i + 1];
    const int state = (int)h[2 * i];
    if (state == ST_PENDING || state == ST_PAUSED) paused.push_back(s);
  }
  LLFEM_TRY(update_many(paused, false));
  for (SolveSys* s : paused)
    if (s->status != kRunning) {
      dirty_ = true;
      if (finished) finished->push_back(s);
    }
  return LLFEM_OK;
}

int BatchSolver::launch_chunk(int iters) {
  hipStream_t st = ctx_->stream;
  const SysDev* sys = static_cast<const SysDev*>(d_sys_);
  for (int k = 0; k < iters; ++k) {
    if (f64v_) {
      k_bs_spmv<double><<<(unsigned)nsb_, kSB, 0, st>>>(sys, d_sb_);
      k_bs_alpha<<<(unsigned)nsys_, kSB, 0, st>>>(sys);
      k_bs_update<double><<<(unsigned)nvb_, kSB, 0, st>>>(sys, d_vb_);
      k_bs_beta<<<(unsigned)nsys_, kSB, 0, st>>>(sys);
      k_bs_pupd<double><<<(unsigned)nvb_, kSB, 0, st>>>(sys, d_vb_);
    } else {
      k_bs_spmv<float><<<(unsigned)nsb_, kSB, 0, st>>>(sys, d_sb_);
      k_bs_alpha<<<(unsigned)nsys_, kSB, 0, st>>>(sys);
      k_bs_update<float><<<(unsigned)nvb_, kSB, 0, st>>>(sys, d_vb_);
      LLFEM_TRY(amg_batch_apply(ctx_, amgb_));
      k_bs_beta<<<(unsigned)nsys_, kSB, 0, st>>>(sys);
      k_bs_pupd<float><<<(unsigned)nvb_, kSB, 0, st>>>(sys, d_vb_);
    }
  }
  return LLFEM_OK;
}

// u = u_fixed on constrained DoFs, xh elsewhere; true relative residual
// ||(K u - f)_free|| / ||f_free|| (test/test_dsl.jl:106).
int BatchSolver::finish(SolveSys* s, double* u) {
  hipStream_t st = ctx_->stream;
  const SysSpec& sp = s->spec;
  const int64_t n = s->n;
  k_fill_u<<<blocks(n), kSB, 0, st>>>(n, sp.fixed, sp.ufix, s->xh, u);
  k_spmv64<<<blocks(sp.N * kSG), kSB, 0, st>>>(sp.N, sp.adj_ptr, sp.adj, sp.val64, u, nullptr, s->y);
  k_ru_resid<<<blocks(n), kSB, 0, st>>>(n, sp.f, s->y, sp.fixed, nullptr, s->part);
  k_sum<<<1, kSB, 0, st>>>(blocks(n), s->part, s->scal + SC_N - 1);
  k_ru_resid<<<blocks(n), kSB, 0, st>>>(n, sp.f, nullptr, sp.fixed, nullptr, s->part);
  k_sum<<<1, kSB, 0, st>>>(blocks(n), s->part, s->scal + SC_N - 2);
  double h[2];
  LLFEM_HIP(hipMemcpyAsync(h, s->scal + SC_N - 2, sizeof(h), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  LLFEM_HIP(hipGetLastError());
  s->relres = h[0] > 0 ? std::sqrt(h[1] / h[0]) : std::sqrt(h[1]);
  return LLFEM_OK;
}

void BatchSolver::remove(SolveSys* s) {
  auto it = std::find(sys_.begin(), sys_.end(), s);
  if (it == sys_.end()) return;
  sys_.erase(it);
  (void)hipStreamSynchronize(ctx_->stream);
  release(s);
  dirty_ = true;
}

bool BatchSolver::any_running() const {
  for (SolveSys* s : sys_)
    if (s->status == 0) return true;
  return false;
}

int64_t sys_iterations(const SolveSys* s) { return s->iters; }
int64_t sys_updates(const SolveSys* s) { return s->book.updates; }
int64_t sys_rollbacks(const SolveSys* s) { return s->book.rollbacks; }
int64_t sys_refreshes(const SolveSys* s) { return s->refreshes; }
int sys_status(const SolveSys* s) { return s->status; }
double sys_relres(const SolveSys* s) { return s->relres; }

}  // namespace llfem
