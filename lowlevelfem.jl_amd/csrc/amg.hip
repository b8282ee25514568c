// Aggregation AMG preconditioner for the elasticity systems of the PCG engine.
//
// Reference: solveDisplacement(K, f; support, iterative=true,
// preconditioner=...) passes its `preconditioner` to cg as Pl
// (src/linear.jl:3711-3743, the call at :3736); the reference's default is
// Identity().  This file provides a multigrid Pl that stays on the device
// across requests on the same mesh and support.
//
// Method (plain aggregation AMG with rigid-body near-null space):
//   * Aggregates: greedy root + strong neighbours (block Frobenius norm
//     >= 0.25 x the row's largest off-diagonal block), then leftovers join
//     the neighbouring aggregate they couple to most.  Coarser levels
//     aggregate the aggregate graph itself.
//   * Tentative prolongator: per aggregate, the six rigid-body modes of its
//     nodes (constrained DoFs zeroed), orthonormalised by modified
//     Gram-Schmidt in FP64; the triangular factor is the next level's
//     near-null space.  Columns that vanish (e.g. aggregates of clamped
//     nodes) are dropped and their coarse DoFs masked.
//   * Coarse operators A_{l+1} = Q^T A_l Q (Galerkin), per request on the
//     device: each coarse 6x6 block sums its fine blocks in a fixed order.
//   * Smoother: Chebyshev (degree 2) on D^-1 A over [0.1, 1.1] x lambda,
//     lambda from 15 power steps; coarse correction scaled by 1.5; dense
//     inverse on the coarsest level (at most 24 aggregates).
//
// Everything a level needs lives in 3-DoF "pseudo-node" block form: a
// coarse aggregate's six DoFs are two pseudo-nodes, so every level uses the
// same block-sparse layout as the fine operator ([slot][c][d] FP32 blocks,
// node-major adjacency).  The V-cycle kernels run table-driven over all
// systems of a batch, one launch per kernel and level.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <future>
#include <numeric>
#include <thread>
#include <vector>

#include "common.h"

namespace llfem {

namespace {

constexpr int kMB = 256;                 // threads per block
constexpr int kMG = 8;                   // lanes per pseudo-node row in the level SpMV
constexpr int kMSpN = kMB / kMG;         // pseudo-nodes per SpMV block
// The level SpMV needs no per-block partials, so its grid is not capped (a
// capped grid-stride version was 6 % slower on the 6.3M-DoF beam); the
// vector and restriction kernels cap their blocks and stride.
constexpr int64_t kMaxSB = int64_t(1) << 24, kMaxVB = 1024, kMaxAB = 1024;
constexpr int kCheb = 2;                 // Chebyshev degree (pre and post)
constexpr double kChebLo = 0.1, kChebHi = 1.1;
constexpr float kOverCorrect = 1.5f;
constexpr double kStrength = 0.25;
constexpr int kPowerSteps = 15;
constexpr int64_t kCoarseAggs = 24;      // coarsest level: at most this many aggregates
constexpr int kMaxLevels = 12;
constexpr int64_t kMaxDense = 1536;      // DoFs of the dense coarsest solve

inline unsigned nblk(int64_t n, int bs = kMB) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }
inline int64_t cap(int64_t n, int64_t per, int64_t mx) { return std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, mx)); }

__device__ __forceinline__ double msum(double v, double* sh) {
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

}  // namespace

// One (system, level) entry of a V-cycle launch.
struct MgEnt {
  int64_t N;                 // pseudo-nodes of the level (3 DoFs each)
  const int64_t* adj_ptr;
  const int32_t* adj;
  const float* val;          // [nadj][c][d]
  const uint8_t* mask;       // [3N] 1 = no DoF (constrained / dropped)
  const float* dinv;         // [3N] inverse diagonal, 0 on masked DoFs
  float* b;                  // right-hand side of the level
  float* x;                  // iterate of the level
  float* y;                  // A x
  float* d;                  // Chebyshev direction
  const int32_t* agg;        // [N] aggregate of each pseudo-node (-1: none)
  const float* Q;            // [N][3][6] tentative prolongator rows
  const int32_t* mem_ptr;    // [nagg+1] members of each aggregate
  const int32_t* mem;
  int64_t nagg;
  float* bc;                 // next level's b
  const float* xc;           // next level's x
  const float* Ainv;         // coarsest: dense inverse [nc][nc]
  int32_t nc;
  float inv_theta, inv_delta;
  const double* state;       // system state word: run only while it reads 0 (NULL: always)
  double* rzpart;            // level 0: partial r.z of vector block lb at rzpart[2 lb + 1]
  int32_t sblk0, nsblk, vblk0, nvblk, ablk0, nablk;
};

namespace {

__device__ __forceinline__ bool mg_skip(const MgEnt& E) { return E.state && *E.state != 0.0; }


// y = mask(A x) for every entry's level
__global__ __launch_bounds__(kMB) void k_mg_spmv(const MgEnt* __restrict__ ents, const int32_t* __restrict__ map) {
  const MgEnt& E = ents[map[blockIdx.x]];
  if (mg_skip(E)) return;
  // local restrict-qualified copies: the output never aliases the inputs
  const int64_t N = E.N;
  const auto* __restrict__ adj_ptr = gp(E.adj_ptr);
  const auto* __restrict__ adj = gp(E.adj);
  const auto* __restrict__ val = gp(E.val);
  const auto* __restrict__ x = gp(E.x);
  const auto* __restrict__ mask = gp(E.mask);
  auto* __restrict__ y = gp(E.y);
  const int lb = blockIdx.x - E.sblk0;
  const int nsblk = E.nsblk;
  const int l = threadIdx.x % kMG;
  for (int64_t j = (int64_t)lb * kMSpN + threadIdx.x / kMG; j - threadIdx.x / kMG < N;
       j += (int64_t)nsblk * kMSpN) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    if (j < N) {
      const int64_t p0 = adj_ptr[j];
      const int R = (int)(adj_ptr[j + 1] - p0);
      const auto* P = val + 9 * p0;
#pragma unroll 2
      for (int t = l; t < R; t += kMG) {
        const int64_t i = adj[p0 + t];
        const float x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        const auto* B = P + 9 * t;
        a0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
        a1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
        a2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
      }
    }
#pragma unroll
    for (int o = kMG / 2; o > 0; o >>= 1) {
      a0 += __shfl_xor(a0, o, kMG);
      a1 += __shfl_xor(a1, o, kMG);
      a2 += __shfl_xor(a2, o, kMG);
    }
    if (j < N && l == 0) {
      y[3 * j] = mask[3 * j] ? 0.f : a0;
      y[3 * j + 1] = mask[3 * j + 1] ? 0.f : a1;
      y[3 * j + 2] = mask[3 * j + 2] ? 0.f : a2;
    }
  }
}

// first Chebyshev step from x = 0: x = d = D^-1 b / theta
__global__ __launch_bounds__(kMB) void k_mg_cheb0(const MgEnt* __restrict__ ents, const int32_t* __restrict__ map) {
  const MgEnt& E = ents[map[blockIdx.x]];
  if (mg_skip(E)) return;
  const int lb = blockIdx.x - E.vblk0;
  const float* __restrict__ dinv = (const float*)gp(E.dinv);
  const float* __restrict__ b = (const float*)gp(E.b);
  float* __restrict__ x = (float*)gp(E.x);
  float* __restrict__ d = (float*)gp(E.d);
  const float it = E.inv_theta;
  for (int64_t i = (int64_t)lb * kMB + threadIdx.x; i < 3 * E.N; i += (int64_t)E.nvblk * kMB) {
    const float v = it * dinv[i] * b[i];
    x[i] = v;
    d[i] = v;
  }
}

// Chebyshev step: d = ca d + (cbt / theta + cbd / delta) D^-1 (b - y); x += d.
// With rz set (level 0, last step): partial b.x of the block (b = r, x = z).
__global__ __launch_bounds__(kMB) void k_mg_cheb(const MgEnt* __restrict__ ents, const int32_t* __restrict__ map,
                                                 float ca, float cbt, float cbd, int rz) {
  __shared__ double sh[kMB];
  const MgEnt& E = ents[map[blockIdx.x]];
  if (mg_skip(E)) return;
  const int lb = blockIdx.x - E.vblk0;
  const float cb = cbt * E.inv_theta + cbd * E.inv_delta;
  const float* __restrict__ b = (const float*)gp(E.b);
  const float* __restrict__ y = (const float*)gp(E.y);
  const float* __restrict__ dinv = (const float*)gp(E.dinv);
  float* __restrict__ d = (float*)gp(E.d);
  float* __restrict__ x = (float*)gp(E.x);
  const int64_t n = 3 * E.N;
  double acc = 0.0;
  for (int64_t i = (int64_t)lb * kMB + threadIdx.x; i < n; i += (int64_t)E.nvblk * kMB) {
    const float bi = b[i];
    const float dv = ca * d[i] + cb * dinv[i] * (bi - y[i]);
    const float xv = x[i] + dv;
    d[i] = dv;
    x[i] = xv;
    acc += (double)bi * (double)xv;
  }
  if (rz) {
    acc = msum(acc, sh);
    if (threadIdx.x == 0) ((double*)gp(E.rzpart))[2 * lb + 1] = acc;
  }
}

// bc = Q^T (b - y): kMG lanes per aggregate stride over its members, then a
// fixed shuffle tree sums the lanes.
__global__ __launch_bounds__(kMB) void k_mg_restrict(const MgEnt* __restrict__ ents, const int32_t* __restrict__ map) {
  const MgEnt& E = ents[map[blockIdx.x]];
  if (mg_skip(E)) return;
  const int lb = blockIdx.x - E.ablk0;
  const int l = threadIdx.x % kMG;
  const int32_t* __restrict__ mem_ptr = (const int32_t*)gp(E.mem_ptr);
  const int32_t* __restrict__ mem = (const int32_t*)gp(E.mem);
  const float* __restrict__ Q = (const float*)gp(E.Q);
  const float* __restrict__ b = (const float*)gp(E.b);
  const float* __restrict__ y = (const float*)gp(E.y);
  float* __restrict__ bc = (float*)gp(E.bc);
  for (int64_t a = (int64_t)lb * kMSpN + threadIdx.x / kMG; a - threadIdx.x / kMG < E.nagg;
       a += (int64_t)E.nablk * kMSpN) {
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (a < E.nagg) {
      for (int32_t m = mem_ptr[a] + l; m < mem_ptr[a + 1]; m += kMG) {
        const int64_t j = mem[m];
        const float* Qj = Q + 18 * j;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float r = b[3 * j + c] - y[3 * j + c];
#pragma unroll
          for (int u = 0; u < 6; ++u) acc[u] += Qj[6 * c + u] * r;
        }
      }
    }
#pragma unroll
    for (int o = kMG / 2; o > 0; o >>= 1)
#pragma unroll
      for (int u = 0; u < 6; ++u) acc[u] += __shfl_xor(acc[u], o, kMG);
    if (a < E.nagg && l == 0)
#pragma unroll
      for (int u = 0; u < 6; ++u) bc[6 * a + u] = acc[u];
  }
}

// x += alpha Q xc
__global__ __launch_bounds__(kMB) void k_mg_prolong(const MgEnt* __restrict__ ents, const int32_t* __restrict__ map,
                                                    float alpha) {
  const MgEnt& E = ents[map[blockIdx.x]];
  if (mg_skip(E)) return;
  const int lb = blockIdx.x - E.vblk0;
  const int32_t* __restrict__ agg = (const int32_t*)gp(E.agg);
  const float* __restrict__ Q = (const float*)gp(E.Q);
  const float* __restrict__ xcv = (const float*)gp(E.xc);
  float* __restrict__ x = (float*)gp(E.x);
  for (int64_t j = (int64_t)lb * kMB + threadIdx.x; j < E.N; j += (int64_t)E.nvblk * kMB) {
    const int32_t a = agg[j];
    if (a < 0) continue;
    const float* Qj = Q + 18 * j;
    const float* xc = xcv + 6 * (int64_t)a;
    float v[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) v[u] = xc[u];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float s = 0.f;
#pragma unroll
      for (int u = 0; u < 6; ++u) s += Qj[6 * c + u] * v[u];
      x[3 * j + c] += alpha * s;
    }
  }
}

// coarsest level: x = Ainv b (block per entry)
__global__ __launch_bounds__(kMB) void k_mg_coarse(const MgEnt* __restrict__ ents) {
  const MgEnt& E = ents[blockIdx.x];
  if (mg_skip(E)) return;
  const int n = E.nc;
  const float* __restrict__ Ainv = (const float*)gp(E.Ainv);
  const float* __restrict__ b = (const float*)gp(E.b);
  float* __restrict__ x = (float*)gp(E.x);
  for (int i = threadIdx.x; i < n; i += kMB) {
    const float* row = Ainv + (int64_t)i * n;
    float s = 0.f;
    for (int k = 0; k < n; ++k) s += row[k] * b[k];
    x[i] = s;
  }
}

// ---- per-request numeric set-up -------------------------------------------------

// Galerkin product: thread per (coarse aggregate pair p, coarse row u).
__global__ __launch_bounds__(kMB) void k_galerkin(int64_t npairs, const int64_t* __restrict__ gptr,
                                                  const int32_t* __restrict__ gblk, const int32_t* __restrict__ gcol,
                                                  const int32_t* __restrict__ adjf, const float* __restrict__ Qf,
                                                  const float* __restrict__ valf, const int32_t* __restrict__ pair_o,
                                                  float* __restrict__ valc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t p = t / 6;
  const int u = (int)(t % 6);
  if (p >= npairs) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t c = gptr[p]; c < gptr[p + 1]; ++c) {
    const int64_t g = gblk[c];
    const int64_t j = gcol[c];
    const int64_t i = adjf[g];
    const float* B = valf + 9 * g;      // B[3c+d] = A[(j,c),(i,d)]
    const float* Qj = Qf + 18 * j;
    const float* Qi = Qf + 18 * i;
    const double q0 = Qj[u], q1 = Qj[6 + u], q2 = Qj[12 + u];
    const double t0 = q0 * B[0] + q1 * B[3] + q2 * B[6];
    const double t1 = q0 * B[1] + q1 * B[4] + q2 * B[7];
    const double t2 = q0 * B[2] + q1 * B[5] + q2 * B[8];
#pragma unroll
    for (int v = 0; v < 6; ++v) acc[v] += t0 * Qi[v] + t1 * Qi[6 + v] + t2 * Qi[12 + v];
  }
  const int s = u / 3, cc = u % 3;
  const int64_t o = pair_o[2 * p + s];
#pragma unroll
  for (int v = 0; v < 6; ++v) valc[9 * (o + v / 3) + 3 * cc + v % 3] = (float)acc[v];
}

__global__ void k_mg_dinv(int64_t N, const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ dslot,
                          const float* __restrict__ val, const uint8_t* __restrict__ mask, float* __restrict__ dinv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const float* B = val + 9 * (adj_ptr[j] + dslot[j]);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float v = B[4 * d];
    dinv[3 * j + d] = (mask[3 * j + d] || !(v > 0.f)) ? 0.f : 1.f / v;
  }
}

__global__ void k_mg_rand(int64_t n, const float* __restrict__ dinv, float* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t h = (uint32_t)i * 2654435761u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  v[i] = dinv[i] > 0.f ? (float)(h & 0xFFFFu) / 65536.f - 0.5f : 0.f;
}

__global__ __launch_bounds__(kMB) void k_mg_spmv1(int64_t N, const int64_t* __restrict__ adj_ptr,
                                                  const int32_t* __restrict__ adj, const float* __restrict__ val,
                                                  const uint8_t* __restrict__ mask, const float* __restrict__ x,
                                                  float* __restrict__ y) {
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = gt / kMG;
  const int l = (int)(gt % kMG);
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  if (j < N) {
    const int64_t p0 = adj_ptr[j];
    const int R = (int)(adj_ptr[j + 1] - p0);
    for (int t = l; t < R; t += kMG) {
      const int64_t i = adj[p0 + t];
      const float* B = val + 9 * (p0 + t);
      a0 += B[0] * x[3 * i] + B[1] * x[3 * i + 1] + B[2] * x[3 * i + 2];
      a1 += B[3] * x[3 * i] + B[4] * x[3 * i + 1] + B[5] * x[3 * i + 2];
      a2 += B[6] * x[3 * i] + B[7] * x[3 * i + 1] + B[8] * x[3 * i + 2];
    }
  }
#pragma unroll
  for (int o = kMG / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kMG);
    a1 += __shfl_xor(a1, o, kMG);
    a2 += __shfl_xor(a2, o, kMG);
  }
  if (j < N && l == 0) {
    y[3 * j] = mask[3 * j] ? 0.f : a0;
    y[3 * j + 1] = mask[3 * j + 1] ? 0.f : a1;
    y[3 * j + 2] = mask[3 * j + 2] ? 0.f : a2;
  }
}

__global__ void k_mg_dscale(int64_t n, const float* __restrict__ dinv, const float* __restrict__ y,
                            float* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = dinv[i] * y[i];
}

// partial (v.y, v.D v) per block
__global__ __launch_bounds__(kMB) void k_mg_rq(int64_t n, const float* __restrict__ v, const float* __restrict__ y,
                                               const float* __restrict__ dinv, double* __restrict__ part) {
  __shared__ double sh[kMB];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (i < n && dinv[i] > 0.f) {
    a = (double)v[i] * y[i];
    b = (double)v[i] * v[i] / dinv[i];
  }
  a = msum(a, sh);
  b = msum(b, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

__global__ __launch_bounds__(kMB) void k_mg_sum2(int64_t nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[kMB];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += kMB) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  a = msum(a, sh);
  b = msum(b, sh);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

// per block: the largest off-diagonal block norm of node j, then strong flags
__global__ void k_mg_strength(int64_t N, const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ dslot,
                              const double* __restrict__ val64, double theta, uint8_t* __restrict__ strong) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int64_t a0 = adj_ptr[j];
  const int64_t R = adj_ptr[j + 1] - a0;
  const double* P = val64 + 9 * a0;    // panel [c][t][d]
  const int64_t self = dslot[j];
  double mx = 0.0;
  for (int64_t t = 0; t < R; ++t) {
    if (t == self) continue;
    double s = 0.0;
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        const double v = P[c * 3 * R + 3 * t + d];
        s += v * v;
      }
    mx = fmax(mx, s);
  }
  const double thr = theta * theta * mx;
  for (int64_t t = 0; t < R; ++t) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        const double v = P[c * 3 * R + 3 * t + d];
        s += v * v;
      }
    strong[a0 + t] = (t != self && s > 0.0 && s >= thr) ? 1 : 0;
  }
}

// ---- host hierarchy -------------------------------------------------------------

}  // namespace

// Host allocator for the large arrays of the hierarchy build: resize() leaves
// the elements uninitialised, so the first touch of a large array happens
// where it is filled -- in parallel threads -- rather than in a serial zero
// fill.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  using value_type = T;
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
template <class T>
using HVec = std::vector<T, NoInitAlloc<T>>;

namespace {

struct HostLevel {
  int64_t N = 0;                   // pseudo-nodes
  std::vector<int64_t> adj_ptr;    // level >= 1 only (level 0: the mesh's)
  std::vector<int32_t> adj, dslot;
  std::vector<uint8_t> mask;       // level >= 1 only
  // transfer (all but the coarsest)
  int64_t nagg = 0;
  std::vector<int32_t> agg;        // [N]
  HVec<float> Q;                   // [N][18]
  std::vector<int32_t> mem_ptr, mem;
  std::vector<int32_t> pair_o;     // [npairs][2]
  std::vector<int64_t> gptr;
  HVec<int32_t> gblk, gcol;
};

// Run f(lo, hi) over [0, n) split across host threads; serial below `grain`
// items (per-aggregate loops pass a small grain: on coarse levels there are
// few aggregates, each with many members).
template <class F>
void parallel_for(int64_t n, F&& f, int64_t grain = 4096) {
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), 16));
  if (n < grain || nt == 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t step = (n + nt - 1) / nt;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t lo = t * step, hi = std::min(n, lo + step);
    if (lo < hi) th.emplace_back([&f, lo, hi]() { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// A graph over real nodes: the neighbours of i are idx[e], e in
// [ptr[i], ptr[i+1]), except i itself and, with `flag`, entries flagged 0.
struct Graph {
  const int64_t* ptr = nullptr;
  const int32_t* idx = nullptr;
  const uint8_t* flag = nullptr;
  template <class F>
  void each(int64_t i, F&& f) const {
    for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
      const int32_t j = idx[e];
      if (j == i || (flag && !flag[e])) continue;
      f(j);
    }
  }
};

// The aggregate the neighbours of i (through `from`) belong to most often;
// ties to the smallest id; -1 if none.
int32_t most_coupled(const Graph& g, int64_t i, const std::vector<int32_t>& from, std::vector<int32_t>& buf) {
  buf.clear();
  g.each(i, [&](int32_t j) {
    if (from[j] >= 0) buf.push_back(from[j]);
  });
  if (buf.empty()) return -1;
  std::sort(buf.begin(), buf.end());
  int32_t best = -1;
  size_t bc = 0;
  for (size_t k = 0; k < buf.size();) {
    size_t e = k;
    while (e < buf.size() && buf[e] == buf[k]) ++e;
    if (e - k > bc) {
      bc = e - k;
      best = buf[k];
    }
    k = e;
  }
  return best;
}

// Greedy aggregation over real nodes.  Returns the aggregate count; agg = -1
// for nodes with no neighbour at all.
int64_t aggregate(int64_t n, const Graph& S, const Graph& F, std::vector<int32_t>& agg) {
  agg.assign((size_t)n, -1);
  int64_t na = 0;
  for (int64_t i = 0; i < n; ++i) {   // phase 1: root + all strong neighbours, all unaggregated
    if (agg[i] >= 0) continue;
    bool ok = true, any = false;
    for (int64_t e = S.ptr[i]; e < S.ptr[i + 1]; ++e) {   // stops at the first aggregated one
      const int32_t j = S.idx[e];
      if (j == i || (S.flag && !S.flag[e])) continue;
      any = true;
      if (agg[j] >= 0) {
        ok = false;
        break;
      }
    }
    if (!ok || !any) continue;
    agg[i] = (int32_t)na;
    S.each(i, [&](int32_t j) { agg[j] = (int32_t)na; });
    ++na;
  }
  // phase 2: join the phase-1 aggregate coupled most strongly.  It reads the
  // phase-1 snapshot only, so the nodes are independent (host threads).
  std::vector<int32_t> snap = agg, buf;
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> b;
    for (int64_t i = lo; i < hi; ++i)
      if (snap[i] < 0) agg[i] = most_coupled(S, i, snap, b);
  });
  for (int64_t i = 0; i < n; ++i) {   // phase 3a: leftovers with unaggregated strong neighbours
    if (agg[i] >= 0) continue;
    bool any = false;
    S.each(i, [&](int32_t j) { any = any || agg[j] < 0; });
    if (!any) continue;
    agg[i] = (int32_t)na;
    S.each(i, [&](int32_t j) {
      if (agg[j] < 0) agg[j] = (int32_t)na;
    });
    ++na;
  }
  for (int64_t i = 0; i < n; ++i) {   // phase 3b: weakly coupled nodes join, isolated ones stay out
    if (agg[i] >= 0) continue;
    const int32_t a = most_coupled(F, i, agg, buf);
    if (a >= 0) {
      agg[i] = a;
    } else {
      bool any = false;
      F.each(i, [&](int32_t) { any = true; });
      if (any) agg[i] = (int32_t)na++;
    }
  }
  return na;
}


// Modified Gram-Schmidt (two passes) of a rows x 6 matrix (row-major) in
// place; R (6x6, row-major) with B = Q R; dropped columns are zeroed in Q and
// their rows in R.
void mgs6(int64_t rows, double* M, double* R, bool* dropped) {
  std::fill(R, R + 36, 0.0);
  for (int u = 0; u < 6; ++u) {
    double n0 = 0.0;
    for (int64_t r = 0; r < rows; ++r) n0 += M[6 * r + u] * M[6 * r + u];
    n0 = std::sqrt(n0);
    for (int pass = 0; pass < 2; ++pass)
      for (int w = 0; w < u; ++w) {
        if (dropped[w]) continue;
        double c = 0.0;
        for (int64_t r = 0; r < rows; ++r) c += M[6 * r + w] * M[6 * r + u];
        for (int64_t r = 0; r < rows; ++r) M[6 * r + u] -= c * M[6 * r + w];
        R[6 * w + u] += c;
      }
    double nv = 0.0;
    for (int64_t r = 0; r < rows; ++r) nv += M[6 * r + u] * M[6 * r + u];
    nv = std::sqrt(nv);
    if (nv > 1e-10 * n0 && nv > 0.0) {
      dropped[u] = false;
      R[6 * u + u] = nv;
      for (int64_t r = 0; r < rows; ++r) M[6 * r + u] /= nv;
    } else {
      dropped[u] = true;
      for (int v = 0; v < 6; ++v) R[6 * u + v] = 0.0;
      for (int64_t r = 0; r < rows; ++r) M[6 * r + u] = 0.0;
    }
  }
}

}  // namespace

struct AmgHost {
  std::vector<HostLevel> lev;      // lev.back(): coarsest
  int64_t N0 = 0;
  double ms = 0.0;
};

// Input of the host build, gathered from the device by amg_gather.
struct AmgInput {
  int64_t N = 0;
  std::vector<int64_t> adj_ptr;
  HVec<int32_t> adj;
  std::vector<int32_t> dslot;
  HVec<uint8_t> strong;            // per block
  std::vector<double> xyz;         // SoA
  std::vector<uint8_t> mask;       // [3N]
};

struct AmgHier {
  struct Lev {
    int64_t N = 0, nadj = 0, nagg = 0, npairs = 0;
    const int64_t* adj_ptr = nullptr;
    const int32_t* adj = nullptr;
    const int32_t* dslot = nullptr;
    const uint8_t* mask = nullptr;
    const int32_t* agg = nullptr;
    const float* Q = nullptr;
    const int32_t* mem_ptr = nullptr;
    const int32_t* mem = nullptr;
    const int32_t* pair_o = nullptr;
    const int64_t* gptr = nullptr;
    const int32_t* gblk = nullptr;
    const int32_t* gcol = nullptr;
  };
  std::vector<Lev> lev;
  std::vector<int64_t> c_adj_ptr;  // host copy of the coarsest pattern
  std::vector<int32_t> c_adj;
  std::vector<uint8_t> c_mask;
  std::vector<std::pair<void*, size_t>> owned;
  size_t bytes = 0;
  double host_ms = 0.0;
};

namespace {

// Build level l+1 from level l.  Real nodes map to `ps` pseudo-nodes each
// (level 0: one; coarser: two).  nullrow(j, c, out6): near-null-space row of
// pseudo-node j, component c.  S, F: strong and full real-node graphs.
// Outputs the coarse level C, its near-null space Bc ([na][6][6]) and the
// aggregate graph (nbp, nbi; each list includes the aggregate itself).
template <class NullRow>
void coarsen(HostLevel& L, const int64_t* adj_ptr, const int32_t* adj, int ps, int64_t n_real, NullRow&& nullrow,
             const Graph& S, const Graph& F, HostLevel& C, std::vector<double>& Bc, std::vector<int64_t>& nbp,
             std::vector<int32_t>& nbi) {
  using clk = std::chrono::steady_clock;
  const bool trace = std::getenv("LLFEM_AMG_TRACE") != nullptr;
  auto tick = [&](const char* what, clk::time_point& t) {
    if (!trace) return;
    const auto now = clk::now();
    std::fprintf(stderr, "[amg]   N %lld %s %.1f ms\n", (long long)L.N, what,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  };
  auto tp = clk::now();
  std::vector<int32_t> ragg;
  const int64_t na = aggregate(n_real, S, F, ragg);
  tick("aggregate", tp);
  L.nagg = na;
  L.agg.assign((size_t)L.N, -1);
  for (int64_t a = 0; a < n_real; ++a)
    for (int s = 0; s < ps; ++s) L.agg[ps * a + s] = ragg[a];
  // members (pseudo-nodes, ascending)
  L.mem_ptr.assign((size_t)na + 1, 0);
  for (int64_t j = 0; j < L.N; ++j)
    if (L.agg[j] >= 0) L.mem_ptr[L.agg[j] + 1]++;
  for (int64_t a = 0; a < na; ++a) L.mem_ptr[a + 1] += L.mem_ptr[a];
  L.mem.assign((size_t)L.mem_ptr[na], 0);
  {
    std::vector<int32_t> fill(L.mem_ptr.begin(), L.mem_ptr.end() - 1);
    for (int64_t j = 0; j < L.N; ++j)
      if (L.agg[j] >= 0) L.mem[fill[L.agg[j]]++] = (int32_t)j;
  }
  tick("members", tp);
  // tentative prolongator and coarse near-null space (aggregates in parallel)
  L.Q.resize((size_t)L.N * 18);
  parallel_for(L.N, [&](int64_t lo, int64_t hi) { std::fill(L.Q.begin() + 18 * lo, L.Q.begin() + 18 * hi, 0.f); });
  Bc.assign((size_t)na * 36, 0.0);
  C.N = 2 * na;
  C.mask.assign((size_t)C.N * 3, 0);
  parallel_for(na, [&](int64_t lo, int64_t hi) {
    std::vector<double> M;
    for (int64_t a = lo; a < hi; ++a) {
      const int64_t m0 = L.mem_ptr[a], m1 = L.mem_ptr[a + 1];
      const int64_t rows = 3 * (m1 - m0);
      M.assign((size_t)rows * 6, 0.0);
      for (int64_t m = m0; m < m1; ++m)
        for (int c = 0; c < 3; ++c) nullrow(L.mem[m], c, &M[6 * (3 * (m - m0) + c)]);
      double R[36];
      bool dropped[6];
      mgs6(rows, M.data(), R, dropped);
      for (int64_t m = m0; m < m1; ++m) {
        const int64_t j = L.mem[m];
        for (int c = 0; c < 3; ++c)
          for (int u = 0; u < 6; ++u) L.Q[18 * j + 6 * c + u] = (float)M[6 * (3 * (m - m0) + c) + u];
      }
      for (int k = 0; k < 36; ++k) Bc[36 * a + k] = R[k];
      for (int u = 0; u < 6; ++u) C.mask[6 * a + u] = dropped[u] ? 1 : 0;
    }
  }, 64);
  tick("prolongator", tp);
  // per aggregate: neighbour aggregates (every fine block couples), then the
  // Galerkin contribution lists of its pairs, fine blocks in ascending order.
  // Two passes over the aggregates in parallel (sizes, then contents at
  // offsets from prefix sums): each aggregate's lists depend on it alone, so
  // the result does not depend on the thread count.
  nbp.assign((size_t)na + 1, 0);
  std::vector<int64_t> nblk((size_t)na + 1, 0);
  // neighbour aggregates of a, ascending, into nb (stamp: per-thread scratch)
  auto neighbours = [&](int64_t a, std::vector<int32_t>& stamp, std::vector<int32_t>& nb) {
    nb.clear();
    stamp[a] = (int32_t)a;
    nb.push_back((int32_t)a);
    for (int64_t m = L.mem_ptr[a]; m < L.mem_ptr[a + 1]; ++m) {
      const int64_t j = L.mem[m];
      for (int64_t g = adj_ptr[j]; g < adj_ptr[j + 1]; ++g) {
        const int32_t B = L.agg[adj[g]];
        if (B >= 0 && stamp[B] != a) {
          stamp[B] = (int32_t)a;
          nb.push_back(B);
        }
      }
    }
    std::sort(nb.begin(), nb.end());
  };
  parallel_for(na, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> stamp((size_t)na, -1), nb;
    for (int64_t a = lo; a < hi; ++a) {
      neighbours(a, stamp, nb);
      nbp[a + 1] = (int64_t)nb.size();
      int64_t c = 0;
      for (int64_t m = L.mem_ptr[a]; m < L.mem_ptr[a + 1]; ++m) {
        const int64_t j = L.mem[m];
        for (int64_t g = adj_ptr[j]; g < adj_ptr[j + 1]; ++g) c += L.agg[adj[g]] >= 0 ? 1 : 0;
      }
      nblk[a + 1] = c;
    }
  }, 64);
  for (int64_t a = 0; a < na; ++a) {
    nbp[a + 1] += nbp[a];
    nblk[a + 1] += nblk[a];
  }
  nbi.assign((size_t)nbp[na], 0);
  L.gptr.assign((size_t)nbp[na] + 1, 0);
  L.gblk.resize((size_t)nblk[na]);   // every element is written below
  L.gcol.resize((size_t)nblk[na]);
  parallel_for(na, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> stamp((size_t)na, -1), pos((size_t)na, 0), nb;
    std::vector<int64_t> cur;
    for (int64_t a = lo; a < hi; ++a) {
      neighbours(a, stamp, nb);
      const size_t k1 = nb.size();
      std::copy(nb.begin(), nb.end(), nbi.begin() + nbp[a]);
      for (size_t k = 0; k < k1; ++k) pos[nb[k]] = (int32_t)k;
      cur.assign(k1 + 1, 0);
      for (int64_t m = L.mem_ptr[a]; m < L.mem_ptr[a + 1]; ++m) {
        const int64_t j = L.mem[m];
        for (int64_t g = adj_ptr[j]; g < adj_ptr[j + 1]; ++g) {
          const int32_t B = L.agg[adj[g]];
          if (B >= 0) cur[pos[B] + 1]++;
        }
      }
      const int64_t base = nblk[a];
      for (size_t k = 0; k < k1; ++k) {
        cur[k + 1] += cur[k];
        L.gptr[nbp[a] + k + 1] = base + cur[k + 1];
      }
      for (int64_t m = L.mem_ptr[a]; m < L.mem_ptr[a + 1]; ++m) {
        const int64_t j = L.mem[m];
        for (int64_t g = adj_ptr[j]; g < adj_ptr[j + 1]; ++g) {
          const int32_t B = L.agg[adj[g]];
          if (B < 0) continue;
          const int64_t o = base + cur[pos[B]]++;
          L.gblk[o] = (int32_t)g;
          L.gcol[o] = (int32_t)j;
        }
      }
    }
  }, 64);
  tick("galerkin lists", tp);
  // coarse pseudo-node pattern: 2A and 2A+1 both list [2B, 2B+1 for B in nb(A)]
  const int64_t npairs = nbp[na];
  C.adj_ptr.assign((size_t)C.N + 1, 0);
  C.adj.assign((size_t)4 * npairs, 0);
  C.dslot.assign((size_t)C.N, 0);
  L.pair_o.assign((size_t)npairs * 2, 0);
  parallel_for(na, [&](int64_t lo, int64_t hi) {   // each aggregate writes its own ranges
    for (int64_t a = lo; a < hi; ++a) {
      const int64_t r = 2 * (nbp[a + 1] - nbp[a]);
      C.adj_ptr[2 * a] = 4 * nbp[a];
      C.adj_ptr[2 * a + 1] = 4 * nbp[a] + r;
      for (int s = 0; s < 2; ++s) {
        const int64_t o = 4 * nbp[a] + s * r;
        for (int64_t k = 0; k < nbp[a + 1] - nbp[a]; ++k) {
          const int32_t B = nbi[nbp[a] + k];
          C.adj[o + 2 * k] = 2 * B;
          C.adj[o + 2 * k + 1] = 2 * B + 1;
          if (B == a) C.dslot[2 * a + s] = (int32_t)(2 * k + s);
          L.pair_o[2 * (nbp[a] + k) + s] = (int32_t)(o + 2 * k);
        }
      }
    }
  }, 64);
  C.adj_ptr[2 * na] = 4 * nbp[na];
  tick("coarse pattern", tp);
}

}  // namespace

// Strong-coupling flags (GPU) and host copies of the pattern, coordinates and
// support mask: the inputs of the host hierarchy build.
int amg_gather(llfem_ctx* ctx, const SysSpec& sp, const double* d_xyz, AmgInput* in) {
  hipStream_t st = ctx->stream;
  const int64_t N = sp.N, nadj = sp.nadj;
  if (nadj >= ((int64_t)1 << 31)) return fail(LLFEM_EUNSUP, "AMG: more than 2^31 node blocks");
  uint8_t* d_strong = nullptr;
  LLFEM_TRY(dalloc(ctx, &d_strong, (size_t)nadj));
  k_mg_strength<<<nblk(N), kMB, 0, st>>>(N, sp.adj_ptr, sp.dslot, sp.val64, kStrength, d_strong);
  in->N = N;
  in->adj_ptr.resize((size_t)N + 1);
  in->adj.resize((size_t)nadj);
  in->dslot.resize((size_t)N);
  in->strong.resize((size_t)nadj);
  in->xyz.resize((size_t)3 * N);
  in->mask.resize((size_t)3 * N);
  // first touch of the uninitialised block arrays on all host threads, while
  // the strength kernel runs: the copies would otherwise take their page
  // faults one 4 KB page at a time in the runtime's copy loop
  auto touch = [](auto& v) {
    uint8_t* p = reinterpret_cast<uint8_t*>(v.data());
    const int64_t bytes = (int64_t)(v.size() * sizeof(v[0]));
    parallel_for((bytes + 4095) / 4096, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) p[i * 4096] = 0;
    });
  };
  touch(in->adj);
  touch(in->strong);
  hipError_t e = hipMemcpyAsync(in->adj_ptr.data(), sp.adj_ptr, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(in->adj.data(), sp.adj, sizeof(int32_t) * nadj, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(in->dslot.data(), sp.dslot, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(in->strong.data(), d_strong, (size_t)nadj, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(in->xyz.data(), d_xyz, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(in->mask.data(), sp.fixed, (size_t)3 * N, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  drelease(ctx, d_strong, (size_t)nadj);
  if (e != hipSuccess) return fail(LLFEM_EHIP, std::string("AMG gather: ") + hipGetErrorString(e));
  return LLFEM_OK;
}

// The aggregation hierarchy (CPU only, no device calls: may run on any thread).
namespace {
// every value of v in [lo, hi), checked across host threads
template <class V>
bool all_in(const V& v, int64_t lo, int64_t hi) {
  std::atomic<bool> ok{true};
  parallel_for((int64_t)v.size(), [&](int64_t b, int64_t e) {
    bool o = true;
    for (int64_t i = b; i < e; ++i) o &= v[i] >= lo && v[i] < hi;
    if (!o) ok = false;
  });
  return ok;
}
}  // namespace

int amg_build_host(const AmgInput& in, AmgHost* H) {
  auto t0 = std::chrono::steady_clock::now();
  const int64_t N = in.N;
  H->N0 = N;
  H->lev.clear();
  H->lev.reserve(kMaxLevels + 1);
  H->lev.emplace_back();
  H->lev[0].N = N;
  // rigid-body modes about the centroid, constrained rows zeroed
  double c0[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    for (int64_t j = 0; j < N; ++j) c0[k] += in.xyz[(size_t)k * N + j];
    c0[k] /= std::max<int64_t>(N, 1);
  }
  auto rbm_row = [&](int64_t j, int c, double* out) {
    for (int u = 0; u < 6; ++u) out[u] = 0.0;
    if (in.mask[3 * j + c]) return;
    const double r[3] = {in.xyz[j] - c0[0], in.xyz[(size_t)N + j] - c0[1], in.xyz[(size_t)2 * N + j] - c0[2]};
    out[c] = 1.0;
    if (c == 0) { out[4] = r[2]; out[5] = -r[1]; }
    if (c == 1) { out[3] = -r[2]; out[5] = r[0]; }
    if (c == 2) { out[3] = r[1]; out[4] = -r[0]; }
  };
  std::vector<double> Bn, Bc;
  std::vector<int64_t> nbp, nbp_prev;
  std::vector<int32_t> nbi, nbi_prev;
  for (int l = 0;; ++l) {
    HostLevel C;
    if (l == 0) {
      const Graph S{in.adj_ptr.data(), in.adj.data(), in.strong.data()};
      const Graph F{in.adj_ptr.data(), in.adj.data(), nullptr};
      coarsen(H->lev[0], in.adj_ptr.data(), in.adj.data(), 1, N, rbm_row, S, F, C, Bc, nbp, nbi);
    } else {
      const HostLevel& L = H->lev[l];
      const Graph G{nbp_prev.data(), nbi_prev.data(), nullptr};
      auto row = [&](int64_t j, int c, double* out) {   // real node j/2, DoF 3 (j%2) + c
        const double* R = &Bn[(size_t)(j / 2) * 36 + 6 * (3 * (j % 2) + c)];
        for (int u = 0; u < 6; ++u) out[u] = R[u];
      };
      coarsen(H->lev[l], L.adj_ptr.data(), L.adj.data(), 2, L.N / 2, row, G, G, C, Bc, nbp, nbi);
    }
    const int64_t na = H->lev[l].nagg;
    const int64_t n_real = l == 0 ? N : H->lev[l].N / 2;
    if (na == 0) return fail(LLFEM_EUNSUP, "AMG: no aggregates (empty operator)");
    H->lev.push_back(std::move(C));
    if (na <= kCoarseAggs) break;
    if ((int)H->lev.size() >= kMaxLevels || 5 * na > 4 * n_real) {
      if (6 * na > kMaxDense) return fail(LLFEM_EUNSUP, "AMG: coarsening stalled");
      break;
    }
    Bn.swap(Bc);
    nbp_prev.swap(nbp);
    nbi_prev.swap(nbi);
  }
  const auto tc = std::chrono::steady_clock::now();
  // every index the kernels follow stays inside its array
  for (size_t l = 0; l + 1 < H->lev.size(); ++l) {
    const HostLevel& L = H->lev[l];
    const HostLevel& C = H->lev[l + 1];
    const int64_t nadj = l == 0 ? (int64_t)in.adj.size() : L.adj_ptr.back();
    const int64_t cnadj = C.adj_ptr.back();
    bool ok = (int64_t)L.agg.size() == L.N && (int64_t)L.Q.size() == 18 * L.N && C.N == 2 * L.nagg &&
              (int64_t)L.mem_ptr.size() == L.nagg + 1 && L.mem_ptr.back() == (int64_t)L.mem.size() &&
              (int64_t)L.gptr.size() == (int64_t)L.pair_o.size() / 2 + 1 &&
              L.gptr.back() == (int64_t)L.gblk.size() && L.gblk.size() == L.gcol.size() &&
              (int64_t)C.dslot.size() == C.N && (int64_t)C.mask.size() == 3 * C.N;
    ok = ok && all_in(L.agg, -1, L.nagg) && all_in(L.mem, 0, L.N) && all_in(L.gblk, 0, nadj) &&
         all_in(L.gcol, 0, L.N) && all_in(L.pair_o, 0, cnadj - 1) && all_in(C.adj, 0, C.N);
    for (int64_t j = 0; j < C.N && ok; ++j)
      ok = C.adj_ptr[j] <= C.adj_ptr[j + 1] && C.dslot[j] >= 0 && C.dslot[j] < C.adj_ptr[j + 1] - C.adj_ptr[j] &&
           C.adj[C.adj_ptr[j] + C.dslot[j]] == j;
    if (!ok) return fail(LLFEM_EINVAL, "AMG: inconsistent hierarchy at level " + std::to_string(l));
  }
  if (3 * H->lev.back().N > kMaxDense) return fail(LLFEM_EUNSUP, "AMG: coarsest level too large");
  H->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (std::getenv("LLFEM_AMG_TRACE"))
    std::fprintf(stderr, "[amg]   index checks %.1f ms\n",
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc).count());
  return LLFEM_OK;
}

void amg_free(llfem_ctx* ctx, AmgHier* h) {
  if (!h) return;
  for (auto& o : h->owned) dfree(ctx, o.first, o.second);
  delete h;
}

int amg_upload(llfem_ctx* ctx, const AmgHost& H, const SysSpec& sp, AmgHier** out) {
  hipStream_t st = ctx->stream;
  AmgHier* h = new AmgHier();
  h->host_ms = H.ms;
  int rc = LLFEM_OK;
  auto up = [&](const auto& v, auto** dst) -> int {
    using T = typename std::decay_t<decltype(v)>::value_type;
    *dst = nullptr;
    if (v.empty()) return LLFEM_OK;
    void* p = nullptr;
    LLFEM_TRY(dmalloc(ctx, &p, sizeof(T) * v.size()));
    h->owned.emplace_back(p, sizeof(T) * v.size());
    h->bytes += sizeof(T) * v.size();
    LLFEM_HIP(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    *dst = static_cast<std::remove_const_t<std::remove_pointer_t<std::decay_t<decltype(*dst)>>>*>(p);
    return LLFEM_OK;
  };
  h->lev.resize(H.lev.size());
  for (size_t l = 0; l < H.lev.size() && rc == LLFEM_OK; ++l) {
    const HostLevel& L = H.lev[l];
    AmgHier::Lev& D = h->lev[l];
    D.N = L.N;
    D.nagg = L.nagg;
    D.npairs = (int64_t)L.pair_o.size() / 2;
    if (l == 0) {
      D.adj_ptr = sp.adj_ptr;
      D.adj = sp.adj;
      D.dslot = sp.dslot;
      D.mask = sp.fixed;
      D.nadj = sp.nadj;
    } else {
      D.nadj = L.adj_ptr.back();
      if ((rc = up(L.adj_ptr, &D.adj_ptr)) || (rc = up(L.adj, &D.adj)) || (rc = up(L.dslot, &D.dslot)) ||
          (rc = up(L.mask, &D.mask)))
        break;
    }
    if ((rc = up(L.agg, &D.agg)) || (rc = up(L.Q, &D.Q)) || (rc = up(L.mem_ptr, &D.mem_ptr)) ||
        (rc = up(L.mem, &D.mem)) || (rc = up(L.pair_o, &D.pair_o)) || (rc = up(L.gptr, &D.gptr)) ||
        (rc = up(L.gblk, &D.gblk)) || (rc = up(L.gcol, &D.gcol)))
      break;
  }
  if (rc == LLFEM_OK) {
    const HostLevel& C = H.lev.back();
    h->c_adj_ptr = C.adj_ptr;
    h->c_adj = C.adj;
    h->c_mask = C.mask;
    if (hipStreamSynchronize(st) != hipSuccess) rc = fail(LLFEM_EHIP, "AMG upload failed");
  }
  if (rc != LLFEM_OK) {
    (void)hipStreamSynchronize(st);
    amg_free(ctx, h);
    return rc;
  }
  *out = h;
  return LLFEM_OK;
}

int64_t amg_levels(const AmgHier* h) { return h ? (int64_t)h->lev.size() : 0; }
size_t amg_bytes(const AmgHier* h) { return h ? h->bytes : 0; }
double amg_host_ms(const AmgHier* h) { return h ? h->host_ms : 0.0; }

// ---- per-system numeric levels ----------------------------------------------------

struct AmgSys {
  const AmgHier* h = nullptr;
  void* mem = nullptr;
  size_t bytes = 0;
  std::vector<MgEnt> ent;          // one per level (the last: coarsest)
  MgEnt* d_solo = nullptr;         // device copy of `ent` with state = NULL (restarts)
  void* solo_tab = nullptr;
  size_t solo_bytes = 0;
  struct Tab {
    MgEnt* ents = nullptr;
    int32_t *smap = nullptr, *vmap = nullptr, *amap = nullptr;
    int nent = 0;
    int64_t nsb = 0, nvb = 0, nab = 0;
    MgEnt* cents = nullptr;
    int ncent = 0;
  };
  std::vector<Tab> solo;
  std::vector<double> lambda;
};

namespace {

struct ChebCoef {
  float ca, cbt, cbd;
};

// Chebyshev recurrence coefficients for the interval [lo, hi] x lambda: the
// same for every level and system (only theta and delta scale with lambda).
void cheb_coefs(std::vector<ChebCoef>& pre, std::vector<ChebCoef>& post) {
  const double theta = 0.5 * (kChebHi + kChebLo), delta = 0.5 * (kChebHi - kChebLo);
  const double sigma = theta / delta;
  pre.clear();
  post.clear();
  double rho = 1.0 / sigma;
  for (int s = 1; s < kCheb; ++s) {   // pre: step 0 is k_mg_cheb0
    const double rn = 1.0 / (2.0 * sigma - rho);
    pre.push_back({(float)(rn * rho), 0.f, (float)(2.0 * rn)});
    rho = rn;
  }
  rho = 1.0 / sigma;
  post.push_back({0.f, 1.f, 0.f});
  for (int s = 1; s < kCheb; ++s) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    post.push_back({(float)(rn * rho), 0.f, (float)(2.0 * rn)});
    rho = rn;
  }
}

// Launch tables for a set of systems (block -> entry maps per level).
struct TabBuild {
  std::vector<std::vector<MgEnt>> ents, cents;
  std::vector<std::vector<int32_t>> smap, vmap, amap;
};

void tab_add(TabBuild& T, const AmgSys* s, const double* state) {
  const int L = (int)s->ent.size() - 1;
  if ((int)T.ents.size() < L + 1) {
    T.ents.resize(L + 1);
    T.cents.resize(L + 1);
    T.smap.resize(L + 1);
    T.vmap.resize(L + 1);
    T.amap.resize(L + 1);
  }
  for (int l = 0; l < L; ++l) {
    MgEnt e = s->ent[l];
    e.state = state;
    const int32_t id = (int32_t)T.ents[l].size();
    e.sblk0 = (int32_t)T.smap[l].size();
    e.vblk0 = (int32_t)T.vmap[l].size();
    e.ablk0 = (int32_t)T.amap[l].size();
    T.smap[l].insert(T.smap[l].end(), (size_t)e.nsblk, id);
    T.vmap[l].insert(T.vmap[l].end(), (size_t)e.nvblk, id);
    T.amap[l].insert(T.amap[l].end(), (size_t)e.nablk, id);
    T.ents[l].push_back(e);
  }
  MgEnt c = s->ent[L];
  c.state = state;
  T.cents[L].push_back(c);
}

int tab_upload(llfem_ctx* ctx, const TabBuild& T, void** mem, size_t* bytes, std::vector<AmgSys::Tab>* out) {
  size_t need = 0;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  for (size_t l = 0; l < T.ents.size(); ++l)
    need += al(sizeof(MgEnt) * T.ents[l].size()) + al(sizeof(MgEnt) * T.cents[l].size()) +
            al(4 * T.smap[l].size()) + al(4 * T.vmap[l].size()) + al(4 * T.amap[l].size());
  need += 256;
  if (*bytes < need) {
    if (*mem) dfree(ctx, *mem, *bytes);
    *mem = nullptr;
    *bytes = 0;
    LLFEM_TRY(dmalloc(ctx, mem, need * 2));
    *bytes = need * 2;
  }
  hipStream_t st = ctx->stream;
  char* m = static_cast<char*>(*mem);
  out->assign(T.ents.size(), AmgSys::Tab{});
  auto put = [&](const void* src, size_t b) -> char* {
    char* p = m;
    if (b) {
      if (hipMemcpyAsync(p, src, b, hipMemcpyHostToDevice, st) != hipSuccess) return nullptr;
    }
    m += al(b);
    return p;
  };
  for (size_t l = 0; l < T.ents.size(); ++l) {
    AmgSys::Tab& t = (*out)[l];
    t.nent = (int)T.ents[l].size();
    t.ncent = (int)T.cents[l].size();
    t.nsb = (int64_t)T.smap[l].size();
    t.nvb = (int64_t)T.vmap[l].size();
    t.nab = (int64_t)T.amap[l].size();
    t.ents = reinterpret_cast<MgEnt*>(put(T.ents[l].data(), sizeof(MgEnt) * T.ents[l].size()));
    t.cents = reinterpret_cast<MgEnt*>(put(T.cents[l].data(), sizeof(MgEnt) * T.cents[l].size()));
    t.smap = reinterpret_cast<int32_t*>(put(T.smap[l].data(), 4 * T.smap[l].size()));
    t.vmap = reinterpret_cast<int32_t*>(put(T.vmap[l].data(), 4 * T.vmap[l].size()));
    t.amap = reinterpret_cast<int32_t*>(put(T.amap[l].data(), 4 * T.amap[l].size()));
    if (!t.ents || !t.cents || !t.smap || !t.vmap || !t.amap) return fail(LLFEM_EHIP, "AMG: table upload failed");
  }
  LLFEM_HIP(hipStreamSynchronize(st));   // host tables go out of scope
  return LLFEM_OK;
}

int run_vcycle(llfem_ctx* ctx, const std::vector<AmgSys::Tab>& T) {
  hipStream_t st = ctx->stream;
  static thread_local std::vector<ChebCoef> pre, post;
  if (post.empty()) cheb_coefs(pre, post);
  const int L = (int)T.size();
  for (int l = 0; l < L; ++l) {
    const AmgSys::Tab& t = T[l];
    if (t.nent) {
      k_mg_cheb0<<<(unsigned)t.nvb, kMB, 0, st>>>(t.ents, t.vmap);
      for (const ChebCoef& c : pre) {
        k_mg_spmv<<<(unsigned)t.nsb, kMB, 0, st>>>(t.ents, t.smap);
        k_mg_cheb<<<(unsigned)t.nvb, kMB, 0, st>>>(t.ents, t.vmap, c.ca, c.cbt, c.cbd, 0);
      }
      k_mg_spmv<<<(unsigned)t.nsb, kMB, 0, st>>>(t.ents, t.smap);
      k_mg_restrict<<<(unsigned)t.nab, kMB, 0, st>>>(t.ents, t.amap);
    }
    if (t.ncent) k_mg_coarse<<<(unsigned)t.ncent, kMB, 0, st>>>(t.cents);
  }
  for (int l = L - 1; l >= 0; --l) {
    const AmgSys::Tab& t = T[l];
    if (!t.nent) continue;
    k_mg_prolong<<<(unsigned)t.nvb, kMB, 0, st>>>(t.ents, t.vmap, kOverCorrect);
    for (size_t s = 0; s < post.size(); ++s) {
      k_mg_spmv<<<(unsigned)t.nsb, kMB, 0, st>>>(t.ents, t.smap);
      k_mg_cheb<<<(unsigned)t.nvb, kMB, 0, st>>>(t.ents, t.vmap, post[s].ca, post[s].cbt, post[s].cbd,
                                                  (l == 0 && s + 1 == post.size()) ? 1 : 0);
    }
  }
  LLFEM_HIP(hipGetLastError());
  return LLFEM_OK;
}

// Dense inverse by Gauss-Jordan with partial pivoting (FP64, n <= kMaxDense).
bool invert_dense(int n, std::vector<double>& A) {
  std::vector<int> piv(n);
  std::iota(piv.begin(), piv.end(), 0);
  std::vector<double> I((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
    if (!(std::fabs(A[(size_t)p * n + k]) > 0.0)) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        std::swap(A[(size_t)p * n + j], A[(size_t)k * n + j]);
        std::swap(I[(size_t)p * n + j], I[(size_t)k * n + j]);
      }
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int j = 0; j < n; ++j) {
      A[(size_t)k * n + j] *= inv;
      I[(size_t)k * n + j] *= inv;
    }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      const double f = A[(size_t)i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
        I[(size_t)i * n + j] -= f * I[(size_t)k * n + j];
      }
    }
  }
  A.swap(I);
  return true;
}

}  // namespace

void amg_sys_free(llfem_ctx* ctx, AmgSys* s) {
  if (!s) return;
  if (s->solo_tab) dfree(ctx, s->solo_tab, s->solo_bytes);
  if (s->mem) dfree(ctx, s->mem, s->bytes);
  delete s;
}

// Numeric levels of one system: Galerkin operators, inverse diagonals,
// Chebyshev bounds, dense coarsest inverse, launch tables for restarts.
// r, z, q: the system's FP32 residual, preconditioned residual and a free
// work vector (level 0's b, x and y); rzpart: where level 0 leaves the
// partial r.z per vector block (vector blocks as the engine's update kernel).
int amg_sys_create(llfem_ctx* ctx, const AmgHier* h, const float* val32, float* r, float* z, float* q,
                   double* rzpart, AmgSys** out) {
  hipStream_t st = ctx->stream;
  const int L = (int)h->lev.size() - 1;   // coarsest index
  AmgSys* s = new AmgSys();
  s->h = h;
  // memory: level 0 dinv + d; levels 1..L: val, dinv, b, x, y, d; Ainv; reduction scratch
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const int64_t nc = 3 * h->lev[L].N;
  size_t need = al(4 * 3 * (size_t)h->lev[0].N) * 2;
  for (int l = 1; l <= L; ++l)
    need += al(4 * 9 * (size_t)h->lev[l].nadj) + 5 * al(4 * 3 * (size_t)h->lev[l].N);
  need += al(4 * (size_t)nc * nc);
  const int64_t maxn = 3 * h->lev[0].N;
  need += al(sizeof(double) * (2 * (size_t)nblk(maxn) + 2 * (size_t)(L + 1) + 8));
  s->bytes = need;
  int rc = dmalloc(ctx, &s->mem, need);
  if (rc != LLFEM_OK) {
    delete s;
    return rc;
  }
  char* m = static_cast<char*>(s->mem);
  auto take = [&](size_t b) -> char* {
    char* p = m;
    m += al(b);
    return p;
  };
  s->ent.assign(L + 1, MgEnt{});
  std::vector<float*> val(L + 1), dinv(L + 1);
  val[0] = const_cast<float*>(val32);
  for (int l = 0; l <= L; ++l) {
    const auto& D = h->lev[l];
    MgEnt& e = s->ent[l];
    e.N = D.N;
    e.adj_ptr = D.adj_ptr;
    e.adj = D.adj;
    e.mask = D.mask;
    if (l == 0) {
      dinv[0] = (float*)take(4 * 3 * (size_t)D.N);
      e.d = (float*)take(4 * 3 * (size_t)D.N);
      e.b = r;
      e.x = z;
      e.y = q;
    } else {
      val[l] = (float*)take(4 * 9 * (size_t)D.nadj);
      dinv[l] = (float*)take(4 * 3 * (size_t)D.N);
      e.b = (float*)take(4 * 3 * (size_t)D.N);
      e.x = (float*)take(4 * 3 * (size_t)D.N);
      e.y = (float*)take(4 * 3 * (size_t)D.N);
      e.d = (float*)take(4 * 3 * (size_t)D.N);
    }
    e.val = val[l];
    e.dinv = dinv[l];
    e.nsblk = (int32_t)cap(D.N, kMSpN, kMaxSB);
    // level 0 uses the engine's vector blocks (256 nodes, at most 1024 blocks),
    // so its r.z partials line up with the engine's r.r partials
    e.nvblk = (int32_t)(l == 0 ? cap(D.N, kMB, kMaxVB) : cap(3 * D.N, kMB, kMaxVB));
    e.agg = D.agg;
    e.Q = D.Q;
    e.mem_ptr = D.mem_ptr;
    e.mem = D.mem;
    e.nagg = D.nagg;
    e.nablk = (int32_t)cap(D.nagg, kMSpN, kMaxAB);
  }
  for (int l = 0; l < L; ++l) {
    s->ent[l].bc = s->ent[l + 1].b;
    s->ent[l].xc = s->ent[l + 1].x;
  }
  s->ent[0].rzpart = rzpart;
  float* Ainv = (float*)take(4 * (size_t)nc * nc);
  double* part = (double*)take(sizeof(double) * (2 * (size_t)nblk(maxn) + 2 * (size_t)(L + 1) + 8));
  double* lam = part + 2 * (size_t)nblk(maxn);
  if ((size_t)(m - static_cast<char*>(s->mem)) > s->bytes) {
    amg_sys_free(ctx, s);
    return fail(LLFEM_EINVAL, "AMG: workspace layout overflow");
  }
  // Galerkin operators, inverse diagonals
  for (int l = 0; l < L; ++l) {
    const auto& D = h->lev[l];
    k_galerkin<<<nblk(6 * D.npairs), kMB, 0, st>>>(D.npairs, D.gptr, D.gblk, D.gcol, D.adj, D.Q, val[l], D.pair_o,
                                                  val[l + 1]);
  }
  for (int l = 0; l <= L; ++l) {
    const auto& D = h->lev[l];
    k_mg_dinv<<<nblk(D.N), kMB, 0, st>>>(D.N, D.adj_ptr, D.dslot, val[l], D.mask, dinv[l]);
  }
  // largest eigenvalue of D^-1 A per smoothed level (power steps, Rayleigh quotient)
  for (int l = 0; l < L; ++l) {
    const auto& D = h->lev[l];
    const MgEnt& e = s->ent[l];
    const int64_t n = 3 * D.N;
    float* v = e.d;
    float* y = e.y;
    k_mg_rand<<<nblk(n), kMB, 0, st>>>(n, dinv[l], v);
    for (int k = 0; k < kPowerSteps; ++k) {
      k_mg_spmv1<<<nblk(D.N * kMG), kMB, 0, st>>>(D.N, D.adj_ptr, D.adj, val[l], D.mask, v, y);
      k_mg_dscale<<<nblk(n), kMB, 0, st>>>(n, dinv[l], y, v);
    }
    k_mg_spmv1<<<nblk(D.N * kMG), kMB, 0, st>>>(D.N, D.adj_ptr, D.adj, val[l], D.mask, v, y);
    k_mg_rq<<<nblk(n), kMB, 0, st>>>(n, v, y, dinv[l], part);
    k_mg_sum2<<<1, kMB, 0, st>>>(nblk(n), part, lam + 2 * l);
  }
  std::vector<double> hl(2 * (size_t)(L + 1), 0.0);
  std::vector<float> cval((size_t)9 * h->lev[L].nadj);
  hipError_t e = hipMemcpyAsync(hl.data(), lam, sizeof(double) * 2 * L, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(cval.data(), val[L], sizeof(float) * cval.size(), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    amg_sys_free(ctx, s);
    return fail(LLFEM_EHIP, std::string("AMG set-up: ") + hipGetErrorString(e));
  }
  s->lambda.assign(L, 0.0);
  for (int l = 0; l < L; ++l) {
    double lmb = hl[2 * l + 1] > 0 ? hl[2 * l] / hl[2 * l + 1] : 0.0;
    if (!(lmb > 0.0) || !std::isfinite(lmb)) lmb = 1.0;   // empty level: any positive scale
    s->lambda[l] = lmb;
    s->ent[l].inv_theta = (float)(1.0 / (0.5 * (kChebHi + kChebLo) * lmb));
    s->ent[l].inv_delta = (float)(1.0 / (0.5 * (kChebHi - kChebLo) * lmb));
  }
  // dense coarsest inverse (masked DoFs: identity in, zero out)
  {
    const int n = (int)nc;
    std::vector<double> A((size_t)n * n, 0.0);
    const int64_t NL = h->lev[L].N;
    for (int64_t j = 0; j < NL; ++j)
      for (int64_t g = h->c_adj_ptr[j]; g < h->c_adj_ptr[j + 1]; ++g) {
        const int64_t i = h->c_adj[g];
        for (int c = 0; c < 3; ++c)
          for (int d = 0; d < 3; ++d) A[(size_t)(3 * j + c) * n + 3 * i + d] = cval[9 * g + 3 * c + d];
      }
    std::vector<uint8_t> dead(n, 0);
    for (int i = 0; i < n; ++i) dead[i] = h->c_mask[i] || !(A[(size_t)i * n + i] > 0.0);
    for (int i = 0; i < n; ++i)
      if (dead[i]) {
        for (int k = 0; k < n; ++k) A[(size_t)i * n + k] = A[(size_t)k * n + i] = 0.0;
        A[(size_t)i * n + i] = 1.0;
      }
    if (!invert_dense(n, A)) {
      amg_sys_free(ctx, s);
      return fail(LLFEM_EINVAL, "AMG: singular coarsest operator");
    }
    std::vector<float> Af((size_t)n * n);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < n; ++k)
        Af[(size_t)i * n + k] = (dead[i] || dead[k]) ? 0.f : (float)A[(size_t)i * n + k];
    LLFEM_HIP(hipMemcpyAsync(Ainv, Af.data(), sizeof(float) * Af.size(), hipMemcpyHostToDevice, st));
    s->ent[L].Ainv = Ainv;
    s->ent[L].nc = n;
    LLFEM_HIP(hipStreamSynchronize(st));
  }
  // restart tables (this system alone, always runs)
  TabBuild T;
  tab_add(T, s, nullptr);
  rc = tab_upload(ctx, T, &s->solo_tab, &s->solo_bytes, &s->solo);
  if (rc != LLFEM_OK) {
    amg_sys_free(ctx, s);
    return rc;
  }
  *out = s;
  return LLFEM_OK;
}

// z = M r for one system (reliable-update restarts); partial r.z as above.
int amg_apply_solo(llfem_ctx* ctx, AmgSys* s) { return run_vcycle(ctx, s->solo); }

struct AmgBatch {
  void* mem = nullptr;
  size_t bytes = 0;
  std::vector<AmgSys::Tab> tab;
};

// Tables of one V-cycle launch over several systems; states[i]: system i's
// state word (its V-cycle runs while the word reads 0).
int amg_batch_build(llfem_ctx* ctx, const std::vector<AmgSys*>& sys, const std::vector<const double*>& states,
                    AmgBatch** b) {
  if (!*b) *b = new AmgBatch();
  TabBuild T;
  for (size_t i = 0; i < sys.size(); ++i) tab_add(T, sys[i], states[i]);
  if (sys.empty()) {
    (*b)->tab.clear();
    return LLFEM_OK;
  }
  return tab_upload(ctx, T, &(*b)->mem, &(*b)->bytes, &(*b)->tab);
}

int amg_batch_apply(llfem_ctx* ctx, AmgBatch* b) {
  if (!b || b->tab.empty()) return LLFEM_OK;
  return run_vcycle(ctx, b->tab);
}

void amg_batch_free(llfem_ctx* ctx, AmgBatch* b) {
  if (!b) return;
  if (b->mem) dfree(ctx, b->mem, b->bytes);
  delete b;
}

const std::vector<double>& amg_lambda(const AmgSys* s) { return s->lambda; }

int amg_build(llfem_ctx* ctx, const SysSpec& sp, const double* d_xyz, AmgHier** out) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  AmgInput in;
  LLFEM_TRY(amg_gather(ctx, sp, d_xyz, &in));
  const auto t1 = clk::now();
  AmgHost H;
  LLFEM_TRY(amg_build_host(in, &H));
  const auto t2 = clk::now();
  const int rc = amg_upload(ctx, H, sp, out);
  if (std::getenv("LLFEM_AMG_TRACE")) {   // set-up phases, wall clock
    (void)hipStreamSynchronize(ctx->stream);
    const auto t3 = clk::now();
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::fprintf(stderr, "[amg] N %lld: gather %.1f ms, host build %.1f ms, upload %.1f ms\n", (long long)sp.N,
                 ms(t0, t1), ms(t1, t2), ms(t2, t3));
  }
  return rc;
}

// Asynchronous build: the device gather runs at once, the host build on a
// thread of its own, the upload when the caller collects the job.
struct AmgJob {
  std::future<int> fut;
  AmgInput in;
  AmgHost host;
  std::string err;
};

int amg_job_start(llfem_ctx* ctx, const SysSpec& sp, const double* d_xyz, AmgJob** out) {
  AmgJob* j = new AmgJob();
  int rc = amg_gather(ctx, sp, d_xyz, &j->in);
  if (rc != LLFEM_OK) {
    delete j;
    return rc;
  }
  j->fut = std::async(std::launch::async, [j]() {
    const int r = amg_build_host(j->in, &j->host);
    if (r != LLFEM_OK) j->err = llfem_last_error();
    return r;
  });
  *out = j;
  return LLFEM_OK;
}

bool amg_job_ready(AmgJob* j) {
  return j->fut.wait_for(std::chrono::seconds(0)) == std::future_status::ready;
}

void amg_job_drop(AmgJob* j) {
  if (!j) return;
  j->fut.wait();
  delete j;
}

int amg_job_finish(llfem_ctx* ctx, AmgJob* j, const SysSpec& sp, AmgHier** out) {
  int rc = j->fut.get();
  if (rc == LLFEM_OK) rc = amg_upload(ctx, j->host, sp, out);
  else set_error(j->err);
  delete j;
  return rc;
}

}  // namespace llfem
