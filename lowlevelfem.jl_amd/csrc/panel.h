// Device code shared by the assembly translation units: element geometry,
// the per-block element integrals and the node-panel wavefront body.  Included
// inside an unnamed namespace, so each translation unit holds its own copy.
#pragma once

#include "common.h"

namespace llfem {
namespace {

constexpr int kMaxQ = 1024;
constexpr size_t kLdsPerCU = 160 * 1024;   // gfx950 LDS per CU / max per workgroup
constexpr int kXcds = 8;                    // MI355X accelerator complex dies

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
__device__ __forceinline__ double inv3(const double* J, double* Mi) {
  // cofactor inverse, as inv3x3_fast (src/linear.jl:89-114); J row-major
  // J[i][al] = dx_i/dxi_al; Mi row-major = J^{-1}, Mi[al][i] = dxi_al/dx_i.
  double c11 = J[4] * J[8] - J[5] * J[7];
  double c12 = -(J[3] * J[8] - J[5] * J[6]);
  double c13 = J[3] * J[7] - J[4] * J[6];
  double c21 = -(J[1] * J[8] - J[2] * J[7]);
  double c22 = J[0] * J[8] - J[2] * J[6];
  double c23 = -(J[0] * J[7] - J[1] * J[6]);
  double c31 = J[1] * J[5] - J[2] * J[4];
  double c32 = -(J[0] * J[5] - J[2] * J[3]);
  double c33 = J[0] * J[4] - J[1] * J[3];
  double det = J[0] * c11 + J[1] * c12 + J[2] * c13;
  double id = 1.0 / det;
  Mi[0] = c11 * id; Mi[1] = c21 * id; Mi[2] = c31 * id;
  Mi[3] = c12 * id; Mi[4] = c22 * id; Mi[5] = c32 * id;
  Mi[6] = c13 * id; Mi[7] = c23 * id; Mi[8] = c33 * id;
  return det;
}

// J[i][al] = sum_a x_i^a dN_a[al], evaluated on coordinates relative to the
// element's first node (sum_a dN_a = 0 makes this the same Jacobian, without
// the cancellation of large absolute coordinates).
__device__ __forceinline__ void jacobian(int64_t E, int64_t N, int npe, int64_t e,
                                         const double* __restrict__ xyz,
                                         const int32_t* __restrict__ conn,
                                         const double* __restrict__ dNq, double* J) {
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = 0.0;
  const int32_t v0 = conn[e];
  const double x0 = xyz[v0], y0 = xyz[N + v0], z0 = xyz[2 * N + v0];
  for (int a = 1; a < npe; ++a) {
    const int32_t v = conn[(int64_t)a * E + e];
    const double x = xyz[v] - x0, y = xyz[N + v] - y0, z = xyz[2 * N + v] - z0;
    const double g0 = dNq[3 * a], g1 = dNq[3 * a + 1], g2 = dNq[3 * a + 2];
    J[0] += x * g0; J[1] += x * g1; J[2] += x * g2;
    J[3] += y * g0; J[4] += y * g1; J[5] += y * g2;
    J[6] += z * g0; J[7] += z * g1; J[8] += z * g2;
  }
}

// flags |= 1 if some element's Jacobian varies over the points, |= 2 if some
// |det J| is zero or not finite (the reference's singular-geometry error).
__global__ void k_geom_check(int64_t E, int64_t N, int npe, int q, const double* __restrict__ xyz,
                             const int32_t* __restrict__ conn, const double* __restrict__ dN,
                             int* __restrict__ flags) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double J0[9], J[9], Mi[9];
  jacobian(E, N, npe, e, xyz, conn, dN, J0);
  double scale = 0.0;
  for (int k = 0; k < 9; ++k) scale = fmax(scale, fabs(J0[k]));
  int f = 0;
  double det = inv3(J0, Mi);
  if (!(fabs(det) > 0.0) || !isfinite(det)) f |= 2;
  for (int p = 1; p < q; ++p) {
    jacobian(E, N, npe, e, xyz, conn, dN + (size_t)p * npe * 3, J);
    double diff = 0.0;
    for (int k = 0; k < 9; ++k) diff = fmax(diff, fabs(J[k] - J0[k]));
    if (diff > 1e-12 * scale) f |= 1;
    double d = inv3(J, Mi);
    if (!(fabs(d) > 0.0) || !isfinite(d)) f |= 2;
  }
  if (f) atomicOr(flags, f);
}

// Geometry records, one 128-byte line each (element-major, so a gather by
// element id touches one cache line):
//   double[0..8] Minv (row-major J^{-1}), double[9] wd,
//   float[20..29] the same ten values in FP32 (for the FP32 integration path).
// Affine: one record per element, wd = |det J|.  Otherwise one record per
// (element, point), wd = |det J_p| w_p.
constexpr int kGeoStride = 16;   // doubles

// F32: also write the FP32 copy that the FP32-integration kernels read
template <bool F32 = true>
__device__ __forceinline__ void geometry_store(double* g, const double* Mi, double wd) {
  double2* g2 = reinterpret_cast<double2*>(g);
  g2[0] = make_double2(Mi[0], Mi[1]);
  g2[1] = make_double2(Mi[2], Mi[3]);
  g2[2] = make_double2(Mi[4], Mi[5]);
  g2[3] = make_double2(Mi[6], Mi[7]);
  g2[4] = make_double2(Mi[8], wd);
  if constexpr (!F32) return;
  float4* f4 = reinterpret_cast<float4*>(g + 10);
  f4[0] = make_float4((float)Mi[0], (float)Mi[1], (float)Mi[2], (float)Mi[3]);
  f4[1] = make_float4((float)Mi[4], (float)Mi[5], (float)Mi[6], (float)Mi[7]);
  f4[2] = make_float4((float)Mi[8], (float)wd, 0.f, 0.f);
}

// scale (nullable): per-element factor of the integrand (material groups),
// folded into wd.
__device__ __forceinline__ void geometry_one(int64_t E, int64_t N, int npe, int q, int affine,
                                             const double* __restrict__ xyz,
                                             const int32_t* __restrict__ conn,
                                             const double* __restrict__ dN,
                                             const double* __restrict__ w, double* __restrict__ geo,
                                             int64_t e, const double* __restrict__ scale = nullptr) {
  double J[9], Mi[9];
  const int np = affine ? 1 : q;
  const double se = scale ? scale[e] : 1.0;
  for (int p = 0; p < np; ++p) {
    jacobian(E, N, npe, e, xyz, conn, dN + (size_t)p * npe * 3, J);
    const double det = inv3(J, Mi);
    const double wd = (affine ? fabs(det) : fabs(det) * w[p]) * se;
    geometry_store(geo + ((size_t)e * np + p) * kGeoStride, Mi, wd);
  }
}

// Affine elements with a compile-time node count: the same sums in the same
// order as jacobian(), with every connectivity and coordinate load issued
// before the first is used (the run-time loop waits for each node's gathers
// in turn).  Coordinates come from the node-major copy (x, y, z, 0 per node:
// two 16-byte loads from one line per node instead of three loads from three
// arrays); F32 selects the FP32 copy in the record.
template <int NPE, bool F32>
__global__ __launch_bounds__(256) void k_geometry_affine(int64_t E, const double4* __restrict__ xyz4,
                                                         const int32_t* __restrict__ conn,
                                                         const double* __restrict__ dN, double* __restrict__ geo,
                                                         const double* __restrict__ scale) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int32_t v[NPE];
#pragma unroll
  for (int a = 0; a < NPE; ++a) v[a] = conn[(int64_t)a * E + e];
  double x[NPE], y[NPE], z[NPE];
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const double4 c = xyz4[v[a]];
    x[a] = c.x;
    y[a] = c.y;
    z[a] = c.z;
  }
  double J[9], Mi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = 0.0;
#pragma unroll
  for (int a = 1; a < NPE; ++a) {
    const double dx = x[a] - x[0], dy = y[a] - y[0], dz = z[a] - z[0];
    const double g0 = dN[3 * a], g1 = dN[3 * a + 1], g2 = dN[3 * a + 2];
    J[0] += dx * g0; J[1] += dx * g1; J[2] += dx * g2;
    J[3] += dy * g0; J[4] += dy * g1; J[5] += dy * g2;
    J[6] += dz * g0; J[7] += dz * g1; J[8] += dz * g2;
  }
  const double det = inv3(J, Mi);
  geometry_store<F32>(geo + (size_t)e * kGeoStride, Mi, fabs(det) * (scale ? scale[e] : 1.0));
}

// dNT[(a * q + p) * 3 + c] = dN[(p * npe + a) * 3 + c]: the gradient table
// with the points fastest.
__global__ __launch_bounds__(256) void k_dn_transpose(int npe, int q, const double* __restrict__ dN,
                                                      double* __restrict__ dNT) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)q * npe * 3) return;
  const int c = (int)(t % 3);
  const int64_t r = t / 3;
  const int p = (int)(r % q);
  const int64_t a = r / q;
  dNT[t] = dN[((int64_t)p * npe + a) * 3 + c];
}

// Curved elements: one block per element, thread per point (blockDim >= q).
// The element's node coordinates, relative to its first node, are gathered
// once into LDS; each point then sums over the nodes with the same sums in
// the same order as jacobian(), its gradients from dNT (the lanes of a wave,
// neighbouring points, load neighbouring doubles).
constexpr int kCurvedMaxNpe = 128;

__global__ __launch_bounds__(128) void k_geometry_curved(int64_t E, int64_t N, int npe, int q,
                                                         const double* __restrict__ xyz,
                                                         const int32_t* __restrict__ conn,
                                                         const double* __restrict__ dNT,
                                                         const double* __restrict__ w, double* __restrict__ geo,
                                                         const double* __restrict__ scale) {
  __shared__ double rel[3][kCurvedMaxNpe];
  const int64_t e = blockIdx.x;
  const int32_t v0 = conn[e];
  const double x0 = xyz[v0], y0 = xyz[N + v0], z0 = xyz[2 * N + v0];
  for (int a = threadIdx.x; a < npe; a += blockDim.x) {
    const int32_t v = conn[(int64_t)a * E + e];
    rel[0][a] = xyz[v] - x0;
    rel[1][a] = xyz[N + v] - y0;
    rel[2][a] = xyz[2 * N + v] - z0;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < q; p += blockDim.x) {
    double J[9], Mi[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) J[k] = 0.0;
    for (int a = 1; a < npe; ++a) {
      const double x = rel[0][a], y = rel[1][a], z = rel[2][a];
      const double* g = dNT + ((int64_t)a * q + p) * 3;
      const double g0 = g[0], g1 = g[1], g2 = g[2];
      J[0] += x * g0; J[1] += x * g1; J[2] += x * g2;
      J[3] += y * g0; J[4] += y * g1; J[5] += y * g2;
      J[6] += z * g0; J[7] += z * g1; J[8] += z * g2;
    }
    const double det = inv3(J, Mi);
    geometry_store(geo + ((size_t)e * q + p) * kGeoStride, Mi, fabs(det) * w[p] * (scale ? scale[e] : 1.0));
  }
}

// Affine: thread per element.  Curved: thread per (element, point), so
// high-order elements with many points keep the device busy.
__global__ void k_geometry(int64_t E, int64_t N, int npe, int q, int affine,
                           const double* __restrict__ xyz, const int32_t* __restrict__ conn,
                           const double* __restrict__ dN, const double* __restrict__ w,
                           double* __restrict__ geo, const double* __restrict__ scale) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (affine) {
    if (t >= E) return;
    geometry_one(E, N, npe, q, 1, xyz, conn, dN, w, geo, t, scale);
    return;
  }
  if (t >= E * q) return;
  const int64_t e = t / q;
  const int p = (int)(t - e * q);
  double J[9], Mi[9];
  jacobian(E, N, npe, e, xyz, conn, dN + (size_t)p * npe * 3, J);
  const double det = inv3(J, Mi);
  geometry_store(geo + (size_t)t * kGeoStride, Mi, fabs(det) * w[p] * (scale ? scale[e] : 1.0));
}

// ---------------------------------------------------------------------------
// node-panel kernel
// ---------------------------------------------------------------------------
struct PanelParams {
  const void* rec;     // step records of the launch
  int rec_bytes;
  const int2* desc;    // node descriptors {adj offset, R} of the launch
  int64_t waves;
  int rounds, cnt, max_r;
  const double* geo;
  int64_t E;
  int q;
  const double* tab;   // double table, see llfem_assemble
  const float* tabf;   // FP32 copy of the same table
  double lam, mu, rho;
  double* out;
  int lds_per_wave;    // doubles
  double C[81];        // general constitutive tensor C[d][al][c][be]
  const double* tab2;  // kOpKM: the mass table beside the stiffness table
  double* out2;        // kOpKM: the mass values
  double* sink;        // 64 doubles: where padding lane groups send their panel stores
};

// Stiffness and consistent mass of one mesh in the same pass (batched affine
// path): each step adds the K block and the mass scalar of its node pair, so
// the records, geometry and panel round trips are shared.
constexpr int kOpKM = 2;

// Panel update: plain read-add-write.  No two lanes of a step address the
// same live row; successive steps stay ordered because LDS executes one wave's
// instructions in order and a compiler barrier separates the steps.
__device__ __forceinline__ void lds_add(double* p, double v) { *p += v; }

template <typename Real>
struct Geo {
  Real Mi[9];
  Real wd;
};

template <typename Real>
__device__ __forceinline__ Geo<Real> load_geo(const double* __restrict__ rec) {
  Geo<Real> g;
  if constexpr (sizeof(Real) == 8) {
    const double2* r2 = reinterpret_cast<const double2*>(rec);
    double2 v0 = r2[0], v1 = r2[1], v2 = r2[2], v3 = r2[3], v4 = r2[4];
    g.Mi[0] = v0.x; g.Mi[1] = v0.y; g.Mi[2] = v1.x; g.Mi[3] = v1.y; g.Mi[4] = v2.x;
    g.Mi[5] = v2.y; g.Mi[6] = v3.x; g.Mi[7] = v3.y; g.Mi[8] = v4.x; g.wd = v4.y;
  } else {
    const float4* r4 = reinterpret_cast<const float4*>(rec + 10);
    float4 v0 = r4[0], v1 = r4[1];
    float2 v2 = reinterpret_cast<const float2*>(rec + 10)[4];
    g.Mi[0] = v0.x; g.Mi[1] = v0.y; g.Mi[2] = v0.z; g.Mi[3] = v0.w; g.Mi[4] = v1.x;
    g.Mi[5] = v1.y; g.Mi[6] = v1.z; g.Mi[7] = v1.w; g.Mi[8] = v2.x; g.wd = v2.y;
  }
  return g;
}

template <typename Real>
__device__ __forceinline__ void load9(const double* d, const float* f, Real* o) {
  if constexpr (sizeof(Real) == 8) {
#pragma unroll
    for (int m = 0; m < 9; ++m) o[m] = d[m];
  } else {
#pragma unroll
    for (int m = 0; m < 9; ++m) o[m] = f[m];
  }
}

// G = Minv^T Gh Minv for one block (affine elements), as explicit fused
// multiply-adds (the same rounding in every kernel that uses it).
template <typename Real>
__device__ __forceinline__ void block_G_affine(const Geo<Real>& g, const Real* Gh, Real* G) {
  const Real* Mi = g.Mi;
  Real T[9];   // T[al][k] = sum_be Gh[al][be] Mi[be][k]
#pragma unroll
  for (int al = 0; al < 3; ++al)
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      T[al * 3 + kk] =
          __builtin_fma(Gh[al * 3 + 2], Mi[6 + kk], __builtin_fma(Gh[al * 3 + 1], Mi[3 + kk], Gh[al * 3] * Mi[kk]));
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      G[i * 3 + kk] = __builtin_fma(Mi[6 + i], T[6 + kk], __builtin_fma(Mi[3 + i], T[3 + kk], Mi[i] * T[kk]));
}

// Hooke's law on one block: K[d][c] = lam G[d][c] + mu G[c][d] + mu tr(G)
// delta_dc, all scaled by `scale`.  The products are written as explicit
// fused multiply-adds so every kernel that shares this function rounds the
// same way whatever the compiler's contraction choices around it.
template <typename Real>
__device__ __forceinline__ void hooke_block(Real lam, Real mu, const Real* G, Real scale, Real* K) {
  const Real l = lam * scale, u = mu * scale;
  const Real tr = u * (G[0] + G[4] + G[8]);
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) K[d * 3 + c] = __builtin_fma(l, G[d * 3 + c], u * G[c * 3 + d]);
  K[0] += tr;
  K[4] += tr;
  K[8] += tr;
}

// Isotropic Hooke on an affine element, accumulated into the panel rows:
//   K0 = Minv^T T' Minv with T' = lam Gh + mu Gh^T   (= lam G + mu G^T),
//   tr(K0) = (lam + mu) tr(G), so mu tr(G) = cmu tr(K0), cmu = mu / (lam + mu),
//   dst[c][d] += wd (K0[d][c] + cmu tr(K0) delta_dc).
// Tp is T' (prime_entry below builds it from Gh); the wd scale is folded into
// the accumulating fma.  Written as explicit fused multiply-adds so every
// kernel that shares it rounds the same way.
template <typename Real>
__device__ __forceinline__ Real prime_entry(Real lam, Real mu, const Real* Gh, int k) {
  return __builtin_fma(lam, Gh[k], mu * Gh[(k % 3) * 3 + k / 3]);
}

template <typename Real>
__device__ __forceinline__ void hooke_affine_acc(const Geo<Real>& g, const Real* Tp, Real cmu, const double* old,
                                                 double* out) {
  const Real* Mi = g.Mi;
  Real M[9];   // M[al][k] = sum_be Tp[al][be] Mi[be][k]
#pragma unroll
  for (int al = 0; al < 3; ++al)
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      M[al * 3 + kk] =
          __builtin_fma(Tp[al * 3 + 2], Mi[6 + kk], __builtin_fma(Tp[al * 3 + 1], Mi[3 + kk], Tp[al * 3] * Mi[kk]));
  Real K0[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      K0[i * 3 + kk] = __builtin_fma(Mi[6 + i], M[6 + kk], __builtin_fma(Mi[3 + i], M[3 + kk], Mi[i] * M[kk]));
  const Real t = cmu * (K0[0] + K0[4] + K0[8]);
  K0[0] += t;
  K0[4] += t;
  K0[8] += t;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      out[c * 3 + d] = __builtin_fma((double)g.wd, (double)K0[d * 3 + c], old[c * 3 + d]);
}

// K[d][c] = Ke[(b,d),(a,c)] from G and the material.
template <bool ISO, typename Real>
__device__ __forceinline__ void block_K(const PanelParams& P, const Real* G, Real scale, Real* K) {
  if constexpr (ISO) {
    hooke_block<Real>((Real)P.lam, (Real)P.mu, G, scale, K);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Real acc = 0;
#pragma unroll
        for (int al = 0; al < 3; ++al)
#pragma unroll
          for (int be = 0; be < 3; ++be)
            acc += (Real)P.C[((d * 3 + al) * 3 + c) * 3 + be] * G[al * 3 + be];
        K[d * 3 + c] = acc * scale;
      }
  }
}

template <typename Real>
struct TabBlk {
  Real v[9];
  Real m;      // kOpKM: the mass table entry
};

// Operator-table entry of (a, b): Ghat (9 values) for stiffness, Mhat for mass.
template <int NPE, int OP, bool AFFINE, typename Real>
__device__ __forceinline__ TabBlk<Real> load_tab(const double* tab, const float* tabf, const double* tab2, int a,
                                                 int b) {
  TabBlk<Real> t;
  if constexpr (OP == kOpKM) {
    load9<Real>(tab + (size_t)(a * NPE + b) * 9, tabf + (size_t)(a * NPE + b) * 9, t.v);
    t.m = (Real)tab2[a * NPE + b];
  } else if constexpr (AFFINE && OP == LLFEM_OP_STIFFNESS) {
    load9<Real>(tab + (size_t)(a * NPE + b) * 9, tabf + (size_t)(a * NPE + b) * 9, t.v);
  } else if constexpr (AFFINE) {
    t.v[0] = (Real)tab[a * NPE + b];
  }
  return t;
}

// Add element block Ke[(b,:),(a,:)] to the panel rows of slot s (or, for a
// non-affine element, integrate it point by point first).
// PRIMED: the table entry tb already holds T' = lam Gh + mu Gh^T (the round
// kernels build it once per block in LDS); otherwise it is formed here.
template <int NPE, typename MapT, int OP, bool AFFINE, bool ISO, typename Real, bool PRIMED = false>
__device__ __forceinline__ void contribute(const PanelParams& P, const double* geo, const double* tab,
                                           const TabBlk<Real>& tb, double* __restrict__ wl, int base, int R,
                                           int b, int e, int a, int s, const Geo<Real>& g, int mofs = 0) {
  if constexpr ((OP == LLFEM_OP_STIFFNESS || OP == kOpKM) && AFFINE && ISO) {
    Real Tp[9];
    if constexpr (PRIMED) {
#pragma unroll
      for (int k = 0; k < 9; ++k) Tp[k] = tb.v[k];
    } else {
#pragma unroll
      for (int k = 0; k < 9; ++k) Tp[k] = prime_entry<Real>((Real)P.lam, (Real)P.mu, tb.v, k);
    }
    const Real cmu = (Real)(P.mu / (P.lam + P.mu));
    double* dst = wl + base + 3 * s;
    double old[9], upd[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) old[c * 3 + d] = dst[c * 3 * R + d];
    hooke_affine_acc<Real>(g, Tp, cmu, old, upd);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) dst[c * 3 * R + d] = upd[c * 3 + d];
    if constexpr (OP == kOpKM) lds_add(wl + base + mofs + s, P.rho * (double)g.wd * (double)tb.m);
    __builtin_amdgcn_wave_barrier();
  } else if constexpr (OP == LLFEM_OP_STIFFNESS || OP == kOpKM) {
    Real G[9];
    Real scale;
    if constexpr (AFFINE) {
      block_G_affine<Real>(g, tb.v, G);
      scale = g.wd;
    } else {
#pragma unroll
      for (int m = 0; m < 9; ++m) G[m] = 0;
      scale = 1;
      for (int p = 0; p < P.q; ++p) {
        const Geo<Real> gq = load_geo<Real>(geo + ((size_t)e * P.q + p) * kGeoStride);
        const double* hb = tab + ((size_t)p * NPE + b) * 3;
        const double* ha = tab + ((size_t)p * NPE + a) * 3;
        Real gb[3], ga[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          gb[i] = gq.Mi[i] * (Real)hb[0] + gq.Mi[3 + i] * (Real)hb[1] + gq.Mi[6 + i] * (Real)hb[2];
          ga[i] = (gq.Mi[i] * (Real)ha[0] + gq.Mi[3 + i] * (Real)ha[1] + gq.Mi[6 + i] * (Real)ha[2]) * gq.wd;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int kk = 0; kk < 3; ++kk) G[i * 3 + kk] += gb[i] * ga[kk];
      }
    }
    Real K[9];
    block_K<ISO, Real>(P, G, scale, K);
    double* dst = wl + base + 3 * s;
    double old[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) old[c * 3 + d] = dst[c * 3 * R + d];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) dst[c * 3 * R + d] = old[c * 3 + d] + (double)K[d * 3 + c];
    if constexpr (OP == kOpKM) lds_add(wl + base + mofs + s, P.rho * (double)g.wd * (double)tb.m);
    __builtin_amdgcn_wave_barrier();
  } else {
    double mv;
    if constexpr (AFFINE) {
      mv = P.rho * (double)g.wd * (double)tb.v[0];
    } else {
      mv = 0.0;
      for (int p = 0; p < P.q; ++p)
        mv += geo[((size_t)e * P.q + p) * kGeoStride + 9] * tab[p * NPE + b] * tab[p * NPE + a];
      mv *= P.rho;
    }
    lds_add(wl + base + s, mv);
    __builtin_amdgcn_wave_barrier();
  }
}

// Node-panel kernel over processing-order step records.
//
// A wavefront owns NPW = 64/NPE lane groups; lane (n, b) handles element node
// b.  Lane group n streams `rounds` nodes one after another: in round i it
// accumulates node p = (w*G + i)*NPW + n's panel from `cnt` step records
// (one per incidence, null-padded), then writes the finished panel to HBM
// and clears it.  Records of the NPW groups of one step are adjacent, so each
// step is one coalesced 16-byte load per lane -- no pointer chasing -- and the
// prefetch ring runs straight through round boundaries:
//   rec[v]  step record, fetched 2*PF steps ahead;
//   g[v]    element geometry, fetched PF steps ahead;
//   tb[v]   reference-table block Ghat[a][b], fetched PF steps ahead.
// All ring slot indices are compile-time (the loop is unrolled by 2*PF) and
// every load is unconditional (clamped into the wave's record range), so the
// waits are exact.  The only LDS traffic in a step is the panel update: a
// plain read-add-write (volatile), conflict-free within a step because the
// lanes of a group address distinct rows and LDS executes one wave's
// instructions in order -- sums are deterministic and need no atomics.
constexpr int PF = 2;

struct StepRec {
  int e, a, s;   // element, local index (-1: null step), panel row of node b
};

// Wide step records (rec_bytes per record: int element, int16 local index,
// then npe map entries of MapT), decoded at load.
template <int NPE, typename MapT>
__device__ __forceinline__ StepRec load_rec(const void* recs, int rec_bytes, int64_t r, int b) {
  StepRec o;
  const uint8_t* base = static_cast<const uint8_t*>(recs) + r * rec_bytes;
  const int2 h = *reinterpret_cast<const int2*>(base);
  o.e = h.x;
  o.a = (int)(int16_t)(h.y & 0xffff);
  o.s = (int)reinterpret_cast<const MapT*>(base + 8)[b];
  return o;
}

// Compact 16-byte step records (NPE <= 11, one-byte maps), kept raw in the
// prefetch ring and decoded where they are used, so a record load is waited
// for only when its step comes.
typedef unsigned RawRec __attribute__((ext_vector_type(4)));

__device__ __forceinline__ RawRec load_raw(const void* recs, int64_t r) {
  return __builtin_nontemporal_load(reinterpret_cast<const RawRec*>(recs) + r);
}

__device__ __forceinline__ int raw_elem(const RawRec& v) { return (int)v.x; }

__device__ __forceinline__ int raw_loc(const RawRec& v) {
  const uint32_t a8 = v.y & 0xffu;
  return a8 == 0xffu ? -1 : (int)a8;
}

__device__ __forceinline__ int raw_slot(const RawRec& v, int b) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  const int byte = 5 + b;
  return (int)((w[byte >> 2] >> ((byte & 3) * 8)) & 0xffu);
}

// Ring entry of the generic wave body: compact records stay raw until used,
// others are decoded at load.
template <int NPE, typename MapT, bool RAW = (NPE <= 11 && sizeof(MapT) == 1)>
struct Ring {
  using T = StepRec;
  static __device__ __forceinline__ T load(const void* recs, int rec_bytes, int64_t r, int b) {
    return load_rec<NPE, MapT>(recs, rec_bytes, r, b);
  }
  static __device__ __forceinline__ int e(const T& x) { return x.e; }
  static __device__ __forceinline__ int a(const T& x) { return x.a; }
  static __device__ __forceinline__ int s(const T& x, int) { return x.s; }
};

template <int NPE, typename MapT>
struct Ring<NPE, MapT, true> {
  using T = RawRec;
  static __device__ __forceinline__ T load(const void* recs, int, int64_t r, int) { return load_raw(recs, r); }
  static __device__ __forceinline__ int e(const T& x) { return raw_elem(x); }
  static __device__ __forceinline__ int a(const T& x) { return raw_loc(x); }
  static __device__ __forceinline__ int s(const T& x, int b) { return raw_slot(x, b); }
};

// Waves [w0, w0 + nw) of a panel launch run by one wavefront, LDS panel space
// wl (P.lds_per_wave doubles).  Consecutive waves' records and node
// descriptors are contiguous, so a span is one longer stream of nw * rounds
// rounds.  Uses no block-level barrier.
//
// KS > 0 (stiffness): every panel leaves in exactly KS 64-lane stores
// (KS * 64 >= 9 * max_r; lanes past the panel repeat its last value, padding
// groups write to P.sink), and the next descriptors are fetched before the
// stores, so the descriptor wait no longer drains the store burst (gfx9
// counts stores and loads in one in-order counter).  The round end is still
// conditional here, so the first wait after it covers the stores; the
// step-unrolled panel_wave_rounds below removes that too.
template <int NPE, typename MapT, int OP, bool AFFINE, bool ISO, typename Real, int KS = 0>
__device__ __forceinline__ void panel_wave(const PanelParams& P, int64_t w0, int nw, double* wl) {
  static_assert(KS == 0 || OP == LLFEM_OP_STIFFNESS, "fixed store counts are for stiffness panels");
  constexpr int NPW = 64 / NPE;
  constexpr int BLK = (OP == LLFEM_OP_STIFFNESS) ? 9 : (OP == kOpKM ? 10 : 1);
  const int lane = threadIdx.x & 63;
  double* const dummy = wl + (size_t)NPW * P.max_r * BLK;   // 9 doubles
  const int n = lane / NPE;
  const int b = lane - n * NPE;
  const bool grp = n < NPW;
  const int G = P.rounds * nw, cnt = P.cnt;
  const int64_t nsteps = (int64_t)G * cnt;
  const int64_t rbase = w0 * P.rounds * (int64_t)cnt * NPW;   // first record of the span
  const int64_t rlast = rbase + nsteps * NPW - 1;
  const int64_t lane_rec = grp ? n : 0;
  double* const pan = wl + (size_t)(grp ? n : 0) * P.max_r * BLK;

  for (int i = lane; i < NPW * P.max_r * BLK + 9; i += 64) wl[i] = 0.0;

  // device-memory pointers, cast to the global address space (when P lives in
  // a device-memory table the compiler cannot infer it and would emit flat
  // loads and stores)
  const void* const recs = (const void*)gp(P.rec);
  const double* const geo = (const double*)gp(P.geo);
  const double* const tab = (const double*)gp(P.tab);
  const float* const tabf = (const float*)gp(P.tabf);
  const double* const tab2 = (const double*)gp(P.tab2);
  double* const out = (double*)gp(P.out);
  double* const out2 = (double*)gp(P.out2);
  const int mofs = 9 * P.max_r;   // kOpKM: mass panel after the stiffness panel

  // node descriptors of the current and next round
  const int2* desc = (const int2*)gp(P.desc) + w0 * (int64_t)P.rounds * NPW;
  int2 dcur = desc[lane_rec];
  int2 dnext = desc[(G > 1 ? NPW : 0) + lane_rec];

  using RG = Ring<NPE, MapT>;
  typename RG::T rec[2 * PF];
  Geo<Real> g[PF];
  TabBlk<Real> tb[PF];
#pragma unroll
  for (int v = 0; v < 2 * PF; ++v)
    rec[v] = RG::load(recs, P.rec_bytes, min(rbase + (int64_t)v * NPW + lane_rec, rlast), b);
#pragma unroll
  for (int v = 0; v < PF; ++v) {
    const int a = RG::a(rec[v]);
    g[v] = load_geo<Real>(geo + (size_t)RG::e(rec[v]) * kGeoStride);
    tb[v] = load_tab<NPE, OP, AFFINE, Real>(tab, tabf, tab2, a < 0 ? 0 : a, b);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();

  int t = 0, round = 0;
  for (int64_t s0 = 0; s0 < nsteps; s0 += 2 * PF) {
#pragma unroll
    for (int u = 0; u < 2 * PF; ++u) {
      const int64_t s = s0 + u;
      if (s < nsteps) {
        const int ca = RG::a(rec[u]);
        const bool live = grp && ca >= 0 && dcur.y > 0;
        contribute<NPE, MapT, OP, AFFINE, ISO, Real>(P, geo, tab, tb[u % PF], live ? pan : dummy,
                                                     0, live ? dcur.y : 0, b, RG::e(rec[u]),
                                                     ca < 0 ? 0 : ca, live ? RG::s(rec[u], b) : 0,
                                                     g[u % PF], live ? mofs : 0);
        if (++t == cnt) {   // round complete: write the NPW panels, clear them
          t = 0;
          int2 dfut = make_int2(0, 0);
          if constexpr (KS > 0) dfut = desc[(int64_t)min(round + 2, G - 1) * NPW + lane_rec];
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          __builtin_amdgcn_wave_barrier();
          if constexpr (KS > 0) {
            double* const sink = (double*)gp(P.sink);
#pragma unroll
            for (int m = 0; m < NPW; ++m) {
              const int Rm = __builtin_amdgcn_readlane(dcur.y, m * NPE);
              const int om = __builtin_amdgcn_readlane(dcur.x, m * NPE);
              double* pm = wl + (size_t)m * P.max_r * BLK;
              const bool real = Rm > 0;
              const int last = 9 * Rm - 1;
              double* const o = real ? out + 9 * (int64_t)om : sink;
              double v[KS];
              // all reads before any clear: lanes past the end re-read the last value
#pragma unroll
              for (int k = 0; k < KS; ++k) v[k] = pm[real ? min(lane + 64 * k, last) : 0];
#pragma unroll
              for (int k = 0; k < KS; ++k)
                __builtin_nontemporal_store(v[k], o + (real ? min(lane + 64 * k, last) : lane));
#pragma unroll
              for (int k = 0; k < KS; ++k)
                if (lane + 64 * k <= last) pm[lane + 64 * k] = 0.0;
            }
          } else {
#pragma unroll
          for (int m = 0; m < NPW; ++m) {
            const int Rm = __shfl(dcur.y, m * NPE);
            const int64_t om = (int64_t)__shfl(dcur.x, m * NPE);
            double* pm = wl + (size_t)m * P.max_r * BLK;
            if constexpr (OP == LLFEM_OP_STIFFNESS || OP == kOpKM) {
              double* o = out + 9 * om;
              for (int i = lane; i < 9 * Rm; i += 64) {
                __builtin_nontemporal_store(pm[i], o + i);
                pm[i] = 0.0;
              }
              if constexpr (OP == kOpKM) {
                double* o2 = out2 + 3 * om;
                const double* pmm = pm + mofs;
                for (int i = lane; i < 3 * Rm; i += 64) {
                  const int c = i / Rm;
                  __builtin_nontemporal_store(pmm[i - c * Rm], o2 + i);
                }
              }
            } else {
              double* o = out + 3 * om;
              for (int i = lane; i < 3 * Rm; i += 64) {
                const int c = i / Rm;
                __builtin_nontemporal_store(pm[i - c * Rm], o + i);
              }
            }
          }
          }
          if constexpr (OP == LLFEM_OP_MASS) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int i = lane; i < NPW * P.max_r; i += 64) wl[i] = 0.0;
          } else if constexpr (OP == kOpKM) {   // clear the mass panels once all three copies are out
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int i = lane; i < NPW * P.max_r; i += 64) {
              const int m = i / P.max_r;
              wl[(size_t)m * P.max_r * BLK + mofs + (i - m * P.max_r)] = 0.0;
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          __builtin_amdgcn_wave_barrier();
          ++round;
          dcur = dnext;
          if constexpr (KS > 0)
            dnext = dfut;
          else
            dnext = desc[(int64_t)min(round + 1, G - 1) * NPW + lane_rec];
        }
      }
      // ring refill: geometry/table for step s+PF, record for step s+2PF
      const typename RG::T& nx = rec[(u + PF) % (2 * PF)];
      const int na = RG::a(nx);
      g[u % PF] = load_geo<Real>(geo + (size_t)RG::e(nx) * kGeoStride);
      tb[u % PF] = load_tab<NPE, OP, AFFINE, Real>(tab, tabf, tab2, na < 0 ? 0 : na, b);
      rec[u] = RG::load(recs, P.rec_bytes, min(rbase + (s + 2 * PF) * NPW + lane_rec, rlast), b);
      __builtin_amdgcn_sched_barrier(0);   // keep the step's loads in the step
    }
  }
}

// One round per loop iteration: affine Hooke stiffness for a launch whose
// nodes all take CNT steps (CNT a multiple of the ring length 2 * PFX).  The
// CNT steps are unrolled and the round end -- descriptor fetch, KS stores per
// panel, clears -- is straight-line code at the end of the body, so every
// iteration issues the same vector-memory sequence and the compiler's wait
// counts are exact everywhere: after the stores, the first PFX steps consume
// loads issued before them without waiting for the stores, and the stores
// have PFX steps to complete before the next wait that covers them.
// TL: the reference table sits in LDS (tl, entries of kTabLds<Real> values)
// instead of being fetched from global memory with the geometry.
template <typename Real>
constexpr int kTabLds = sizeof(Real) == 8 ? 10 : 12;   // 16-byte aligned entries

template <int NPE, typename Real, int CNT, int PFX, int KS, bool TL = false>
__device__ __forceinline__ void panel_wave_rounds(const PanelParams& P, int64_t w0, double* wl,
                                                  const Real* tl = nullptr) {
  static_assert(CNT % (2 * PFX) == 0, "the ring must wrap at round ends");
  constexpr int NPW = 64 / NPE;
  constexpr int OP = LLFEM_OP_STIFFNESS;
  constexpr int BLK = 9;
  const int lane = threadIdx.x & 63;
  double* const dummy = wl + (size_t)NPW * P.max_r * BLK;   // 9 doubles
  const int n = lane / NPE;
  const int b = lane - n * NPE;
  const bool grp = n < NPW;
  const int G = P.rounds;
  const int64_t rbase = w0 * G * (int64_t)CNT * NPW;
  const int64_t rlast = rbase + (int64_t)G * CNT * NPW - 1;
  const int64_t lane_rec = grp ? n : 0;
  double* const pan = wl + (size_t)(grp ? n : 0) * P.max_r * BLK;

  for (int i = lane; i < NPW * P.max_r * BLK + 9; i += 64) wl[i] = 0.0;

  const void* const recs = (const void*)gp(P.rec);
  const double* const geo = (const double*)gp(P.geo);
  const double* const tab = (const double*)gp(P.tab);
  const float* const tabf = (const float*)gp(P.tabf);
  double* const out = (double*)gp(P.out);
  double* const sink = (double*)gp(P.sink);

  const int2* desc = (const int2*)gp(P.desc) + w0 * (int64_t)G * NPW;
  int2 dcur = desc[lane_rec];
  int2 dnext = desc[(G > 1 ? NPW : 0) + lane_rec];

  auto tab_at = [&](int a) -> TabBlk<Real> {
    if constexpr (TL) {
      TabBlk<Real> t;
      const Real* e = tl + (a * NPE + b) * kTabLds<Real>;
#pragma unroll
      for (int k = 0; k < 9; ++k) t.v[k] = e[k];
      return t;
    } else {
      return load_tab<NPE, OP, true, Real>(tab, tabf, nullptr, a, b);
    }
  };
  RawRec rec[2 * PFX];
  Geo<Real> g[PFX];
  TabBlk<Real> tb[PFX];
#pragma unroll
  for (int v = 0; v < 2 * PFX; ++v) rec[v] = load_raw(recs, min(rbase + (int64_t)v * NPW + lane_rec, rlast));
#pragma unroll
  for (int v = 0; v < PFX; ++v) {
    const int a = raw_loc(rec[v]);
    g[v] = load_geo<Real>(geo + (size_t)raw_elem(rec[v]) * kGeoStride);
    tb[v] = tab_at(a < 0 ? 0 : a);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();

  for (int r = 0; r < G; ++r) {
    const bool node_live = grp && dcur.y > 0;
#pragma unroll
    for (int u = 0; u < CNT; ++u) {
      const int64_t s = (int64_t)r * CNT + u;
      const RawRec& cr = rec[u % (2 * PFX)];
      const int ca = raw_loc(cr);
      const bool live = node_live && ca >= 0;
      contribute<NPE, uint8_t, OP, true, true, Real, TL>(P, geo, tab, tb[u % PFX], live ? pan : dummy, 0,
                                                     live ? dcur.y : 0, b, raw_elem(cr), ca < 0 ? 0 : ca,
                                                     live ? raw_slot(cr, b) : 0, g[u % PFX], 0);
      // ring refill: geometry/table for step s+PFX, record for step s+2PFX
      const RawRec& nx = rec[(u + PFX) % (2 * PFX)];
      const int na = raw_loc(nx);
      g[u % PFX] = load_geo<Real>(geo + (size_t)raw_elem(nx) * kGeoStride);
      tb[u % PFX] = tab_at(na < 0 ? 0 : na);
      rec[u % (2 * PFX)] = load_raw(recs, min(rbase + (s + 2 * PFX) * NPW + lane_rec, rlast));
      // keep each step's loads in their step: the scheduler would otherwise
      // sink them towards their uses and shorten the prefetch distance
      __builtin_amdgcn_sched_barrier(0);
    }
    // round end: the descriptors of round r+2 first, then KS stores per panel
    const int2 dfut = desc[(int64_t)min(r + 2, G - 1) * NPW + lane_rec];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int m = 0; m < NPW; ++m) {
      const int Rm = __builtin_amdgcn_readlane(dcur.y, m * NPE);
      const int om = __builtin_amdgcn_readlane(dcur.x, m * NPE);
      double* pm = wl + (size_t)m * P.max_r * BLK;
      const bool real = Rm > 0;
      const int last = 9 * Rm - 1;
      double* const o = real ? out + 9 * (int64_t)om : sink;
      // each store covers one 512-byte-aligned window, so 64-byte sectors
      // are split between two stores only at the ends of a panel
      const int shift = real ? (int)(((uintptr_t)o >> 3) & 7) : 0;
      double v[KS];
#pragma unroll
      for (int k = 0; k < KS; ++k) v[k] = pm[real ? min(max(lane + 64 * k - shift, 0), last) : 0];
#pragma unroll
      for (int k = 0; k < KS; ++k)
        __builtin_nontemporal_store(v[k], o + (real ? min(max(lane + 64 * k - shift, 0), last) : lane));
#pragma unroll
      for (int k = 0; k < KS; ++k) {
        const int i = lane + 64 * k - shift;
        if (i >= 0 && i <= last) pm[i] = 0.0;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    dcur = dnext;
    dnext = dfut;
  }
}

template <int NPE, typename Real, int CNT, int PFX, int KS, bool TL = false>
__global__ __launch_bounds__(512) void k_panel_rounds(PanelParams P) {
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6;
  Real* tl = nullptr;
  if constexpr (TL) {   // the table after the waves' panels (16-byte aligned); every wave reaches the barrier
    tl = reinterpret_cast<Real*>(lds + (((size_t)(blockDim.x >> 6) * P.lds_per_wave + 1) & ~(size_t)1));
    const double* tab = (const double*)gp(P.tab);
    const float* tabf = (const float*)gp(P.tabf);
    // T' = lam Gh + mu Gh^T per (a, b) entry (isotropic stiffness:
    // contribute<..., PRIMED>), one entry per thread
    constexpr int TS = kTabLds<Real>;
    for (int ent = threadIdx.x; ent < NPE * NPE; ent += blockDim.x) {
      Real gh[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        if constexpr (sizeof(Real) == 8) gh[m] = tab[ent * 9 + m];
        else gh[m] = tabf[ent * 9 + m];
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) tl[ent * TS + k] = prime_entry<Real>((Real)P.lam, (Real)P.mu, gh, k);
#pragma unroll
      for (int k = 9; k < TS; ++k) tl[ent * TS + k] = 0;
    }
    __syncthreads();
  }
  // XCD-aware work order, as in k_panel
  const int64_t per = gridDim.x / kXcds;
  const int64_t lb = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  const int64_t w = lb * (blockDim.x >> 6) + wave;
  if (w >= P.waves) return;
  panel_wave_rounds<NPE, Real, CNT, PFX, KS, TL>(P, w, lds + (size_t)wave * P.lds_per_wave, tl);
}

template <int NPE, typename MapT, int OP, bool AFFINE, bool ISO, typename Real, int KS = 0>
__global__ __launch_bounds__(ISO ? 512 : 256) void k_panel(PanelParams P) {
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6;
  // XCD-aware work order: the dispatcher deals blocks round-robin to the 8
  // XCDs, so block b runs on XCD b % 8.  Give XCD x the contiguous block range
  // [x*per, (x+1)*per) of the processing order; neighbouring node groups then
  // share an L2 and re-read element geometry from it rather than from HBM.
  // The host pads the grid to a multiple of kXcds, which makes the remap a
  // bijection of [0, gridDim.x); padding blocks fail the w < waves test.
  const int64_t per = gridDim.x / kXcds;
  const int64_t lb = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;   // logical block
  const int64_t w = lb * (blockDim.x >> 6) + wave;
  if (w >= P.waves) return;                       // no block-level barriers below
  panel_wave<NPE, MapT, OP, AFFINE, ISO, Real, KS>(P, w, 1, lds + (size_t)wave * P.lds_per_wave);
}

// lumped mass: row sums of the consistent mass panels (src/linear.jl:1999-2001)
__global__ void k_lump(int64_t N, const int64_t* __restrict__ adj_ptr, const double* __restrict__ M,
                       double* __restrict__ out) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int64_t a0 = adj_ptr[j];
  const int R = (int)(adj_ptr[j + 1] - a0);
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int i = 0; i < R; ++i) s += M[3 * a0 + (int64_t)c * R + i];
    out[3 * j + c] = s;
  }
}

}  // namespace
}  // namespace llfem
