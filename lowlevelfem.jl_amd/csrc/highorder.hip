// High-order hexahedra (order >= 3: hex64, hex125) on the matrix cores.
//
// Reference: the same B'DB element integration (src/linear.jl:272-307,
// src/multifield.jl:1857-2052), which for n = 125 nodes and q = 125 points is
// a dense (3n x 6q)(6q x 3n) product per element.  Here it is computed as
// the Gram matrix of the weighted physical gradients:
//   Z[(b,al), p] = dN_b/dx_al (xi_p) * sqrt(w_p |det J_p|)      (3n x q)
//   G = Z Z^T,   K_ba[d][c] = lam G_ba[d][c] + mu G_ba[c][d] + mu tr(G_ba) delta_dc
// which is the same quadrature sum regrouped (2*(3n)^2*q flops per element
// instead of 2*(3n)^2*6q).  G runs on MFMA:
//   FP64    v_mfma_f64_16x16x4f64,
//   BF16x3  v_mfma_f32_16x16x16bf16_1k with Z split into bf16 hi + lo parts,
//           G ~ Zh Zh^T + Zh Zl^T + Zl Zh^T, FP32 accumulation.
// Element blocks are staged in FP64 and summed into the CSC by a
// deterministic segmented gather (contributions in ascending element order).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "common.h"

namespace llfem {

constexpr int kHoTile = 96;     // G tile: 96 rows = 32 nodes x 3 components
constexpr int kHoMaxRows = 384; // padded 3n: n <= 128
constexpr int kHoMaxQ = 128;    // padded number of points
constexpr int kGeoRec = 16;     // doubles per geometry record written by k_geometry
constexpr int kHoXcds = 8;      // XCDs of the device (blocks are dealt round-robin)

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rn(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(short h) {
  return __uint_as_float(((uint32_t)(uint16_t)h) << 16);
}

// Z for one (element, point); rows = padded 3n (multiple of 96), qp = padded
// q (multiple of 16).  Layouts:
//   FP64  Z[e][p/4][row][p%4]           (k-blocks of 4: one MFMA fragment = 512 contiguous B)
//   BF16  Zh/Zl[e][p/16][row][p%16]     (k-blocks of 16)
// Padding rows (>= 3n) and points (>= q) are written as zero.
// Row position b holds local node perm[b] (the element's nodes in ascending
// global node id, so the Gram epilogue's neighbouring lanes write
// neighbouring CSC slots).
template <bool BF16>
__global__ void k_hex_z(int64_t E, int npe, int q, int rows, int affine, const double* __restrict__ geo,
                        const double* __restrict__ dN, const double* __restrict__ w,
                        const int32_t* __restrict__ perm,
                        double* __restrict__ Z, short* __restrict__ Zh, short* __restrict__ Zl) {
  const int64_t e = blockIdx.x;
  const int p = threadIdx.x;           // point (blockDim = qp)
  const int qp = blockDim.x;
  double Mi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double sw = 0.0;
  if (p < q) {
    const double* g = geo + (affine ? (size_t)e : ((size_t)e * q + p)) * kGeoRec;
#pragma unroll
    for (int m = 0; m < 9; ++m) Mi[m] = g[m];
    sw = sqrt(affine ? g[9] * w[p] : g[9]);
  }
  for (int b = 0; b < rows / 3; ++b) {
    double z[3] = {0.0, 0.0, 0.0};
    if (p < q && b < npe) {
      const double* h = dN + ((size_t)p * npe + perm[b]) * 3;
#pragma unroll
      for (int al = 0; al < 3; ++al) z[al] = sw * (Mi[al] * h[0] + Mi[3 + al] * h[1] + Mi[6 + al] * h[2]);
    }
#pragma unroll
    for (int al = 0; al < 3; ++al) {
      const int row = 3 * b + al;
      if constexpr (!BF16) {
        Z[(((size_t)e * (qp / 4) + p / 4) * rows + row) * 4 + (p & 3)] = z[al];
      } else {
        const float f = (float)z[al];
        const short hi = bf16_rn(f);
        const short lo = bf16_rn((float)(z[al] - (double)bf16_to_f(hi)));
        const size_t o = (((size_t)e * (qp / 16) + p / 16) * rows + row) * 16 + (p & 15);
        Zh[o] = hi;
        Zl[o] = lo;
      }
    }
  }
}

struct GramParams {
  int64_t E;
  int npe;
  int ntiles;             // node tiles of 32 (rows = 96 * ntiles)
  int npairs;             // upper tile pairs per element
  int qp;                 // padded points
  const double* Z;
  const short* Zh;
  const short* Zl;
  double lam, mu;
  double C[81];
  int iso;
  double* S;              // staged blocks S[e][col node a][row node b][c][d]
  // a block that is the only contribution to its CSC block goes straight to
  // the CSC values (direct[staged index] = CSC block, -1: staged)
  const int32_t* direct;
  const int32_t* perm;    // [npe] local node of each row position
  const int32_t* blk_node;
  const int64_t* adj_ptr;
  double* out;
};

// Upper-triangle tile pair index -> (ti, tj), ti <= tj.
__device__ __forceinline__ void tile_pair(int idx, int nt, int& ti, int& tj) {
  ti = 0;
  while (idx >= nt - ti) { idx -= nt - ti; ++ti; }
  tj = ti + idx;
}

// Block -> (element, tile pair).  Hardware deals consecutive blocks round-robin
// over the 8 XCDs; the remap gives each XCD a contiguous run of (element,
// pair) ids so all pairs of an element share one L2 (Z is fetched from HBM
// once per element).  The grid is padded to a multiple of 8.
__device__ __forceinline__ bool gram_block(const GramParams& P, int64_t& e, int& ti, int& tj) {
  const unsigned L = blockIdx.x, nb = gridDim.x;
  const unsigned id = (L % kHoXcds) * (nb / kHoXcds) + L / kHoXcds;
  e = id / P.npairs;
  if (e >= P.E) return false;
  tile_pair((int)(id - e * P.npairs), P.ntiles, ti, tj);
  return true;
}

// Epilogue: LDS holds rows [48h, 48h+48) of the 96x96 G tile (row nodes
// 16h..16h+15 of tile ti, all 32 column nodes of tile tj) -> staged K blocks.
__device__ void gram_epilogue(const GramParams& P, const double* Gt, int64_t e, int ti, int tj, int h) {
  const int npe = P.npe;
  double* S = P.S + (size_t)e * npe * npe * 9;
  for (int pass = 0; pass < (ti == tj ? 1 : 2); ++pass) {
    for (int idx = threadIdx.x; idx < 16 * 32; idx += blockDim.x) {
      // pass 0: block (row b in ti, col a in tj), b fastest across threads;
      // pass 1: its transpose (row a in tj, col b in ti), a fastest.
      const int hl = pass == 0 ? (idx & 15) : (idx >> 5);   // row node within the half
      const int al = pass == 0 ? (idx >> 4) : (idx & 31);   // col node in tile tj
      const int b = ti * 32 + 16 * h + hl, a = tj * 32 + al;
      if (b >= npe || a >= npe) continue;
      // G_ba[x][y] of (row node b, col node a); pass 1 needs G_ab = G_ba^T
      double G[9];
#pragma unroll
      for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) {
          const double g = Gt[(3 * hl + x) * kHoTile + 3 * al + y];
          if (pass == 0) G[x * 3 + y] = g; else G[y * 3 + x] = g;
        }
      double K[9];
      if (P.iso) {
        const double tr = P.mu * (G[0] + G[4] + G[8]);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int c = 0; c < 3; ++c) K[d * 3 + c] = P.lam * G[d * 3 + c] + P.mu * G[c * 3 + d];
        K[0] += tr; K[4] += tr; K[8] += tr;
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
              for (int y = 0; y < 3; ++y) acc += P.C[((d * 3 + x) * 3 + c) * 3 + y] * G[x * 3 + y];
            K[d * 3 + c] = acc;
          }
      }
      // block of column node cn, row node rn (local nodes): [c][d] = K[d][c]
      const int bl = P.perm[b], al_ = P.perm[a];
      const int cn = pass == 0 ? al_ : bl, rn = pass == 0 ? bl : al_;
      const int32_t p = P.direct[((size_t)e * npe + cn) * npe + rn];
      if (p >= 0) {   // CSC panel of node j: value (c, slot s, d) at 9 a0 + c 3R + 3 s + d
        const int64_t j = P.blk_node[p];
        const int64_t a0 = P.adj_ptr[j];
        const int64_t R = P.adj_ptr[j + 1] - a0;
        double* o = P.out + 9 * a0 + 3 * (p - a0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) o[c * 3 * R + d] = K[d * 3 + c];
      } else {
        double* o = S + ((size_t)cn * npe + rn) * 9;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) o[c * 3 + d] = K[d * 3 + c];
      }
    }
  }
}

// One block per (element, upper tile pair); 4 waves, each a 3x3 grid of 16x16
// subtiles of the 96x96 tile (wave w: rows 48*(w/2).., cols 48*(w%2)..).
// Fragments of step kb+1 are loaded while the MFMAs of step kb run.  The tile
// goes through LDS in two 48-row halves (36 KB), so four blocks fit a CU.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_gram_f64(GramParams P) {
  extern __shared__ double Gt[];   // 48 x 96
  int64_t e;
  int ti, tj;
  if (!gram_block(P, e, ti, tj)) return;   // block-uniform
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rows = P.ntiles * kHoTile;
  const int lr = lane & 15, lk = lane >> 4;
  const double* Za = P.Z + (size_t)e * P.qp * rows + (size_t)(ti * kHoTile + (wave >> 1) * 48 + lr) * 4 + lk;
  const double* Zb = P.Z + (size_t)e * P.qp * rows + (size_t)(tj * kHoTile + (wave & 1) * 48 + lr) * 4 + lk;
  const size_t kstride = (size_t)rows * 4;
  const int nk = P.qp / 4;   // even: qp is a multiple of 16
  d4 acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double a0[3], b0[3], a1[3], b1[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) { a0[i] = Za[64 * i]; b0[i] = Zb[64 * i]; }
  for (int kb = 0; kb < nk; kb += 2) {
    const size_t o1 = (size_t)(kb + 1) * kstride;
    const size_t o2 = (size_t)(kb + 2 < nk ? kb + 2 : nk - 1) * kstride;
#pragma unroll
    for (int i = 0; i < 3; ++i) { a1[i] = Za[o1 + 64 * i]; b1[i] = Zb[o1 + 64 * i]; }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i], b0[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 3; ++i) { a0[i] = Za[o2 + 64 * i]; b0[i] = Zb[o2 + 64 * i]; }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i], b1[j], acc[i][j], 0, 0, 0);
  }
  // f64 16x16x4 output layout: lane l, reg r -> row (l/16) + 4r, col l%16
  const int tc0 = (wave & 1) * 48;
  for (int h = 0; h < 2; ++h) {
    if ((wave >> 1) == h) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            Gt[(16 * i + lk + 4 * r) * kHoTile + tc0 + 16 * j + lr] = acc[i][j][r];
    }
    __syncthreads();
    gram_epilogue(P, Gt, e, ti, tj, h);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_gram_bf16x3(GramParams P) {
  extern __shared__ double Gt[];   // 48 x 96
  int64_t e;
  int ti, tj;
  if (!gram_block(P, e, ti, tj)) return;   // block-uniform
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rows = P.ntiles * kHoTile;
  const int lr = lane & 15, lk = lane >> 4;
  const size_t eo = (size_t)e * P.qp * rows;
  const size_t oa = eo + (size_t)(ti * kHoTile + (wave >> 1) * 48 + lr) * 16 + 4 * lk;
  const size_t ob = eo + (size_t)(tj * kHoTile + (wave & 1) * 48 + lr) * 16 + 4 * lk;
  const s4* Ah = reinterpret_cast<const s4*>(P.Zh + oa);
  const s4* Al = reinterpret_cast<const s4*>(P.Zl + oa);
  const s4* Bh = reinterpret_cast<const s4*>(P.Zh + ob);
  const s4* Bl = reinterpret_cast<const s4*>(P.Zl + ob);
  const size_t kstride = (size_t)rows * 4;   // in s4 units (16 shorts per row per k-block)
  const int nk = P.qp / 16;
  f4 acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  s4 ah[2][3], al[2][3], bh[2][3], bl[2][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ah[0][i] = Ah[64 * i]; al[0][i] = Al[64 * i];
    bh[0][i] = Bh[64 * i]; bl[0][i] = Bl[64 * i];
  }
  for (int kb = 0; kb < nk; ++kb) {
    // nk = qp/16 can be odd: the buffer index is kb & 1, resolved by unrolling two steps
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (u == 1) { if (++kb >= nk) break; }
      const int cur = u, nxt = u ^ 1;
      const size_t on = (size_t)(kb + 1 < nk ? kb + 1 : nk - 1) * kstride;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        ah[nxt][i] = Ah[on + 64 * i]; al[nxt][i] = Al[on + 64 * i];
        bh[nxt][i] = Bh[on + 64 * i]; bl[nxt][i] = Bl[on + 64 * i];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[cur][i], bh[cur][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[cur][i], bl[cur][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[cur][i], bh[cur][j], acc[i][j], 0, 0, 0);
        }
    }
  }
  // bf16 16x16x16 output layout: lane l, reg r -> row 4(l/16) + r, col l%16
  const int tc0 = (wave & 1) * 48;
  for (int h = 0; h < 2; ++h) {
    if ((wave >> 1) == h) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            Gt[(16 * i + 4 * lk + r) * kHoTile + tc0 + 16 * j + lr] = (double)acc[i][j][r];
    }
    __syncthreads();
    gram_epilogue(P, Gt, e, ti, tj, h);
    __syncthreads();
  }
}

// One thread per node-pair block p: sum its staged element blocks in ascending
// element order and write the 3x3 block into the node panel.
// CSC blocks with more than one contribution: sum the staged blocks in
// ascending element order (the single-contribution ones were written by the
// Gram epilogue).
__global__ void k_stage_gather(int64_t nmulti, const int32_t* __restrict__ multi, const int32_t* __restrict__ blk_node,
                               const int64_t* __restrict__ adj_ptr, const int64_t* __restrict__ cptr,
                               const int32_t* __restrict__ csrc, const double* __restrict__ S,
                               double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nmulti) return;
  const int64_t p = multi[k];
  const int64_t j = blk_node[p];
  const int64_t a0 = adj_ptr[j];
  const int64_t R = adj_ptr[j + 1] - a0;
  const int64_t s = p - a0;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t c = cptr[p]; c < cptr[p + 1]; ++c) {
    const double* blk = S + (size_t)csrc[c] * 9;
#pragma unroll
    for (int m = 0; m < 9; ++m) acc[m] += blk[m];
  }
  double* o = out + 9 * a0 + 3 * s;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) o[c * 3 * R + d] = acc[c * 3 + d];
}

// contribution keys: (block p << 32) | staged-block index e*npe^2 + a*npe + b,
// generated per (incidence, element node b)
__global__ void k_contrib_keys(int64_t ninc, int npe, const int32_t* __restrict__ inc_ptr_node,
                               const int32_t* __restrict__ inc_elem, const uint8_t* __restrict__ inc_loc,
                               const uint8_t* __restrict__ map8, const uint16_t* __restrict__ map16,
                               const int64_t* __restrict__ adj_ptr, uint64_t* __restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ninc * npe) return;
  const int64_t k = t / npe;
  const int b = (int)(t - k * npe);
  const int64_t j = inc_ptr_node[k];
  const int64_t e = inc_elem[k];
  const int a = inc_loc[k];
  const int s = map8 ? map8[t] : map16[t];
  const uint64_t pblk = (uint64_t)(adj_ptr[j] + s);
  const uint64_t src = (uint64_t)((e * npe + a) * npe + b);
  keys[t] = (pblk << 32) | src;
}

__global__ void k_node_of_incidence(int64_t N, const int32_t* __restrict__ inc_ptr, int32_t* __restrict__ node) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  for (int32_t k = inc_ptr[j]; k < inc_ptr[j + 1]; ++k) node[k] = (int32_t)j;
}

__global__ void k_blk_node(int64_t N, const int64_t* __restrict__ adj_ptr, int32_t* __restrict__ bn) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  for (int64_t p = adj_ptr[j]; p < adj_ptr[j + 1]; ++p) bn[p] = (int32_t)j;
}

__global__ void k_direct_map(int64_t nblk, const int64_t* __restrict__ cptr, const int32_t* __restrict__ csrc,
                             int32_t* __restrict__ direct, uint8_t* __restrict__ multi_flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nblk) return;
  const int64_t n = cptr[p + 1] - cptr[p];
  if (n == 1) direct[csrc[cptr[p]]] = (int32_t)p;
  multi_flag[p] = n > 1 ? 1 : 0;
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

__global__ void k_split_contrib(int64_t n, const uint64_t* __restrict__ keys, int32_t* __restrict__ src,
                                int64_t* __restrict__ ptr, int64_t nblk) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n) return;
  if (t < n) src[t] = (int32_t)(uint32_t)keys[t];
  const int64_t cur = t < n ? (int64_t)(keys[t] >> 32) : nblk;
  const int64_t prev = t > 0 ? (int64_t)(keys[t - 1] >> 32) : -1;
  for (int64_t v = prev + 1; v <= cur; ++v) ptr[v] = t;
}

}  // namespace

// Contribution lists for the staged gather, built once per mesh.
int build_contrib(llfem_mesh* m) {
  if (m->d_cptr) return LLFEM_OK;
  llfem_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = m->ninc * m->npe;
  if (m->E * (int64_t)m->npe * m->npe >= ((int64_t)1 << 31))
    return fail(LLFEM_EUNSUP, "staged assembly: E*npe^2 >= 2^31");
  int32_t* node = nullptr;
  uint64_t *keys = nullptr, *tmp = nullptr;
  LLFEM_TRY(dalloc(ctx, &node, (size_t)m->ninc));
  LLFEM_TRY(dalloc(ctx, &keys, (size_t)n));
  LLFEM_TRY(dalloc(ctx, &tmp, (size_t)n));
  k_node_of_incidence<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, m->d_inc_ptr, node);
  k_contrib_keys<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
      m->ninc, m->npe, node, m->d_inc_elem, m->d_inc_loc,
      m->map_bytes == 1 ? (const uint8_t*)m->d_inc_map : nullptr,
      m->map_bytes == 2 ? (const uint16_t*)m->d_inc_map : nullptr, m->d_adj_ptr, keys);
  size_t tb = 0;
  int end_bit = 64;
  LLFEM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys, tmp, (int)n, 0, end_bit, st));
  void* d_tmp = nullptr;
  LLFEM_TRY(dmalloc(ctx, &d_tmp, tb));
  LLFEM_HIP(hipcub::DeviceRadixSort::SortKeys(d_tmp, tb, keys, tmp, (int)n, 0, end_bit, st));
  dfree(ctx, d_tmp, tb);
  LLFEM_TRY(dalloc(ctx, &m->d_csrc, (size_t)n));
  LLFEM_TRY(dalloc(ctx, &m->d_cptr, (size_t)m->nadj + 1));
  LLFEM_TRY(dalloc(ctx, &m->d_blk_node, (size_t)m->nadj));
  k_split_contrib<<<(unsigned)((n + 256) / 256), 256, 0, st>>>(n, tmp, m->d_csrc, m->d_cptr, m->nadj);
  k_blk_node<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, m->d_adj_ptr, m->d_blk_node);
  // direct-write map of single-contribution blocks, list of the others
  const int64_t nstaged = m->E * (int64_t)m->npe * m->npe;
  uint8_t* flags = nullptr;
  int64_t* d_num = nullptr;
  LLFEM_TRY(dalloc(ctx, &m->d_direct, (size_t)nstaged));
  LLFEM_TRY(dalloc(ctx, &flags, (size_t)m->nadj));
  LLFEM_TRY(dalloc(ctx, &m->d_multi, (size_t)std::max<int64_t>(m->nadj, 1)));
  LLFEM_TRY(dalloc(ctx, &d_num, 1));
  k_fill_i32<<<(unsigned)((nstaged + 255) / 256), 256, 0, st>>>(nstaged, -1, m->d_direct);
  k_direct_map<<<(unsigned)((m->nadj + 255) / 256), 256, 0, st>>>(m->nadj, m->d_cptr, m->d_csrc, m->d_direct, flags);
  hipcub::CountingInputIterator<int32_t> ids(0);
  size_t sb = 0;
  LLFEM_HIP(hipcub::DeviceSelect::Flagged(nullptr, sb, ids, flags, m->d_multi, d_num, (int)m->nadj, st));
  void* d_sel = nullptr;
  LLFEM_TRY(dmalloc(ctx, &d_sel, sb));
  LLFEM_HIP(hipcub::DeviceSelect::Flagged(d_sel, sb, ids, flags, m->d_multi, d_num, (int)m->nadj, st));
  int64_t nm = 0;
  LLFEM_HIP(hipMemcpyAsync(&nm, d_num, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipGetLastError());
  LLFEM_HIP(hipStreamSynchronize(st));
  dfree(ctx, d_sel, sb);
  drelease(ctx, flags, (size_t)m->nadj);
  drelease(ctx, d_num, 1);
  drelease(ctx, node, (size_t)m->ninc);
  drelease(ctx, keys, (size_t)n);
  drelease(ctx, tmp, (size_t)n);
  m->nmulti = nm;
  m->ncontrib = n;
  return LLFEM_OK;
}

// Stiffness of order >= 3 hexahedra (positive quadrature weights).  The
// geometry records (k_geometry) and the dN / w tables are on the device.
// ctx->ev[1] / ev[2] bracket the Gram kernel.
int assemble_highorder(llfem_mesh* m, const llfem_op_desc* d, bool affine, const double* d_dN,
                       const double* d_w, double lam, double mu, const double* C81, double* out) {
  llfem_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int npe = m->npe, q = d->q;
  if (3 * npe > kHoMaxRows || q > kHoMaxQ)
    return fail(LLFEM_EUNSUP, "high-order path supports npe <= 128 and q <= 128");
  for (int p = 0; p < q; ++p)
    if (!(d->weights[p] > 0.0))
      return fail(LLFEM_EUNSUP, "high-order path needs positive quadrature weights");
  LLFEM_TRY(build_contrib(m));
  const bool bf = d->precision == LLFEM_BF16X3;
  const int ntiles = (3 * npe + kHoTile - 1) / kHoTile;
  const int rows = ntiles * kHoTile;
  const int qp = (q + 15) / 16 * 16;
  const size_t zd = (size_t)m->E * qp * rows;   // values per Z copy
  const size_t sd = (size_t)m->E * npe * npe * 9;
  const size_t need = (bf ? 2 * zd * sizeof(short) : zd * sizeof(double)) + sd * sizeof(double);
  if (m->ho_bytes < need) {
    if (m->d_ho) dfree(ctx, m->d_ho, m->ho_bytes);
    m->d_ho = nullptr;
    m->ho_bytes = 0;
    LLFEM_TRY(dmalloc(ctx, &m->d_ho, need));
    m->ho_bytes = need;
  }
  char* base = static_cast<char*>(m->d_ho);
  double* S = reinterpret_cast<double*>(base);
  double* Z = reinterpret_cast<double*>(base + sd * sizeof(double));
  short* Zh = reinterpret_cast<short*>(Z);
  short* Zl = Zh + zd;
  if (!m->d_ho_perm) {   // element 0's nodes in ascending global id
    std::vector<int32_t> perm(npe);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t x, int32_t y) { return m->h_conn0[x] < m->h_conn0[y]; });
    LLFEM_TRY(dalloc(ctx, &m->d_ho_perm, (size_t)npe));
    LLFEM_HIP(hipMemcpyAsync(m->d_ho_perm, perm.data(), sizeof(int32_t) * npe, hipMemcpyHostToDevice, st));
  }
  if (bf)
    k_hex_z<true><<<(unsigned)m->E, qp, 0, st>>>(m->E, npe, q, rows, affine ? 1 : 0, m->d_geo, d_dN, d_w,
                                                 m->d_ho_perm, nullptr, Zh, Zl);
  else
    k_hex_z<false><<<(unsigned)m->E, qp, 0, st>>>(m->E, npe, q, rows, affine ? 1 : 0, m->d_geo, d_dN, d_w,
                                                  m->d_ho_perm, Z, nullptr, nullptr);
  GramParams P{};
  P.E = m->E;
  P.npe = npe;
  P.ntiles = ntiles;
  P.npairs = ntiles * (ntiles + 1) / 2;
  P.qp = qp;
  P.Z = Z;
  P.Zh = Zh;
  P.Zl = Zl;
  P.lam = lam;
  P.mu = mu;
  P.iso = C81 == nullptr;
  if (C81) std::memcpy(P.C, C81, sizeof(P.C));
  P.S = S;
  P.direct = m->d_direct;
  P.perm = m->d_ho_perm;
  P.blk_node = m->d_blk_node;
  P.adj_ptr = m->d_adj_ptr;
  P.out = out;
  const size_t lds = sizeof(double) * 48 * kHoTile;
  auto fn = bf ? k_gram_bf16x3 : k_gram_f64;
  LLFEM_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int64_t nblocks = (m->E * P.npairs + kHoXcds - 1) / kHoXcds * kHoXcds;
  if (nblocks >= ((int64_t)1 << 31)) return fail(LLFEM_EUNSUP, "high-order path: too many elements");
  const dim3 grid((unsigned)nblocks);
  LLFEM_HIP(hipEventRecord(ctx->ev[1], st));
  hipLaunchKernelGGL(fn, grid, dim3(256), lds, st, P);
  LLFEM_HIP(hipEventRecord(ctx->ev[2], st));
  if (m->nmulti > 0)
    k_stage_gather<<<(unsigned)((m->nmulti + 255) / 256), 256, 0, st>>>(m->nmulti, m->d_multi, m->d_blk_node,
                                                                         m->d_adj_ptr, m->d_cptr, m->d_csrc, S, out);
  LLFEM_HIP(hipGetLastError());
  return LLFEM_OK;
}

}  // namespace llfem
