// High-order hexahedra (order >= 3: hex64, hex125) on the matrix cores.
//
// Reference: the same B'DB element integration (src/linear.jl:272-307,
// src/multifield.jl:1857-2052), which for n = 125 nodes and q = 125 points is
// a dense (3n x 6q)(6q x 3n) product per element.  Here it is computed as
// the Gram matrix of the weighted physical gradients:
//   Z[(b,al), p] = dN_b/dx_al (xi_p) * sqrt(w_p |det J_p|)      (3n x q)
//   G = Z Z^T,   K_ba[d][c] = lam G_ba[d][c] + mu G_ba[c][d] + mu tr(G_ba) delta_dc
// which is the same quadrature sum regrouped (2*(3n)^2*q flops per element
// instead of 2*(3n)^2*6q).  Z is never stored: each Gram block builds the Z
// rows of its two tiles in LDS, 16 points at a time, from the element's
// geometry and the reference gradients.  G runs on MFMA:
//   FP64    v_mfma_f64_16x16x4f64,
//   BF16x3  v_mfma_f32_16x16x16bf16_1k with Z split into bf16 hi + lo parts,
//           G ~ Zh Zh^T + Zh Zl^T + Zl Zh^T, FP32 accumulation.
// An element block that is the only contribution to its CSC block is written
// by the Gram epilogue straight into the column's panel; the others are
// staged in FP64, contiguous per CSC block in ascending element order, and
// summed by a gather.  The epilogue's destinations come from per-mesh tables
// indexed by the Gram row positions (no dependent index loads per block).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>

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

// Reference gradients in Gram row-position order, padded: dN_{perm[pos]}/dxi_c
// at point p, zero for padding points (p >= q) and positions (pos >= npe),
// stored [p][pos][c] (points_major) or [pos][p][c].  Row position pos holds
// local node perm[pos] (the element's nodes in ascending global node id, so
// the Gram epilogue's neighbouring lanes write neighbouring CSC slots).
__global__ void k_dn_positions(int q, int qp, int npe, int npos, int points_major, const int32_t* __restrict__ perm,
                               const double* __restrict__ dN, double* __restrict__ dNp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)qp * npos * 3) return;
  const int c = (int)(i % 3);
  const int pos = points_major ? (int)((i / 3) % npos) : (int)(i / (3 * qp));
  const int p = points_major ? (int)(i / (3 * npos)) : (int)((i / 3) % qp);
  dNp[i] = (p < q && pos < npe) ? dN[((size_t)p * npe + perm[pos]) * 3 + c] : 0.0;
}

constexpr int kGramThreads = 256;   // 4 waves; the epilogue deals 2 x 256 blocks per pass
// Row stride (doubles) of the G half-tile in LDS.  The epilogue's lanes read
// rows 3 apart (pass 0) or columns 3 apart (pass 1); with 100 = 4 (mod 32)
// both patterns touch each 8-byte bank pair at most twice per wavefront.
constexpr int kGtStride = 100;

struct GramParams {
  int64_t E;
  int npe;
  int ntiles;             // node tiles of 32 (rows = 96 * ntiles)
  int npairs;             // upper tile pairs per element
  int q, qp;              // points, padded to a multiple of 16
  int npos;               // row positions (32 per tile)
  int affine;             // one geometry record per element (else per point)
  const double* geo;      // geometry records: inverse Jacobian [9], |det J| (affine) or w |det J|
  const double* w;        // [q] quadrature weights
  const double* dNp;      // reference gradients by row position: FP64 [qp][npos][3], BF16 [npos][qp][3]
  double lam, mu;
  const double* C;        // general material [d][x][c][y] in device memory (NULL: isotropic lam, mu)
  double* S;              // staged blocks [k][c][d]
  const int32_t* gslot;   // [E][col pos][row pos]: >= 0 offset in the column panel, < 0: -(k+1) staged
  const int64_t* gcol;    // [E][col pos][2]: 9 a0 and 3 R of the column node's panel
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

// Column references of a tile pair into LDS (read after the first barrier):
// slots 0..31 the positions of tile tj, 32..63 those of tile ti.
__device__ __forceinline__ void load_colref(const GramParams& P, int64_t e, int ti, int tj, int64_t* colref) {
  const int t = threadIdx.x;
  if (t < 128) {
    const int slot = t >> 1;
    const int pos = slot < 32 ? tj * 32 + slot : ti * 32 + slot - 32;
    colref[t] = pos < P.npe ? P.gcol[2 * ((size_t)e * P.npe + pos) + (t & 1)] : 0;
  }
}

// Epilogue: LDS holds rows [48h, 48h+48) of the 96x96 G tile (row positions
// 16h..16h+15 of tile ti, all 32 column positions of tile tj) -> K blocks,
// written in place (single contribution) or staged.  Each thread takes two
// blocks per pass; the destination references of all its blocks are loaded
// before any block is formed, so their latencies overlap.
template <bool ISO>
__device__ void gram_epilogue(const GramParams& P, const double* Gt, const int64_t* colref, const double* Cs,
                              int64_t e, int ti, int tj, int h) {
  const int npe = P.npe;
  const int npass = ti == tj ? 1 : 2;
  int32_t sl[4];
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // pass 0: block (row b in ti, col a in tj), b fastest across threads;
    // pass 1: its transpose (row a in tj, col b in ti), a fastest.
    const int pass = k >> 1, idx = threadIdx.x + kGramThreads * (k & 1);
    const int hl = pass == 0 ? (idx & 15) : (idx >> 5);   // row position within the half
    const int al = pass == 0 ? (idx >> 4) : (idx & 31);   // column position in tile tj
    const int b = ti * 32 + 16 * h + hl, a = tj * 32 + al;
    ok[k] = pass < npass && b < npe && a < npe;
    sl[k] = ok[k] ? P.gslot[((size_t)e * npe + (pass == 0 ? a : b)) * npe + (pass == 0 ? b : a)] : 0;
  }
  // one block at a time (the other half's accumulators are still live)
#pragma unroll 1
  for (int k = 0; k < 4; ++k) {
    const int32_t slk = k == 0 ? sl[0] : k == 1 ? sl[1] : k == 2 ? sl[2] : sl[3];
    const bool okk = k == 0 ? ok[0] : k == 1 ? ok[1] : k == 2 ? ok[2] : ok[3];
    if (!okk) continue;
    const int pass = k >> 1, idx = threadIdx.x + kGramThreads * (k & 1);
    const int hl = pass == 0 ? (idx & 15) : (idx >> 5);
    const int al = pass == 0 ? (idx >> 4) : (idx & 31);
    // G_ba[x][y] of (row position b, column position a); pass 1 needs G_ab = G_ba^T
    double G[9];
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) {
        const double g = Gt[(3 * hl + x) * kGtStride + 3 * al + y];
        if (pass == 0) G[x * 3 + y] = g; else G[y * 3 + x] = g;
      }
    // destination of the block of (column node, row node): [c][d] at o[c * len + d]
    double* o;
    int64_t len;
    if (slk >= 0) {   // CSC panel: value (c, slot s, d) at 9 a0 + c 3R + 3 s + d
      // column reference in LDS: tile tj's 32 positions, then tile ti's
      const int cslot = pass == 0 ? al : 32 + 16 * h + hl;
      o = P.out + colref[2 * cslot] + slk;
      len = colref[2 * cslot + 1];
    } else {
      o = P.S + (size_t)(-(int64_t)slk - 1) * 9;
      len = 3;
    }
    if constexpr (ISO) {
      double K[9];
      const double tr = P.mu * (G[0] + G[4] + G[8]);
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) K[d * 3 + c] = P.lam * G[d * 3 + c] + P.mu * G[c * 3 + d];
      K[0] += tr; K[4] += tr; K[8] += tr;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) o[c * len + d] = K[d * 3 + c];
    } else {
      // K[d][c] = sum_xy C[d][x][c][y] G[x][y], one entry at a time
#pragma unroll 1
      for (int dc = 0; dc < 9; ++dc) {
        const int d = dc / 3, c = dc - 3 * d;
        double acc = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
          for (int y = 0; y < 3; ++y) acc += Cs[((d * 3 + x) * 3 + c) * 3 + y] * G[x * 3 + y];
        o[c * len + d] = acc;
      }
    }
  }
}

// Z rows of the chunk's 16 points, built in LDS.  Per point: the inverse
// Jacobian and sqrt(w |det J|) (pm); per (tile, row position, point): the
// three weighted physical gradients
//   z[al] = sw * (Mi[al] h0 + Mi[3+al] h1 + Mi[6+al] h2),  h = dN/dxi.
// FP64 layout: zs[tile][p][row] with row stride kZrs (the f64 MFMA reads a
// 16-row x 4-point fragment); BF16: hi and lo parts zh/zl[tile][row][p]
// (the bf16 MFMA reads 4 consecutive points of a row).
constexpr int kZpts = 16;     // points per chunk
constexpr int kZrs = 112;     // FP64 row stride of a point's Z column (96 rows + pad: 2 LDS cycles per fragment)

template <bool BF16>
__device__ __forceinline__ void gram_build_z(const GramParams& P, const double* pm, int ti, int tj, int c0,
                                             double* zs, short* zh, short* zl) {
  const int ntile = ti == tj ? 1 : 2;
  for (int f = threadIdx.x; f < ntile * 32 * kZpts; f += kGramThreads) {
    const int tile = f / (32 * kZpts);
    // FP64: row positions fastest across lanes (the zs column of a point is
    // contiguous); BF16: points fastest (a row's 16 points are contiguous)
    const int p = BF16 ? f % kZpts : (f / 32) % kZpts;
    const int n = BF16 ? (f / kZpts) % 32 : f % 32;
    const int pos = (tile == 0 ? ti : tj) * 32 + n;
    const double* h = BF16 ? P.dNp + ((size_t)pos * P.qp + c0 + p) * 3 : P.dNp + ((size_t)(c0 + p) * P.npos + pos) * 3;
    const double h0 = h[0], h1 = h[1], h2 = h[2];
    const double* M = pm + p * 10;
    const double sw = M[9];
#pragma unroll
    for (int al = 0; al < 3; ++al) {
      const double z = sw * (M[al] * h0 + M[3 + al] * h1 + M[6 + al] * h2);
      const int row = 3 * n + al;
      if constexpr (!BF16) {
        zs[(tile * kZpts + p) * kZrs + row] = z;
      } else {
        const short hi = bf16_rn((float)z);
        const short lo = bf16_rn((float)(z - (double)bf16_to_f(hi)));
        zh[(tile * kHoTile + row) * kZpts + p] = hi;
        zl[(tile * kHoTile + row) * kZpts + p] = lo;
      }
    }
  }
}

// One block per (element, upper tile pair); 4 waves, each a 3x3 grid of 16x16
// subtiles of the 96x96 tile (wave w: rows 48*(w/2).., cols 48*(w%2)..).
// Per chunk of 16 points: per-point geometry into LDS, the two tiles' Z rows
// into LDS, then the MFMAs of the chunk read their fragments from LDS.  The
// tile then goes through the same LDS in two 48-row halves (36 KB), so four
// blocks fit a CU.
template <bool BF16, bool ISO>
__global__ __launch_bounds__(kGramThreads) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_gram(GramParams P) {
  extern __shared__ double Gt[];   // 48 rows x kGtStride doubles; during the k loop the Z chunk
  __shared__ int64_t colref[2 * 64];
  __shared__ double Cs[ISO ? 1 : 81];
  __shared__ double pm[kZpts * 10];
  if constexpr (!ISO)
    for (int i = threadIdx.x; i < 81; i += blockDim.x) Cs[i] = P.C[i];
  int64_t e;
  int ti, tj;
  if (!gram_block(P, e, ti, tj)) return;   // block-uniform
  load_colref(P, e, ti, tj, colref);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int tb = ti == tj ? 0 : 1;        // tile of the B operand in the chunk buffers
  const int ra = (wave >> 1) * 48 + lr, rb = (wave & 1) * 48 + lr;
  double* zs = Gt;
  short* zh = reinterpret_cast<short*>(Gt);
  short* zl = zh + 2 * kHoTile * kZpts;
  using Acc = typename std::conditional<BF16, f4, d4>::type;
  Acc acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = Acc{0, 0, 0, 0};
  for (int c0 = 0; c0 < P.qp; c0 += kZpts) {
    // per-point geometry of the chunk (the previous chunk's MFMAs are done:
    // barrier at the end of the loop body)
    if (threadIdx.x < kZpts * 10) {
      const int p = threadIdx.x / 10, m = threadIdx.x % 10;
      const int pg = c0 + p;
      double v = 0.0;
      if (pg < P.q) {
        const double* g = P.geo + (P.affine ? (size_t)e : ((size_t)e * P.q + pg)) * kGeoRec;
        v = m < 9 ? g[m] : sqrt(P.affine ? g[9] * P.w[pg] : g[9]);
      }
      pm[threadIdx.x] = v;
    }
    __syncthreads();
    gram_build_z<BF16>(P, pm, ti, tj, c0, zs, zh, zl);
    __syncthreads();
    if constexpr (!BF16) {
#pragma unroll
      for (int kb = 0; kb < kZpts / 4; ++kb) {
        double a[3], b[3];
        const double* za = zs + (kb * 4 + lk) * kZrs;
        const double* zb = zs + (tb * kZpts + kb * 4 + lk) * kZrs;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          a[i] = za[ra + 16 * i];
          b[i] = zb[rb + 16 * i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    } else {
      s4 ah[3], al_[3], bh[3], bl[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        ah[i] = *reinterpret_cast<const s4*>(zh + (ra + 16 * i) * kZpts + 4 * lk);
        al_[i] = *reinterpret_cast<const s4*>(zl + (ra + 16 * i) * kZpts + 4 * lk);
        bh[i] = *reinterpret_cast<const s4*>(zh + (tb * kHoTile + rb + 16 * i) * kZpts + 4 * lk);
        bl[i] = *reinterpret_cast<const s4*>(zl + (tb * kHoTile + rb + 16 * i) * kZpts + 4 * lk);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al_[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // output layouts: f64 16x16x4: lane l, reg r -> row (l/16) + 4r, col l%16;
  // bf16 16x16x16: row 4(l/16) + r, col l%16
  const int tc0 = (wave & 1) * 48;
  for (int h = 0; h < 2; ++h) {
    if ((wave >> 1) == h) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = BF16 ? 16 * i + 4 * lk + r : 16 * i + lk + 4 * r;
            Gt[row * kGtStride + tc0 + 16 * j + lr] = (double)acc[i][j][r];
          }
    }
    __syncthreads();
    gram_epilogue<ISO>(P, Gt, colref, Cs, e, ti, tj, h);
    __syncthreads();
  }
}

// CSC blocks with more than one contribution: nine lanes per block, one per
// entry [c][d]; its staged blocks are contiguous (ascending element order)
// and summed in that order, so neighbouring lanes read neighbouring doubles.
__global__ void k_stage_gather(int64_t nmulti, const int32_t* __restrict__ mptr, const int64_t* __restrict__ mref,
                               const double* __restrict__ S, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nmulti * 9) return;
  const int64_t k = t / 9;
  const int m = (int)(t - 9 * k);
  const int64_t c0 = mptr[k], c1 = mptr[k + 1];
  const int64_t base = mref[2 * k], len = mref[2 * k + 1];
  // eight contributions per round, all loads issued before the sums (a
  // hexahedral mesh has at most eight elements at a node pair)
  double acc = 0.0;
  for (int64_t c = c0; c < c1; c += 8) {
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = c + i < c1 ? S[(size_t)(c + i) * 9 + m] : 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (c + i < c1) acc += v[i];
  }
  out[base + (m / 3) * len + m % 3] = acc;
}

// contribution keys: (block p << 32) | element block e*npe^2 + a*npe + b
// (column local node a, row local node b), generated per (incidence, b)
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

// per sorted contribution t: 1 if its CSC block has several contributions
// (slot n: 0, so the exclusive sum ends with the staged count)
__global__ void k_multi_flags(int64_t n, const uint64_t* __restrict__ keys, int32_t* __restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n) return;
  int32_t f = 0;
  if (t < n) {
    const uint64_t p = keys[t] >> 32;
    f = (t > 0 && (keys[t - 1] >> 32) == p) || (t + 1 < n && (keys[t + 1] >> 32) == p);
  }
  flag[t] = f;
}

// gslot per sorted contribution t: staged blocks get -(sid+1); a single
// contribution gets its value offset 3 s in the column panel.  Multi blocks
// record their first staged block and their value offset (at the first of
// their contributions).
__global__ void k_gslot(int64_t n, int npe, const uint64_t* __restrict__ keys, const int32_t* __restrict__ flag,
                        const int32_t* __restrict__ sid, const int32_t* __restrict__ iperm,
                        const int32_t* __restrict__ blk_node, const int64_t* __restrict__ adj_ptr,
                        int32_t* __restrict__ gslot, int32_t* __restrict__ mfirst, int64_t* __restrict__ mref_of_blk) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t p = (int64_t)(keys[t] >> 32);
  const int64_t src = (int64_t)(uint32_t)keys[t];
  const int64_t nn = (int64_t)npe * npe;
  const int64_t e = src / nn;
  const int a = (int)((src / npe) % npe), b = (int)(src % npe);
  const int64_t pos = (e * npe + iperm[a]) * npe + iperm[b];
  const int64_t j = blk_node[p];
  const int64_t a0 = adj_ptr[j];
  if (flag[t]) {
    gslot[pos] = -(sid[t] + 1);
    if (t == 0 || (keys[t - 1] >> 32) != (uint64_t)p) {
      mfirst[p] = sid[t];
      mref_of_blk[2 * p] = 9 * a0 + 3 * (p - a0);
      mref_of_blk[2 * p + 1] = 3 * (adj_ptr[j + 1] - a0);
    }
  } else {
    gslot[pos] = (int32_t)(3 * (p - a0));
  }
}

// per multi block m (ascending p): its staged range start and value reference
__global__ void k_multi_tables(int64_t nmulti, const int32_t* __restrict__ multi, const int32_t* __restrict__ mfirst,
                               const int64_t* __restrict__ mref_of_blk, int64_t nstaged, int32_t* __restrict__ mptr,
                               int64_t* __restrict__ mref) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m > nmulti) return;
  if (m == nmulti) {
    mptr[m] = (int32_t)nstaged;
    return;
  }
  const int64_t p = multi[m];
  mptr[m] = mfirst[p];
  mref[2 * m] = mref_of_blk[2 * p];
  mref[2 * m + 1] = mref_of_blk[2 * p + 1];
}

// gcol per (element, column position): the column node's panel reference
__global__ void k_gcol(int64_t E, int npe, const int32_t* __restrict__ conn, const int32_t* __restrict__ perm,
                       const int64_t* __restrict__ adj_ptr, int64_t* __restrict__ gcol) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E * npe) return;
  const int64_t e = i / npe;
  const int apos = (int)(i - e * npe);
  const int64_t j = conn[(size_t)perm[apos] * E + e];
  const int64_t a0 = adj_ptr[j];
  gcol[2 * i] = 9 * a0;
  gcol[2 * i + 1] = 3 * (adj_ptr[j + 1] - a0);
}

__global__ void k_mark_multi(int64_t nadj, const int32_t* __restrict__ mfirst, uint8_t* __restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nadj) flag[p] = mfirst[p] >= 0 ? 1 : 0;
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

}  // namespace

// Gram tables of a mesh (d_ho_perm, d_gslot, d_gcol, d_mptr, d_mref), built
// once: the element blocks' contributions sorted by (CSC block, element
// block), then split into single-contribution blocks (written in place by
// the Gram epilogue) and staged ones (contiguous per CSC block).
int build_contrib(llfem_mesh* m) {
  if (m->d_gslot) return LLFEM_OK;
  llfem_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int npe = m->npe;
  const int64_t n = m->ninc * npe;
  if (m->E * (int64_t)npe * npe >= ((int64_t)1 << 31))
    return fail(LLFEM_EUNSUP, "staged assembly: E*npe^2 >= 2^31");
  if (n != m->E * (int64_t)npe * npe) return fail(LLFEM_EINVAL, "staged assembly: incidence count mismatch");
  // row positions: element 0's nodes in ascending global id
  std::vector<int32_t> perm(npe), iperm(npe);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int32_t x, int32_t y) { return m->h_conn0[x] < m->h_conn0[y]; });
  for (int k = 0; k < npe; ++k) iperm[perm[k]] = k;
  int32_t *d_iperm = nullptr, *node = nullptr, *flag = nullptr, *sid = nullptr, *mfirst = nullptr,
          *blk_node = nullptr, *multi = nullptr;
  uint64_t *keys = nullptr, *tmp = nullptr;
  int64_t *mref_of_blk = nullptr, *d_num = nullptr;
  uint8_t* mflag = nullptr;
  void *d_sort = nullptr, *d_scan = nullptr, *d_sel = nullptr;
  size_t sort_b = 0, scan_b = 0, sel_b = 0;
  auto cleanup = [&]() {
    drelease(ctx, d_iperm, (size_t)npe);
    drelease(ctx, node, (size_t)m->ninc);
    drelease(ctx, flag, (size_t)n + 1);
    drelease(ctx, sid, (size_t)n + 1);
    drelease(ctx, mfirst, (size_t)m->nadj);
    drelease(ctx, blk_node, (size_t)m->nadj);
    drelease(ctx, multi, (size_t)m->nadj);
    drelease(ctx, keys, (size_t)n);
    drelease(ctx, tmp, (size_t)n);
    drelease(ctx, mref_of_blk, 2 * (size_t)m->nadj);
    drelease(ctx, d_num, 1);
    drelease(ctx, mflag, (size_t)m->nadj);
    if (d_sort) dfree(ctx, d_sort, sort_b);
    if (d_scan) dfree(ctx, d_scan, scan_b);
    if (d_sel) dfree(ctx, d_sel, sel_b);
  };
  auto run = [&]() -> int {
    LLFEM_TRY(dalloc(ctx, &m->d_ho_perm, (size_t)npe));
    LLFEM_TRY(dalloc(ctx, &d_iperm, (size_t)npe));
    LLFEM_HIP(hipMemcpyAsync(m->d_ho_perm, perm.data(), sizeof(int32_t) * npe, hipMemcpyHostToDevice, st));
    LLFEM_HIP(hipMemcpyAsync(d_iperm, iperm.data(), sizeof(int32_t) * npe, hipMemcpyHostToDevice, st));
    LLFEM_TRY(dalloc(ctx, &node, (size_t)m->ninc));
    LLFEM_TRY(dalloc(ctx, &keys, (size_t)n));
    LLFEM_TRY(dalloc(ctx, &tmp, (size_t)n));
    k_node_of_incidence<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, m->d_inc_ptr, node);
    k_contrib_keys<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        m->ninc, npe, node, m->d_inc_elem, m->d_inc_loc,
        m->map_bytes == 1 ? (const uint8_t*)m->d_inc_map : nullptr,
        m->map_bytes == 2 ? (const uint16_t*)m->d_inc_map : nullptr, m->d_adj_ptr, keys);
    LLFEM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sort_b, keys, tmp, (int)n, 0, 64, st));
    LLFEM_TRY(dmalloc(ctx, &d_sort, sort_b));
    LLFEM_HIP(hipcub::DeviceRadixSort::SortKeys(d_sort, sort_b, keys, tmp, (int)n, 0, 64, st));
    // staged ids: exclusive sum of the multi flags over the sorted contributions
    LLFEM_TRY(dalloc(ctx, &flag, (size_t)n + 1));
    LLFEM_TRY(dalloc(ctx, &sid, (size_t)n + 1));
    k_multi_flags<<<(unsigned)((n + 256) / 256), 256, 0, st>>>(n, tmp, flag);
    LLFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, flag, sid, (int)(n + 1), st));
    LLFEM_TRY(dmalloc(ctx, &d_scan, scan_b));
    LLFEM_HIP(hipcub::DeviceScan::ExclusiveSum(d_scan, scan_b, flag, sid, (int)(n + 1), st));
    LLFEM_TRY(dalloc(ctx, &blk_node, (size_t)m->nadj));
    LLFEM_TRY(dalloc(ctx, &mfirst, (size_t)m->nadj));
    LLFEM_TRY(dalloc(ctx, &mref_of_blk, 2 * (size_t)m->nadj));
    LLFEM_TRY(dalloc(ctx, &m->d_gslot, (size_t)n));
    k_blk_node<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, m->d_adj_ptr, blk_node);
    k_fill_i32<<<(unsigned)((m->nadj + 255) / 256), 256, 0, st>>>(m->nadj, -1, mfirst);
    k_gslot<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, npe, tmp, flag, sid, d_iperm, blk_node, m->d_adj_ptr,
                                                        m->d_gslot, mfirst, mref_of_blk);
    int32_t ns = 0;
    LLFEM_HIP(hipMemcpyAsync(&ns, sid + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    // multi blocks in ascending p
    LLFEM_TRY(dalloc(ctx, &mflag, (size_t)m->nadj));
    LLFEM_TRY(dalloc(ctx, &multi, (size_t)m->nadj));
    LLFEM_TRY(dalloc(ctx, &d_num, 1));
    k_mark_multi<<<(unsigned)((m->nadj + 255) / 256), 256, 0, st>>>(m->nadj, mfirst, mflag);
    hipcub::CountingInputIterator<int32_t> ids(0);
    LLFEM_HIP(hipcub::DeviceSelect::Flagged(nullptr, sel_b, ids, mflag, multi, d_num, (int)m->nadj, st));
    LLFEM_TRY(dmalloc(ctx, &d_sel, sel_b));
    LLFEM_HIP(hipcub::DeviceSelect::Flagged(d_sel, sel_b, ids, mflag, multi, d_num, (int)m->nadj, st));
    int64_t nm = 0;
    LLFEM_HIP(hipMemcpyAsync(&nm, d_num, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipStreamSynchronize(st));
    m->nstaged = ns;
    m->nmulti = nm;
    LLFEM_TRY(dalloc(ctx, &m->d_mptr, (size_t)nm + 1));
    LLFEM_TRY(dalloc(ctx, &m->d_mref, 2 * (size_t)nm));
    k_multi_tables<<<(unsigned)((nm + 256) / 256), 256, 0, st>>>(nm, multi, mfirst, mref_of_blk, ns, m->d_mptr,
                                                                 m->d_mref);
    LLFEM_TRY(dalloc(ctx, &m->d_gcol, 2 * (size_t)m->E * npe));
    k_gcol<<<(unsigned)((m->E * npe + 255) / 256), 256, 0, st>>>(m->E, npe, m->d_conn, m->d_ho_perm, m->d_adj_ptr,
                                                                 m->d_gcol);
    LLFEM_HIP(hipGetLastError());
    LLFEM_HIP(hipStreamSynchronize(st));
    return LLFEM_OK;
  };
  const int rc = run();
  cleanup();
  if (rc != LLFEM_OK) {
    drelease(ctx, m->d_ho_perm, (size_t)npe);
    drelease(ctx, m->d_gslot, (size_t)n);
    drelease(ctx, m->d_gcol, 2 * (size_t)m->E * npe);
    drelease(ctx, m->d_mptr, (size_t)m->nmulti + 1);
    drelease(ctx, m->d_mref, 2 * (size_t)m->nmulti);
    m->nmulti = m->nstaged = 0;
  }
  return rc;
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
  const int npos = 32 * ntiles;
  const int qp = (q + kZpts - 1) / kZpts * kZpts;
  auto al256 = [](size_t b) { return (b + 255) / 256 * 256; };
  // d_ho: staged blocks | general material | reference gradients by row position
  const size_t sbytes = al256((size_t)m->nstaged * 9 * sizeof(double));
  const size_t cbytes = al256(81 * sizeof(double));
  const size_t nbytes = al256((size_t)qp * npos * 3 * sizeof(double));
  const size_t need = sbytes + cbytes + nbytes;
  if (m->ho_bytes < need) {
    if (m->d_ho) dfree(ctx, m->d_ho, m->ho_bytes);
    m->d_ho = nullptr;
    m->ho_bytes = 0;
    LLFEM_TRY(dmalloc(ctx, &m->d_ho, need));
    m->ho_bytes = need;
  }
  char* base = static_cast<char*>(m->d_ho);
  double* S = reinterpret_cast<double*>(base);
  double* dC = reinterpret_cast<double*>(base + sbytes);
  double* dNp = reinterpret_cast<double*>(base + sbytes + cbytes);
  if (C81) LLFEM_HIP(hipMemcpyAsync(dC, C81, 81 * sizeof(double), hipMemcpyHostToDevice, st));
  const int64_t ndn = (int64_t)qp * npos * 3;
  k_dn_positions<<<(unsigned)((ndn + 255) / 256), 256, 0, st>>>(q, qp, npe, npos, bf ? 0 : 1, m->d_ho_perm, d_dN,
                                                                 dNp);
  GramParams P{};
  P.E = m->E;
  P.npe = npe;
  P.ntiles = ntiles;
  P.npairs = ntiles * (ntiles + 1) / 2;
  P.q = q;
  P.qp = qp;
  P.npos = npos;
  P.affine = affine ? 1 : 0;
  P.geo = m->d_geo;
  P.w = d_w;
  P.dNp = dNp;
  P.lam = lam;
  P.mu = mu;
  P.C = C81 ? dC : nullptr;
  P.S = S;
  P.gslot = m->d_gslot;
  P.gcol = m->d_gcol;
  P.out = out;
  const size_t lds = sizeof(double) * 48 * kGtStride;
  auto fn = bf ? (C81 ? k_gram<true, false> : k_gram<true, true>) : (C81 ? k_gram<false, false> : k_gram<false, true>);
  LLFEM_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int64_t nblocks = (m->E * P.npairs + kHoXcds - 1) / kHoXcds * kHoXcds;
  if (nblocks >= ((int64_t)1 << 31)) return fail(LLFEM_EUNSUP, "high-order path: too many elements");
  const dim3 grid((unsigned)nblocks);
  LLFEM_HIP(hipEventRecord(ctx->ev[1], st));
  hipLaunchKernelGGL(fn, grid, dim3(kGramThreads), lds, st, P);
  LLFEM_HIP(hipEventRecord(ctx->ev[2], st));
  if (m->nmulti > 0)
    k_stage_gather<<<(unsigned)((9 * m->nmulti + 255) / 256), 256, 0, st>>>(m->nmulti, m->d_mptr, m->d_mref, S,
                                                                             out);
  LLFEM_HIP(hipGetLastError());
  return LLFEM_OK;
}

}  // namespace llfem
