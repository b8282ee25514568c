// Operator assembly: element geometry + node-panel accumulation.
//
// Reference algorithm: K = sum_e P_e^T Ke P_e with
//   Ke = sum_q B_q^T D B_q |det J_q| w_q        (src/multifield.jl:1857-2052,
//                                                 src/linear.jl:272-307)
// scattered into the CSC pattern block by block (src/multifield.jl:2465-2493).
//
// MI355X design (owner computes, no atomics on HBM, no staging of Ke):
//   * In node-major CSC the three columns of node j are one contiguous run of
//     9*R_j doubles ("panel": [column c][row slot s][row comp d]) where R_j is
//     the number of adjacent nodes.  Each panel is owned by one wavefront.
//   * The wavefront walks j's incidences (element e, local index a) in
//     ascending element order; lane b computes the 3x3 block Ke[(b,:),(a,:)]
//     and adds it to LDS at the slot given by the element->nnz map.  Within
//     one step every lane of a node hits a different slot and LDS executes a
//     wave's instructions in order, so the sums are deterministic and
//     conflict-free (ds_add_f64 is used as a plain add).
//   * The finished panel is written once, coalesced.  HBM traffic is the
//     nzval write plus index/geometry reads (no read-modify-write of nzval).
//   * Nodes are processed in classes of equal incidence count so the
//     incidence loop never diverges inside a wave; 64/npe nodes share a wave.
//   * Affine elements (constant Jacobian: straight-sided simplices,
//     parallelepipeds) use precomputed reference integrals
//     Ghat_ba = sum_q w_q dNhat_b dNhat_a^T, so a block costs two 3x3x3
//     products: G = Minv^T Ghat Minv |det J|.  Curved elements integrate per
//     Gauss point from per-GP inverse Jacobians.
//   * K_ba[d][c] = lam G[d][c] + mu G[c][d] + mu tr(G) delta_dc for Hooke's
//     law, or sum C[d][al][c][be] G[al][be] for a general 6x6 D.
#include <algorithm>
#include <numeric>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "common.h"

#include "panel.h"

namespace llfem {

namespace {

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
using PanelFn = void (*)(PanelParams);

template <int NPE, typename MapT>
PanelFn pick(int op, bool affine, bool iso, bool fp32) {
  if (op == LLFEM_OP_MASS)
    return affine ? k_panel<NPE, MapT, LLFEM_OP_MASS, true, true, double>
                  : k_panel<NPE, MapT, LLFEM_OP_MASS, false, true, double>;
  if (affine) {
    if (iso) return fp32 ? k_panel<NPE, MapT, 0, true, true, float> : k_panel<NPE, MapT, 0, true, true, double>;
    return fp32 ? k_panel<NPE, MapT, 0, true, false, float> : k_panel<NPE, MapT, 0, true, false, double>;
  }
  if (iso) return fp32 ? k_panel<NPE, MapT, 0, false, true, float> : k_panel<NPE, MapT, 0, false, true, double>;
  return fp32 ? k_panel<NPE, MapT, 0, false, false, float> : k_panel<NPE, MapT, 0, false, false, double>;
}

// Stiffness, affine, Hooke's law, one-byte maps (tet4, hex8, tet10): the
// variant whose panels leave in a compile-time number of stores per round.
template <int NPE, typename Real>
PanelFn pick_fixed_stores(int ks) {
  switch (ks) {
    case 2: return k_panel<NPE, uint8_t, 0, true, true, Real, 2>;
    case 3: return k_panel<NPE, uint8_t, 0, true, true, Real, 3>;
    case 4: return k_panel<NPE, uint8_t, 0, true, true, Real, 4>;
    case 5: return k_panel<NPE, uint8_t, 0, true, true, Real, 5>;
    case 6: return k_panel<NPE, uint8_t, 0, true, true, Real, 6>;
    case 8: return k_panel<NPE, uint8_t, 0, true, true, Real, 8>;
    case 10: return k_panel<NPE, uint8_t, 0, true, true, Real, 10>;
    case 12: return k_panel<NPE, uint8_t, 0, true, true, Real, 12>;
    case 16: return k_panel<NPE, uint8_t, 0, true, true, Real, 16>;
    default: return nullptr;
  }
}

// Smallest instantiated store count covering a panel of 9 * max_r doubles
// (0: none, use the run-time store loop).
int fixed_store_count(int max_r) {
  const int need = (9 * max_r + 63) / 64;
  for (int ks : {2, 3, 4, 5, 6, 8, 10, 12, 16})
    if (ks >= need) return ks;
  return 0;
}

PanelFn select_fixed_stores(int npe, int map_bytes, int op, bool affine, bool iso, bool fp32, int max_r) {
  if (map_bytes != 1 || op != LLFEM_OP_STIFFNESS || !affine || !iso) return nullptr;
  const int ks = fixed_store_count(max_r);
  if (ks == 0) return nullptr;
  switch (npe) {
    case 4: return fp32 ? pick_fixed_stores<4, float>(ks) : pick_fixed_stores<4, double>(ks);
    case 8: return fp32 ? pick_fixed_stores<8, float>(ks) : pick_fixed_stores<8, double>(ks);
    case 10: return fp32 ? pick_fixed_stores<10, float>(ks) : pick_fixed_stores<10, double>(ks);
    default: return nullptr;
  }
}

// Round-per-iteration variants (k_panel_rounds) for the node classes of
// structured tet10 and hex8 meshes: (npe, steps per node, ring, stores per
// panel).  A launch whose (npe, cnt) is listed and whose panels fit the
// store count runs it; every other launch keeps k_panel.
template <typename Real, bool TL>
PanelFn select_rounds_real(int npe, int cnt, int ks_needed) {
  struct V {
    int npe, cnt, ks;
    PanelFn fn;
  };
  static const V table[] = {
      {10, 4, 4, k_panel_rounds<10, Real, 4, 2, 4, TL>},   {10, 4, 6, k_panel_rounds<10, Real, 4, 2, 6, TL>},
      {10, 6, 4, k_panel_rounds<10, Real, 6, 3, 4, TL>},   {10, 6, 6, k_panel_rounds<10, Real, 6, 3, 6, TL>},
      {10, 24, 10, k_panel_rounds<10, Real, 24, 3, 10, TL>}, {10, 24, 12, k_panel_rounds<10, Real, 24, 3, 12, TL>},
      {8, 8, 4, k_panel_rounds<8, Real, 8, 2, 4, TL>},     {8, 8, 6, k_panel_rounds<8, Real, 8, 2, 6, TL>},
  };
  for (const V& v : table)
    if (v.npe == npe && v.cnt == cnt && v.ks >= ks_needed) return v.fn;
  return nullptr;
}

// LLFEM_PANEL_OCC=0 sizes panel blocks by LDS alone (A/B).
bool occupancy_sizing() {
  const char* env = std::getenv("LLFEM_PANEL_OCC");
  return !env || std::atoi(env) != 0;
}

// Waves of `fn` resident on a CU with c-wave blocks (LDS, and registers
// through the occupancy calculator unless LLFEM_PANEL_OCC=0).  Cached.
size_t resident_waves(PanelFn fn, size_t wave_bytes, size_t block_extra, int c) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, size_t, size_t, int>, size_t> cache;
  const bool occ = occupancy_sizing();
  const auto key = std::make_tuple((const void*)fn, wave_bytes, block_extra, c * (occ ? 1 : -1));
  {
    std::lock_guard<std::mutex> lk(mu);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const size_t lds = c * wave_bytes + block_extra;
  size_t blocks = lds <= kLdsPerCU ? kLdsPerCU / lds : 0;
  if (occ && blocks > 0) {
    (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsPerCU);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)fn, 64 * c, lds) == hipSuccess && nb > 0)
      blocks = std::min(blocks, (size_t)nb);
  }
  std::lock_guard<std::mutex> lk(mu);
  return cache[key] = (size_t)c * blocks;
}

// Waves per block (1..max_wpb) that keep the most waves of `fn` resident on
// a CU: by LDS (c waves take c * wave_bytes + block_extra bytes) and by what
// the runtime's occupancy calculator allows for the kernel's registers.  The
// round kernels use up to ~240 VGPRs (two waves per SIMD): sized by LDS alone,
// a 6-wave block left a CU with one block, 6 of its 8 wave slots.
int resident_wpb(PanelFn fn, size_t wave_bytes, size_t block_extra, int max_wpb) {
  const bool occ = occupancy_sizing();
  int wpb = 0;
  size_t best = 0;
  for (int c = 1; c <= max_wpb && c * wave_bytes + block_extra <= kLdsPerCU; ++c) {
    const size_t resident = resident_waves(fn, wave_bytes, block_extra, c);
    // ties: with a per-block table in LDS, the block nearest 4 waves (1-wave
    // blocks repeat the table fill per wave, 8-wave blocks leave a CU idle
    // while their last wave finishes); without one, the smallest block
    // (measured on the tet10 beam's FP64 and FP32 round kernels)
    const bool table = block_extra > 64;
    if (resident > best || (occ && table && resident == best && std::abs(c - 4) < std::abs(wpb - 4))) {
      best = resident;
      wpb = c;
    }
  }
  return wpb;
}

// Reference table in LDS for the round kernels: each lane reads its 9 table
// values from LDS instead of through the texture-address unit, the busiest
// unit of these kernels (FP64 assembly of the tet10 beam 2.09 -> 2.04 ms; the
// FP32 kernels measured 2 % slower with it until the launches were sized by
// registers too, 3.5 % faster since).  LLFEM_PANEL_TABLE_LDS=0 and
// LLFEM_PANEL_TABLE_LDS_F32=0 switch it off for A/B runs.
bool rounds_table_lds_f32() {
  const char* e = std::getenv("LLFEM_PANEL_TABLE_LDS_F32");
  return !e || std::atoi(e) != 0;
}

bool rounds_table_lds() {
  const char* e = std::getenv("LLFEM_PANEL_TABLE_LDS");
  return !e || std::atoi(e) != 0;
}

PanelFn select_rounds(int npe, int map_bytes, int op, bool affine, bool iso, bool fp32, int cnt, int max_r) {
  if (map_bytes != 1 || op != LLFEM_OP_STIFFNESS || !affine || !iso) return nullptr;
  const int need = (9 * max_r + 7 + 63) / 64;   // up to 7 doubles of window alignment
  if (fp32)
    return rounds_table_lds_f32() ? select_rounds_real<float, true>(npe, cnt, need)
                                  : select_rounds_real<float, false>(npe, cnt, need);
  return rounds_table_lds() ? select_rounds_real<double, true>(npe, cnt, need)
                            : select_rounds_real<double, false>(npe, cnt, need);
}

PanelFn select_panel(int npe, int map_bytes, int op, bool affine, bool iso, bool fp32) {
  if (map_bytes == 1) {
    switch (npe) {
      case 4: return pick<4, uint8_t>(op, affine, iso, fp32);
      case 6: return pick<6, uint8_t>(op, affine, iso, fp32);
      case 8: return pick<8, uint8_t>(op, affine, iso, fp32);
      case 10: return pick<10, uint8_t>(op, affine, iso, fp32);
      case 27: return pick<27, uint8_t>(op, affine, iso, fp32);
      default: return nullptr;
    }
  }
  switch (npe) {
    case 8: return pick<8, uint16_t>(op, affine, iso, fp32);
    case 10: return pick<10, uint16_t>(op, affine, iso, fp32);
    case 27: return pick<27, uint16_t>(op, affine, iso, fp32);
    case 64: return pick<64, uint16_t>(op, affine, iso, fp32);
    default: return nullptr;
  }
}

__global__ void k_add_into(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

int ensure_buf(llfem_ctx* ctx, double** p, size_t* cur, size_t bytes) {
  if (*p && *cur >= bytes) return LLFEM_OK;
  if (*p) dfree(ctx, *p, *cur);
  *p = nullptr;
  *cur = 0;
  LLFEM_TRY(dmalloc(ctx, (void**)p, bytes));
  *cur = bytes;
  return LLFEM_OK;
}

}  // namespace

// Voigt index of the symmetric pair {i,j} in SymGrad row order
// [xx,yy,zz,xy,xz,yz] (src/multifield.jl:635-657).
int voigt(int i, int j) {
  if (i == j) return i;
  int a = i < j ? i : j, b = i < j ? j : i;
  if (a == 0 && b == 1) return 3;
  if (a == 0 && b == 2) return 4;
  return 5;
}

// Reference operator table of a panel launch.  Stiffness: affine -> Ghat
// [a][b][al][be] = sum_q w_q dN_b[al] dN_a[be]; curved -> dN itself.  Mass:
// affine -> Mhat[b][a] = sum_q w_q N_b N_a; curved -> N itself.
std::vector<double> operator_table(const llfem_op_desc* d, int npe, bool affine) {
  const int q = d->q;
  std::vector<double> tab;
  if (d->op == LLFEM_OP_STIFFNESS) {
    if (affine) {
      tab.assign((size_t)npe * npe * 9, 0.0);
      for (int b = 0; b < npe; ++b)
        for (int a = 0; a < npe; ++a)
          for (int al = 0; al < 3; ++al)
            for (int be = 0; be < 3; ++be) {
              double s = 0.0;
              for (int p = 0; p < q; ++p)
                s += d->weights[p] * d->dN_ref[((size_t)p * npe + b) * 3 + al] *
                     d->dN_ref[((size_t)p * npe + a) * 3 + be];
              tab[((size_t)a * npe + b) * 9 + al * 3 + be] = s;   // [a][b] order
            }
    } else {
      tab.assign(d->dN_ref, d->dN_ref + (size_t)q * npe * 3);
    }
  } else {
    if (affine) {
      tab.assign((size_t)npe * npe, 0.0);
      for (int b = 0; b < npe; ++b)
        for (int a = 0; a < npe; ++a) {
          double s = 0.0;
          for (int p = 0; p < q; ++p)
            s += d->weights[p] * d->N_ref[(size_t)p * npe + b] * d->N_ref[(size_t)p * npe + a];
          tab[(size_t)b * npe + a] = s;
        }
    } else {
      tab.assign(d->N_ref, d->N_ref + (size_t)q * npe);
    }
  }
  return tab;
}

// Offset (doubles) in m->d_tab of the point-fastest gradient table: after
// [dN | w | op table (double) | op table (float)].
size_t dnt_offset(int npe, int q) {
  const size_t geo_tab = (size_t)q * npe * 3 + q;
  const size_t op_max = std::max<size_t>((size_t)npe * npe * 9, (size_t)q * npe * 3);
  return geo_tab + op_max + op_max / 2 + 2;
}

// Upload [dN | w | op-table space] and run the geometry check once per rule.
int geometry_state(llfem_mesh* m, const llfem_op_desc* d, bool* affine_out) {
  llfem_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int npe = m->npe, q = d->q;
  uint64_t key = hash_bytes(d->dN_ref, sizeof(double) * (size_t)q * npe * 3);
  key = hash_bytes(d->weights, sizeof(double) * (size_t)q, key);
  RuleState& rs = m->rules[key];
  // [dN | w | op table (double) | op table (float) | dN points fastest]
  LLFEM_TRY(ensure_buf(ctx, &m->d_tab, &m->tab_bytes, sizeof(double) * (dnt_offset(npe, q) + (size_t)q * npe * 3)));
  LLFEM_HIP(hipMemcpyAsync(m->d_tab, d->dN_ref, sizeof(double) * (size_t)q * npe * 3,
                           hipMemcpyHostToDevice, st));
  LLFEM_HIP(hipMemcpyAsync(m->d_tab + (size_t)q * npe * 3, d->weights, sizeof(double) * q,
                           hipMemcpyHostToDevice, st));
  if (!rs.checked) {
    int* d_flags = nullptr;
    LLFEM_TRY(dalloc(ctx, &d_flags, 1));
    LLFEM_HIP(hipMemsetAsync(d_flags, 0, sizeof(int), st));
    k_geom_check<<<(unsigned)((m->E + 255) / 256), 256, 0, st>>>(m->E, m->N, npe, q, m->d_xyz,
                                                                 m->d_conn, m->d_tab, d_flags);
    int hf = 0;
    LLFEM_HIP(hipMemcpyAsync(&hf, d_flags, sizeof(int), hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipStreamSynchronize(st));
    drelease(ctx, d_flags, 1);
    rs.checked = true;
    rs.affine = !(hf & 1);
    rs.singular = (hf & 2) != 0;
  }
  if (rs.singular)
    return fail(LLFEM_EGEOM, "compute_element_geometry!: singular geometry at integration point.");
  *affine_out = rs.affine;
  return LLFEM_OK;
}

int run_geometry(llfem_mesh* m, const llfem_op_desc* d, bool affine, const double* d_scale, bool f32) {
  llfem_ctx* ctx = m->ctx;
  const int npe = m->npe, q = d->q;
  const size_t geo_doubles = (size_t)kGeoStride * m->E * (affine ? 1 : q);
  LLFEM_TRY(ensure_buf(ctx, &m->d_geo, &m->geo_bytes, sizeof(double) * geo_doubles));
  const int64_t threads = affine ? m->E : m->E * q;
  if (affine) {
    const unsigned g = (unsigned)((m->E + 255) / 256);
    const double* dN0 = m->d_tab;   // gradients at the first point
#define LLFEM_GEO_AFFINE(n)                                                                        \
  case n:                                                                                          \
    if (f32)                                                                                       \
      k_geometry_affine<n, true><<<g, 256, 0, ctx->stream>>>(m->E, m->d_xyz4, m->d_conn, dN0, m->d_geo, d_scale); \
    else                                                                                           \
      k_geometry_affine<n, false><<<g, 256, 0, ctx->stream>>>(m->E, m->d_xyz4, m->d_conn, dN0, m->d_geo, d_scale); \
    return LLFEM_OK;
    switch (npe) {
      LLFEM_GEO_AFFINE(4)
      LLFEM_GEO_AFFINE(6)
      LLFEM_GEO_AFFINE(8)
      LLFEM_GEO_AFFINE(10)
      LLFEM_GEO_AFFINE(20)
      LLFEM_GEO_AFFINE(27)
      default: break;
    }
#undef LLFEM_GEO_AFFINE
  }
  if (!affine) {
    double* dNT = m->d_tab + dnt_offset(npe, q);
    const int64_t nt = (int64_t)q * npe * 3;
    k_dn_transpose<<<(unsigned)((nt + 255) / 256), 256, 0, ctx->stream>>>(npe, q, m->d_tab, dNT);
    if (npe <= kCurvedMaxNpe) {
      k_geometry_curved<<<(unsigned)m->E, 128, 0, ctx->stream>>>(m->E, m->N, npe, q, m->d_xyz, m->d_conn, dNT,
                                                                 m->d_tab + (size_t)q * npe * 3, m->d_geo, d_scale);
      return LLFEM_OK;
    }
  }
  k_geometry<<<(unsigned)((threads + 255) / 256), 256, 0, ctx->stream>>>(
      m->E, m->N, npe, q, affine ? 1 : 0, m->d_xyz, m->d_conn, m->d_tab,
      m->d_tab + (size_t)q * npe * 3, m->d_geo, d_scale);
  return LLFEM_OK;
}

}  // namespace llfem

using namespace llfem;

namespace llfem {

// llfem_assemble with an optional device target: dev_out != NULL receives the
// values (pdim*nadj*blk doubles) instead of the mesh's own buffer, and the
// mesh's assembled-operator state is left unchanged (used by the server,
// whose concurrent requests on one mesh have their own materials).
int assemble_impl(llfem_mesh* m, const llfem_op_desc* d, double* nzval_out, llfem_stats* stats,
                  double* dev_out) {
  auto t0 = std::chrono::steady_clock::now();
  if (!m || !d) return fail(LLFEM_EINVAL, "llfem_assemble: NULL argument");
  if (d->op != LLFEM_OP_STIFFNESS && d->op != LLFEM_OP_MASS)
    return fail(LLFEM_EINVAL, "llfem_assemble: unknown operator");
  if (d->q <= 0 || d->q > kMaxQ || !d->weights || !d->dN_ref)
    return fail(LLFEM_EINVAL, "llfem_assemble: bad quadrature tables");
  if (d->op == LLFEM_OP_MASS && !d->N_ref)
    return fail(LLFEM_EINVAL, "llfem_assemble: mass needs N_ref");
  if (d->precision != LLFEM_FP64 && d->precision != LLFEM_FP32_INT_FP64_SUM &&
      d->precision != LLFEM_BF16X3)
    return fail(LLFEM_EINVAL, "llfem_assemble: unknown precision");
  // order >= 3 hexahedra (hex64, hex125): Gram formulation on the matrix cores.
  // hex64 FP64 stays on the panel kernel unless LLFEM_HIGHORDER=1.
  const bool ho_type = (m->etype == 92 || m->etype == 93) && d->op == LLFEM_OP_STIFFNESS;
  bool highorder = ho_type && (m->npe > 64 || d->precision == LLFEM_BF16X3);
  if (ho_type)
    if (const char* env = std::getenv("LLFEM_HIGHORDER")) highorder = highorder || std::atoi(env) != 0;
  if (d->precision == LLFEM_BF16X3 && !highorder)
    return fail(LLFEM_EUNSUP, "llfem_assemble: BF16x3 is for order >= 3 hexahedra (stiffness)");
  const bool acc = d->accumulate != 0;
  if (acc) {
    if (dev_out) return fail(LLFEM_EINVAL, "llfem_assemble: accumulate is for the mesh's resident values");
    if (d->op == LLFEM_OP_MASS && d->lumped)
      return fail(LLFEM_EINVAL, "llfem_assemble: accumulate the consistent mass, then lump it");
    const bool have = d->op == LLFEM_OP_STIFFNESS ? (m->have_K || m->partial[0]) : (m->have_M || m->partial[1]);
    if (!have) return fail(LLFEM_EINVAL, "llfem_assemble: accumulate needs a resident operator to add to");
  }
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int npe = m->npe, q = d->q, op = d->op;
  const bool iso = (d->D == nullptr);
  const bool fp32 = (d->precision == LLFEM_FP32_INT_FP64_SUM) && op == LLFEM_OP_STIFFNESS;

  LLFEM_HIP(hipEventRecord(ctx->ev[0], st));
  bool affine = false;
  LLFEM_TRY(geometry_state(m, d, &affine));

  PanelFn fn = highorder ? nullptr : select_panel(npe, m->map_bytes, op, affine, iso, fp32);
  if (!fn && !highorder)
    return fail(LLFEM_EUNSUP, "llfem_assemble: no panel kernel for npe=" + std::to_string(npe));

  // per-operator reference tables (host, FP64); the Gram kernels read dN and w directly
  std::vector<double> tab;
  if (!highorder) tab = operator_table(d, npe, affine);
  const size_t off = (size_t)q * npe * 3 + q;
  std::vector<float> tabf(tab.begin(), tab.end());
  float* d_tabf = reinterpret_cast<float*>(m->d_tab + off + tab.size());
  LLFEM_HIP(hipMemcpyAsync(m->d_tab + off, tab.data(), sizeof(double) * tab.size(),
                           hipMemcpyHostToDevice, st));
  LLFEM_HIP(hipMemcpyAsync(d_tabf, tabf.data(), sizeof(float) * tabf.size(), hipMemcpyHostToDevice, st));
  // per-element integrand factors (material groups): uploaded for this call
  double* d_scale = nullptr;
  if (d->element_scale) {
    for (int64_t e = 0; e < m->E; ++e)
      if (!(d->element_scale[e] >= 0.0) || !std::isfinite(d->element_scale[e]))
        return fail(LLFEM_EINVAL, "llfem_assemble: element_scale must be finite and >= 0");
    LLFEM_TRY(dalloc(ctx, &d_scale, (size_t)m->E));
    LLFEM_HIP(hipMemcpyAsync(d_scale, d->element_scale, sizeof(double) * m->E, hipMemcpyHostToDevice, st));
  }
  const int grc = run_geometry(m, d, affine, d_scale, fp32);
  if (d_scale) {
    LLFEM_HIP(hipStreamSynchronize(st));   // the geometry kernel has read it
    drelease(ctx, d_scale, (size_t)m->E);
  }
  LLFEM_TRY(grc);

  double** outp = (op == LLFEM_OP_STIFFNESS) ? &m->d_valK : &m->d_valM;
  const int blk = (op == LLFEM_OP_STIFFNESS) ? 9 : 3;
  const int64_t nnz = m->nadj * blk;
  if (dev_out && op == LLFEM_OP_MASS && d->lumped)
    return fail(LLFEM_EINVAL, "assemble: lumped mass has no device target");
  if (!dev_out && !*outp) LLFEM_TRY(dalloc(ctx, outp, (size_t)nnz));
  double* target = dev_out ? dev_out : *outp;
  // accumulate: this call's values go to a scratch buffer and are added to
  // the resident ones after the panel launches
  double* acc_tmp = nullptr;
  if (acc) {
    LLFEM_TRY(dalloc(ctx, &acc_tmp, (size_t)nnz));
    target = acc_tmp;
  }

  PanelParams P{};
  P.rec_bytes = m->rec_bytes;
  P.geo = m->d_geo;
  P.E = m->E;
  P.q = q;
  P.tab = m->d_tab + off;
  P.tabf = d_tabf;
  P.lam = 0; P.mu = 0;
  if (op == LLFEM_OP_STIFFNESS && iso) {
    // elastic_constants (src/general.jl:171-175)
    P.mu = d->E / (2.0 * (1.0 + d->nu));
    P.lam = 2.0 * P.mu * d->nu / (1.0 - 2.0 * d->nu);
  }
  if (op == LLFEM_OP_STIFFNESS && !iso)
    for (int dd = 0; dd < 3; ++dd)
      for (int al = 0; al < 3; ++al)
        for (int c = 0; c < 3; ++c)
          for (int be = 0; be < 3; ++be)
            P.C[((dd * 3 + al) * 3 + c) * 3 + be] = d->D[voigt(dd, al) * 6 + voigt(c, be)];
  P.rho = d->rho;
  P.out = target;
  if (!ctx->d_sink) LLFEM_TRY(dalloc(ctx, &ctx->d_sink, 64));
  P.sink = ctx->d_sink;
  const int npw = 64 / npe;
  const int blkl = (op == LLFEM_OP_STIFFNESS) ? 9 : 1;
  if (highorder)
    LLFEM_TRY(assemble_highorder(m, d, affine, m->d_tab, m->d_tab + (size_t)q * npe * 3, P.lam, P.mu,
                                 iso ? nullptr : P.C, target));
  else
    LLFEM_HIP(hipEventRecord(ctx->ev[1], st));
  // Node classes write disjoint panels, so their launches may overlap on two
  // streams (a tuned count, or LLFEM_PANEL_STREAMS=n (1..3), sets it), which
  // fills the tail of one class's launch with the others.
  static const std::vector<PanelLaunch> kNoLaunches;
  const std::vector<PanelLaunch>& plan = highorder ? kNoLaunches : m->launches;
  const auto tuned = ctx->tune.find(m->etype);
  int nstreams = 2;
  if (tuned != ctx->tune.end() && tuned->second.streams > 0) nstreams = std::min(3, tuned->second.streams);
  if (const char* env = std::getenv("LLFEM_PANEL_STREAMS")) nstreams = std::max(1, std::min(3, std::atoi(env)));
  const bool rounds_on =
      !std::getenv("LLFEM_PANEL_NO_ROUNDS") && !(tuned != ctx->tune.end() && tuned->second.rounds == 0);
  nstreams = std::min<int>(nstreams, (int)plan.size());
  // per launch: kernel, waves per block, LDS, grid, and resident waves per CU
  struct Cfg {
    PanelFn fn;
    int wpb;
    size_t lds_bytes;
    unsigned blocks;
    double cost;   // steps / resident waves per CU: the launch's time alone, in relative units
  };
  std::vector<Cfg> cfg(plan.size());
  int lerr = LLFEM_OK;
  for (size_t li = 0; li < plan.size() && lerr == LLFEM_OK; ++li) {
    const PanelLaunch& L = plan[li];
    const size_t wave_bytes = sizeof(double) * (size_t)(npw * L.max_r * blkl + 9);   // + dummy block
    size_t max_wpb = (iso || op == LLFEM_OP_MASS) ? 8 : 4;   // matches k_panel launch bounds
    if (const char* env = std::getenv("LLFEM_PANEL_MAX_WAVES"))     // tuning knob
      max_wpb = std::max<size_t>(1, std::min<size_t>(max_wpb, (size_t)std::atoi(env)));
    // kernel of the launch; a round kernel may keep the reference table in LDS
    // after the block's panels (block_extra bytes, 16-byte aligned start)
    PanelFn lfn = fn;
    size_t block_extra = 0;
    if (!std::getenv("LLFEM_PANEL_RUNTIME_STORES")) {   // A/B knob: the run-time store loop
      if (PanelFn f = select_fixed_stores(npe, m->map_bytes, op, affine, iso, fp32, L.max_r)) lfn = f;
      if (rounds_on)                                    // A/B knob / tuned: the step-unrolled kernel
        if (PanelFn f = select_rounds(npe, m->map_bytes, op, affine, iso, fp32, L.cnt, L.max_r)) {
          lfn = f;
          if (!fp32 && rounds_table_lds()) block_extra = 8 + 10 * sizeof(double) * npe * npe;
          if (fp32 && rounds_table_lds_f32()) block_extra = 8 + 12 * sizeof(float) * npe * npe;
        }
    }
    // waves per block: the size that keeps the most waves resident on a CU
    // (several small blocks beat one large block with LDS left over)
    int wpb = resident_wpb(lfn, wave_bytes, block_extra, (int)max_wpb);
    if (wpb < 1) {
      lerr = fail(LLFEM_EUNSUP, "llfem_assemble: node panel exceeds LDS");
      break;
    }
    int want_wpb = (tuned != ctx->tune.end()) ? tuned->second.wpb : 0;
    if (const char* env = std::getenv("LLFEM_PANEL_WPB")) want_wpb = std::atoi(env);   // A/B knob
    if (want_wpb > 0 && (size_t)want_wpb <= max_wpb && (size_t)want_wpb * wave_bytes + block_extra <= kLdsPerCU)
      wpb = want_wpb;
    const int64_t used = (L.waves + wpb - 1) / wpb;
    Cfg& c = cfg[li];
    c.fn = lfn;
    c.wpb = wpb;
    c.blocks = (unsigned)((used + kXcds - 1) / kXcds * kXcds);   // see k_panel
    c.lds_bytes = wave_bytes * wpb + block_extra;
    const size_t res = std::max<size_t>(1, resident_waves(lfn, wave_bytes, block_extra, wpb));
    c.cost = (double)L.waves * L.rounds * L.cnt / (double)res;
  }
  LLFEM_TRY(lerr);
  // dynamic LDS above 64 KB: allow the kernel a CU's whole LDS (a limit, not
  // a reservation), so no later call lowers it below what another launch uses
  for (const Cfg& c : cfg)
    if (c.lds_bytes > 65536) {
      const hipError_t e =
          hipFuncSetAttribute((const void*)c.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsPerCU);
      if (e != hipSuccess) return fail(LLFEM_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    }
  // streams: each launch to the least-loaded stream, costliest first, and
  // issued costliest first on its stream, so the longest launch starts at
  // once instead of queueing behind a shorter one.  The cost counts the
  // resident waves, so an LDS-bound class (fewer waves per CU) weighs more
  // than its steps alone say.  LLFEM_PANEL_ORDER=plan issues in plan order
  // and balances by steps (A/B).
  std::vector<int> stream_of(plan.size(), 0);
  std::vector<size_t> issue(plan.size());
  std::iota(issue.begin(), issue.end(), size_t(0));
  if (nstreams > 1) {
    const char* ord = std::getenv("LLFEM_PANEL_ORDER");
    const bool plan_order = ord && std::strcmp(ord, "plan") == 0;
    auto cost = [&](size_t i) {
      return plan_order ? (double)plan[i].waves * plan[i].rounds * plan[i].cnt : cfg[i].cost;
    };
    std::vector<size_t> by_cost(plan.size());
    std::iota(by_cost.begin(), by_cost.end(), size_t(0));
    std::stable_sort(by_cost.begin(), by_cost.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
    if (!plan_order) issue = by_cost;
    std::vector<double> load(nstreams, 0.0);
    for (size_t i : by_cost) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      stream_of[i] = k;
      load[k] += cost(i);
    }
    LLFEM_HIP(hipEventRecord(ctx->fork_ev, st));   // after the geometry
    for (int k = 0; k + 1 < nstreams; ++k) LLFEM_HIP(hipStreamWaitEvent(ctx->side[k], ctx->fork_ev, 0));
  }
  for (size_t li : issue) {
    const PanelLaunch& L = plan[li];
    const Cfg& c = cfg[li];
    hipStream_t ls = stream_of[li] == 0 ? st : ctx->side[stream_of[li] - 1];
    P.rec = static_cast<const uint8_t*>(m->d_rec) + L.rec_base * m->rec_bytes;
    P.desc = m->d_desc + L.desc_base;
    P.waves = L.waves;
    P.rounds = L.rounds;
    P.cnt = L.cnt;
    P.max_r = L.max_r;
    P.lds_per_wave = npw * L.max_r * blkl + 9;   // + dummy block
    hipLaunchKernelGGL(c.fn, dim3(c.blocks), dim3(64 * c.wpb), c.lds_bytes, ls, P);
  }
  LLFEM_HIP(hipGetLastError());
  for (int k = 0; k + 1 < nstreams; ++k) {   // join
    LLFEM_HIP(hipEventRecord(ctx->join_ev[k], ctx->side[k]));
    LLFEM_HIP(hipStreamWaitEvent(st, ctx->join_ev[k], 0));
  }
  if (!highorder) LLFEM_HIP(hipEventRecord(ctx->ev[2], st));
  if (acc) {   // resident += this call's values (group order: the earlier groups first)
    k_add_into<<<(unsigned)((nnz + 255) / 256), 256, 0, st>>>(nnz, acc_tmp, *outp);
    LLFEM_HIP(hipGetLastError());
    LLFEM_HIP(hipStreamSynchronize(st));
    drelease(ctx, acc_tmp, (size_t)nnz);
  }
  if (op == LLFEM_OP_MASS && d->lumped) {
    if (!m->d_lumped) LLFEM_TRY(dalloc(ctx, &m->d_lumped, 3 * m->N));
    k_lump<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, m->d_adj_ptr, m->d_valM, m->d_lumped);
  }
  LLFEM_HIP(hipEventRecord(ctx->ev[3], st));
  if (dev_out) {
    // the mesh's own operator buffers are untouched
  } else {
    // complete = every element contributed since the last replacing call
    const int oi = op == LLFEM_OP_STIFFNESS ? 0 : 1;
    bool complete = true;
    if (d->element_scale || acc) {
      std::vector<uint8_t>& cov = m->covered[oi];
      if (!acc) cov.assign((size_t)m->E, 0);
      else if (cov.empty()) cov.assign((size_t)m->E, 1);   // added to a complete operator
      for (int64_t e = 0; e < m->E; ++e)
        if (!d->element_scale || d->element_scale[e] != 0.0) cov[e] = 1;
      for (int64_t e = 0; e < m->E && complete; ++e) complete = cov[e] != 0;
      if (complete) cov.clear();
    } else {
      m->covered[oi].clear();
    }
    m->partial[oi] = !complete;
    if (op == LLFEM_OP_STIFFNESS) {
      m->have_K = complete;
      m->solver_ready = false;
    } else {
      m->have_M = complete;
      m->have_lumped = complete && d->lumped != 0;
    }
  }
  if (nzval_out) {
    if (op == LLFEM_OP_MASS && d->lumped)
      LLFEM_TRY(copy_to_host(ctx, nzval_out, m->d_lumped, 3 * (size_t)m->N));
    else
      LLFEM_TRY(copy_to_host(ctx, nzval_out, acc ? *outp : target, (size_t)nnz));
  }
  LLFEM_HIP(hipEventSynchronize(ctx->ev[3]));
  if (stats) {
    float ms = 0.f, ms_main = 0.f;
    LLFEM_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
    LLFEM_HIP(hipEventElapsedTime(&ms_main, ctx->ev[1], ctx->ev[2]));
    stats->kernel_ms = ms;
    stats->main_kernel_ms = ms_main;
    stats->nnz = nnz;
    stats->elements = m->E;
    stats->alg_bytes = 24.0 * m->N + 4.0 * npe * m->E + 8.0 * nnz;
    // element integration flops: symmetric Gram (3n)(3n+1) q on the high-order
    // path; on the panel path 2*9 q flops per (a,b) block pair for non-affine
    // rules, 0 (table lookup) for affine ones
    const double n3 = 3.0 * npe;
    stats->alg_flops = highorder ? n3 * (n3 + 1.0) * q * (double)m->E
                                 : (affine ? 0.0 : 18.0 * q * (double)npe * npe * (double)m->E);
    stats->iterations = 0;
    stats->refinements = 0;
    stats->relres = 0;
    stats->setup_ms = 0;
    stats->rollbacks = 0;
    stats->restarts = 0;
    stats->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LLFEM_OK;
}

}  // namespace llfem

extern "C" int llfem_assemble(llfem_mesh* m, const llfem_op_desc* d, double* nzval_out,
                              llfem_stats* stats) {
  return assemble_impl(m, d, nzval_out, stats, nullptr);
}
