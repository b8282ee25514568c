// Mesh upload and the once-per-mesh structural pattern.
//
// The reference builds its pattern on the host per assembly call
// (build_csc_pattern, src/multifield.jl:2083-2250) and locates every node
// pair by binary search inside the element loop (find_csc_position,
// src/multifield.jl:2260-2283).  Here both are built once, on the device:
//   incidences  (node, element, local index) sorted by node then element;
//   adjacency   the sorted unique row nodes of every column node, i.e. the
//               node-level CSC pattern (K = adjacency x I_pdim^2);
//   inc_map     for incidence k of node j and element node b, the position
//               ("slot") of conn[e][b] in j's row list -- the element->nnz
//               map that replaces find_csc_position.
// Sorting uses hipcub radix sort; everything stays int32 on the device.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <numeric>

#include <memory>
#include <thread>

#include <cstdlib>

#include "common.h"

namespace llfem {

namespace {

__global__ void k_inc_keys(int64_t E, int npe, const int32_t* __restrict__ conn,
                           uint64_t* __restrict__ keys) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * npe) return;
  int64_t e = t / npe;
  int a = (int)(t - e * npe);
  uint32_t node = (uint32_t)conn[(int64_t)a * E + e];
  keys[t] = ((uint64_t)node << 32) | (uint32_t)t;
}

// ptr[v] = first position whose high word >= v, for v in [0, N].
__global__ void k_segment_ptr(int64_t n, int64_t N, const uint64_t* __restrict__ keys,
                              int64_t* __restrict__ ptr64, int32_t* __restrict__ ptr32) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  int64_t cur = k < n ? (int64_t)(keys[k] >> 32) : N;
  int64_t prev = k > 0 ? (int64_t)(keys[k - 1] >> 32) : -1;
  for (int64_t v = prev + 1; v <= cur; ++v) {
    if (ptr64) ptr64[v] = k;
    if (ptr32) ptr32[v] = (int32_t)k;
  }
}

__global__ void k_inc_split(int64_t n, int npe, const uint64_t* __restrict__ keys,
                            int32_t* __restrict__ elem, uint8_t* __restrict__ loc) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t low = (uint32_t)keys[k];
  elem[k] = (int32_t)(low / (uint32_t)npe);
  loc[k] = (uint8_t)(low % (uint32_t)npe);
}

__global__ void k_pair_keys(int64_t ninc, int npe, int64_t E, const uint64_t* __restrict__ inc_keys,
                            const int32_t* __restrict__ inc_elem, const int32_t* __restrict__ conn,
                            uint64_t* __restrict__ keys) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ninc * npe) return;
  int64_t k = t / npe;
  int b = (int)(t - k * npe);
  uint64_t j = inc_keys[k] >> 32;
  int32_t e = inc_elem[k];
  uint32_t i = (uint32_t)conn[(int64_t)b * E + e];
  keys[t] = (j << 32) | i;
}

__global__ void k_low_word(int64_t n, const uint64_t* __restrict__ keys, int32_t* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (int32_t)(uint32_t)keys[k];
}

__device__ __forceinline__ int lower_bound_i32(const int32_t* a, int n, int32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename MapT>
__global__ void k_inc_map(int64_t ninc, int npe, int64_t E, const uint64_t* __restrict__ inc_keys,
                          const int32_t* __restrict__ inc_elem, const int32_t* __restrict__ conn,
                          const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj,
                          MapT* __restrict__ map, int* __restrict__ err) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ninc * npe) return;
  int64_t k = t / npe;
  int b = (int)(t - k * npe);
  int64_t j = (int64_t)(inc_keys[k] >> 32);
  int32_t e = inc_elem[k];
  int32_t i = conn[(int64_t)b * E + e];
  int64_t p0 = adj_ptr[j];
  int n = (int)(adj_ptr[j + 1] - p0);
  int s = lower_bound_i32(adj + p0, n, i);
  if (s >= n || adj[p0 + s] != i) {
    atomicOr(err, 1);
    s = 0;
  }
  map[t] = (MapT)s;
}

__global__ void k_diag_slot(int64_t N, const int64_t* __restrict__ adj_ptr,
                            const int32_t* __restrict__ adj, int32_t* __restrict__ diag) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  int64_t p0 = adj_ptr[j];
  int n = (int)(adj_ptr[j + 1] - p0);
  diag[j] = n > 0 ? lower_bound_i32(adj + p0, n, (int32_t)j) : -1;
}

inline unsigned grid(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

inline int nbits(uint64_t v) {
  int b = 0;
  while (v) { ++b; v >>= 1; }
  return b ? b : 1;
}

int sort_keys(llfem_ctx* ctx, uint64_t*& keys, uint64_t*& tmp, int64_t n, int end_bit) {
  size_t tb = 0;
  LLFEM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys, tmp, (int)n, 0, end_bit,
                                              ctx->stream));
  void* d_tmp = nullptr;
  LLFEM_TRY(dmalloc(ctx, &d_tmp, tb));
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(d_tmp, tb, keys, tmp, (int)n, 0, end_bit,
                                                   ctx->stream);
  dfree(ctx, d_tmp, tb);
  if (e != hipSuccess) return fail(LLFEM_EHIP, std::string("radix sort: ") + hipGetErrorString(e));
  std::swap(keys, tmp);
  return LLFEM_OK;
}

// One record per (wavefront, round, step, lane group):
//   r = ((w*G + i)*cnt + t)*NPW + n   <->   node slot p = (w*G + i)*NPW + n.
// Compact record (npe <= 11, byte slots): {int32 e; uint8 a; uint8 slot[11]}.
// Generic record: {int32 e; int16 a; int16 pad; MapT slot[npe]} padded to 16 B.
// Null steps (node has fewer incidences than cnt, or padding node) carry
// a = all-ones and element 0.
template <typename MapT>
__global__ void k_build_records(int64_t nrec, int npe, int npw, int cnt, int rec_bytes,
                                int compact, int64_t nodes, const int32_t* __restrict__ order,
                                const int32_t* __restrict__ inc_ptr,
                                const int32_t* __restrict__ inc_elem,
                                const uint8_t* __restrict__ inc_loc, const MapT* __restrict__ map,
                                const int64_t* __restrict__ adj_ptr, uint8_t* __restrict__ rec,
                                int2* __restrict__ desc) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const int n = (int)(r % npw);
  const int64_t q1 = r / npw;
  const int t = (int)(q1 % cnt);
  const int64_t p = (q1 / cnt) * npw + n;
  int32_t e = 0, a = -1;
  int64_t k = -1;
  int32_t j = -1;
  if (p < nodes) {
    j = order[p];
    const int32_t k0 = inc_ptr[j];
    if (t < inc_ptr[j + 1] - k0) {
      k = k0 + t;
      e = inc_elem[k];
      a = inc_loc[k];
    }
  }
  uint8_t* o = rec + r * rec_bytes;
  for (int i = 0; i < rec_bytes; ++i) o[i] = 0;
  *reinterpret_cast<int32_t*>(o) = e;
  if (compact) {
    o[4] = (uint8_t)(a < 0 ? 0xff : a);
    if (k >= 0)
      for (int b = 0; b < npe; ++b) o[5 + b] = (uint8_t)map[k * npe + b];
  } else {
    *reinterpret_cast<int16_t*>(o + 4) = (int16_t)a;
    MapT* sl = reinterpret_cast<MapT*>(o + 8);
    if (k >= 0)
      for (int b = 0; b < npe; ++b) sl[b] = map[k * npe + b];
  }
  if (t == 0) {
    int2 d = make_int2(0, 0);
    if (j >= 0) d = make_int2((int32_t)adj_ptr[j], (int32_t)(adj_ptr[j + 1] - adj_ptr[j]));
    desc[p] = d;
  }
}

}  // namespace

// Plan the panel launches and build the processing-order records.
// Nodes are sorted by incidence count; consecutive counts are merged into one
// launch while the padding waste stays small (or the class is tiny), so a
// mesh needs only a handful of launches.
int build_records(llfem_mesh* m, const std::vector<int32_t>& h_inc_ptr) {
  llfem_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int64_t N = m->N;
  const int npe = m->npe;
  const int npw = 64 / npe;
  m->launches.clear();
  if (npw < 1) return LLFEM_OK;   // npe > 64: no node panels (high-order path)
  auto count = [&](int64_t j) { return h_inc_ptr[j + 1] - h_inc_ptr[j]; };
  auto rows = [&](int64_t j) { return (int32_t)(m->h_adj_ptr[j + 1] - m->h_adj_ptr[j]); };
  std::vector<int32_t> order;
  order.reserve(N);
  for (int64_t j = 0; j < N; ++j)
    if (count(j) > 0) order.push_back((int32_t)j);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return count(x) < count(y); });
  const bool compact = (npe <= 11 && m->map_bytes == 1);
  m->rec_bytes = compact ? 16 : (int)(((8 + (int64_t)npe * m->map_bytes) + 15) / 16 * 16);
  m->launches.clear();
  int64_t rec_total = 0, desc_total = 0;
  std::vector<std::pair<int64_t, int64_t>> ranges;   // order ranges of the launches
  const int64_t ncu = std::max(1, ctx->num_cus);
  for (int64_t p = 0; p < (int64_t)order.size();) {
    const int32_t cmin = count(order[p]);
    int64_t q = p;
    int32_t cmax = cmin, rmax = 0;
    while (q < (int64_t)order.size()) {
      const int32_t c = count(order[q]);
      // merge while the padded steps stay within 25% of the smallest count,
      // or while the launch is still tiny
      if (c > cmax && c * 4 > cmin * 5 && (q - p) >= 4096) break;
      cmax = std::max(cmax, c);
      rmax = std::max(rmax, rows(order[q]));
      ++q;
    }
    const int64_t nodes = q - p;
    const int blk = 9;   // LDS sized for stiffness; mass uses less
    const int64_t wave_bytes = 8 * ((int64_t)npw * rmax * blk + 9);
    const int64_t wpc = std::max<int64_t>(1, std::min<int64_t>(16, (160 * 1024) / wave_bytes));
    // wavefronts per launch: mult x resident waves (kPanelMult unless tuned;
    // LLFEM_PANEL_MULT=n sets it for A/B runs)
    const auto tu = ctx->tune.find(m->etype);
    int64_t mult = (tu != ctx->tune.end() && tu->second.mult > 0) ? tu->second.mult : kPanelMult;
    if (const char* env = std::getenv("LLFEM_PANEL_MULT")) mult = std::max(1, std::min(64, std::atoi(env)));
    const int64_t target = mult * ncu * wpc;
    int64_t G = (nodes + npw * target - 1) / (npw * target);
    G = std::max<int64_t>(1, std::min<int64_t>(G, 256));
    const int64_t per_wave = npw * G;
    const int64_t waves = (nodes + per_wave - 1) / per_wave;
    const int64_t slots = waves * per_wave;
    PanelLaunch L;
    L.rec_base = rec_total;
    L.desc_base = desc_total;
    L.waves = waves;
    L.cnt = cmax;
    L.rounds = (int32_t)G;
    L.max_r = rmax;
    L.nodes = (int32_t)nodes;
    m->launches.push_back(L);
    ranges.push_back({p, q});
    rec_total += slots * cmax;
    desc_total += slots;
    p = q;
  }
  m->rec_count = rec_total;
  m->desc_count = desc_total;
  LLFEM_TRY(dmalloc(ctx, &m->d_rec, (size_t)rec_total * m->rec_bytes));
  LLFEM_TRY(dalloc(ctx, &m->d_desc, (size_t)desc_total));
  int32_t* d_order = nullptr;
  LLFEM_TRY(dalloc(ctx, &d_order, order.size() + 1));
  LLFEM_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice, st));
  for (size_t li = 0; li < m->launches.size(); ++li) {
    const PanelLaunch& L = m->launches[li];
    const int64_t nrec = L.waves * npw * L.rounds * (int64_t)L.cnt;
    uint8_t* rec = static_cast<uint8_t*>(m->d_rec) + L.rec_base * m->rec_bytes;
    const unsigned g = (unsigned)((nrec + 255) / 256);
    if (m->map_bytes == 1)
      k_build_records<uint8_t><<<g, 256, 0, st>>>(nrec, npe, npw, L.cnt, m->rec_bytes, compact ? 1 : 0,
                                                  L.nodes, d_order + ranges[li].first, m->d_inc_ptr,
                                                  m->d_inc_elem, m->d_inc_loc,
                                                  (const uint8_t*)m->d_inc_map, m->d_adj_ptr, rec,
                                                  m->d_desc + L.desc_base);
    else
      k_build_records<uint16_t><<<g, 256, 0, st>>>(nrec, npe, npw, L.cnt, m->rec_bytes, 0, L.nodes,
                                                   d_order + ranges[li].first, m->d_inc_ptr,
                                                   m->d_inc_elem, m->d_inc_loc,
                                                   (const uint16_t*)m->d_inc_map, m->d_adj_ptr, rec,
                                                   m->d_desc + L.desc_base);
  }
  LLFEM_HIP(hipGetLastError());
  LLFEM_HIP(hipStreamSynchronize(st));
  drelease(ctx, d_order, order.size() + 1);
  return LLFEM_OK;
}

// Re-plan the panel launches (after a tuning change); the pattern and the
// element->nnz map are kept.
int rebuild_records(llfem_mesh* m) {
  llfem_ctx* ctx = m->ctx;
  if (m->npe > 64) return LLFEM_OK;
  std::vector<int32_t> h_inc_ptr(m->N + 1);
  LLFEM_HIP(hipMemcpyAsync(h_inc_ptr.data(), m->d_inc_ptr, sizeof(int32_t) * (m->N + 1),
                           hipMemcpyDeviceToHost, ctx->stream));
  LLFEM_HIP(hipStreamSynchronize(ctx->stream));
  if (m->d_rec) dfree(ctx, m->d_rec, (size_t)m->rec_count * m->rec_bytes);
  m->d_rec = nullptr;
  drelease(ctx, m->d_desc, m->desc_count);
  m->rec_count = m->desc_count = 0;
  return build_records(m, h_inc_ptr);
}

int build_pattern(llfem_mesh* m) {
  llfem_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int64_t N = m->N, E = m->E;
  const int npe = m->npe;
  const int64_t ninc = E * npe;
  const int64_t npair = ninc * npe;
  if (ninc >= (int64_t)1 << 31 || npair >= (int64_t)1 << 31)
    return fail(LLFEM_EUNSUP, "mesh too large for 32-bit incidence indexing (E*npe^2 >= 2^31)");
  m->ninc = ninc;
  const int end_bit = 32 + nbits((uint64_t)(N > 0 ? N - 1 : 0));

  // ---- incidences sorted by (node, element, local) -------------------------
  uint64_t *inc_keys = nullptr, *tmp = nullptr;
  LLFEM_TRY(dalloc(ctx, &inc_keys, ninc));
  LLFEM_TRY(dalloc(ctx, &tmp, ninc));
  k_inc_keys<<<grid(ninc), 256, 0, st>>>(E, npe, m->d_conn, inc_keys);
  LLFEM_TRY(sort_keys(ctx, inc_keys, tmp, ninc, end_bit));
  drelease(ctx, tmp, ninc);
  LLFEM_TRY(dalloc(ctx, &m->d_inc_ptr, N + 1));
  LLFEM_TRY(dalloc(ctx, &m->d_inc_elem, ninc));
  LLFEM_TRY(dalloc(ctx, &m->d_inc_loc, ninc));
  k_segment_ptr<<<grid(ninc + 1), 256, 0, st>>>(ninc, N, inc_keys, nullptr, m->d_inc_ptr);
  k_inc_split<<<grid(ninc), 256, 0, st>>>(ninc, npe, inc_keys, m->d_inc_elem, m->d_inc_loc);

  // ---- node adjacency = sorted unique (column node, row node) pairs --------
  uint64_t *pk = nullptr, *pt = nullptr;
  LLFEM_TRY(dalloc(ctx, &pk, npair));
  LLFEM_TRY(dalloc(ctx, &pt, npair));
  k_pair_keys<<<grid(npair), 256, 0, st>>>(ninc, npe, E, inc_keys, m->d_inc_elem, m->d_conn, pk);
  LLFEM_TRY(sort_keys(ctx, pk, pt, npair, end_bit));
  int* d_nsel = nullptr;
  LLFEM_TRY(dalloc(ctx, &d_nsel, 1));
  size_t tb = 0;
  LLFEM_HIP(hipcub::DeviceSelect::Unique(nullptr, tb, pk, pt, d_nsel, (int)npair, st));
  void* d_tmp = nullptr;
  LLFEM_TRY(dmalloc(ctx, &d_tmp, tb));
  LLFEM_HIP(hipcub::DeviceSelect::Unique(d_tmp, tb, pk, pt, d_nsel, (int)npair, st));
  dfree(ctx, d_tmp, tb);
  int nsel = 0;
  LLFEM_HIP(hipMemcpyAsync(&nsel, d_nsel, sizeof(int), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  drelease(ctx, d_nsel, 1);
  drelease(ctx, pk, npair);
  m->nadj = nsel;
  LLFEM_TRY(dalloc(ctx, &m->d_adj, (size_t)nsel));
  LLFEM_TRY(dalloc(ctx, &m->d_adj_ptr, N + 1));
  k_low_word<<<grid(nsel), 256, 0, st>>>(nsel, pt, m->d_adj);
  k_segment_ptr<<<grid(nsel + 1), 256, 0, st>>>(nsel, N, pt, m->d_adj_ptr, nullptr);
  drelease(ctx, pt, npair);

  // ---- host side: row counts, processing order, launch classes ------------
  m->h_adj_ptr.resize(N + 1);
  std::vector<int32_t> h_inc_ptr(N + 1);
  LLFEM_HIP(hipMemcpyAsync(m->h_adj_ptr.data(), m->d_adj_ptr, sizeof(int64_t) * (N + 1),
                           hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipMemcpyAsync(h_inc_ptr.data(), m->d_inc_ptr, sizeof(int32_t) * (N + 1),
                           hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  int32_t max_r = 0;
  for (int64_t j = 0; j < N; ++j)
    max_r = std::max<int32_t>(max_r, (int32_t)(m->h_adj_ptr[j + 1] - m->h_adj_ptr[j]));
  m->max_r = max_r;
  m->map_bytes = max_r <= 256 ? 1 : 2;
  if (max_r > 65535) return fail(LLFEM_EUNSUP, "node with more than 65535 neighbours");

  // ---- element->nnz map ------------------------------------------------------
  int* d_err = nullptr;
  LLFEM_TRY(dalloc(ctx, &d_err, 1));
  LLFEM_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
  LLFEM_TRY(dmalloc(ctx, &m->d_inc_map, (size_t)npair * m->map_bytes));
  if (m->map_bytes == 1)
    k_inc_map<uint8_t><<<grid(npair), 256, 0, st>>>(ninc, npe, E, inc_keys, m->d_inc_elem,
                                                    m->d_conn, m->d_adj_ptr, m->d_adj,
                                                    (uint8_t*)m->d_inc_map, d_err);
  else
    k_inc_map<uint16_t><<<grid(npair), 256, 0, st>>>(ninc, npe, E, inc_keys, m->d_inc_elem,
                                                     m->d_conn, m->d_adj_ptr, m->d_adj,
                                                     (uint16_t*)m->d_inc_map, d_err);
  LLFEM_TRY(dalloc(ctx, &m->d_diag_slot, N));
  k_diag_slot<<<grid(N), 256, 0, st>>>(N, m->d_adj_ptr, m->d_adj, m->d_diag_slot);
  LLFEM_TRY(build_records(m, h_inc_ptr));
  int herr = 0;
  LLFEM_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  LLFEM_HIP(hipStreamSynchronize(st));
  drelease(ctx, d_err, 1);
  drelease(ctx, inc_keys, ninc);
  LLFEM_HIP(hipGetLastError());
  if (herr) return fail(LLFEM_EPATTERN, "find_csc_position: row is missing from the CSC pattern.");
  return LLFEM_OK;
}

namespace {

// caller's node-major coordinates -> SoA x[N] y[N] z[N] and the padded
// node-major copy (x, y, z, 0) the affine geometry kernel gathers
__global__ void k_soa_coords(int64_t N, const double* __restrict__ xyz, double* __restrict__ soa,
                             double4* __restrict__ xyz4) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
  soa[j] = x;
  soa[N + j] = y;
  soa[2 * N + j] = z;
  xyz4[j] = make_double4(x, y, z, 0.0);
}

}  // namespace

}  // namespace llfem

using namespace llfem;

static void mesh_free(llfem_mesh* m) {
  llfem_ctx* c = m->ctx;
  drelease(c, m->d_xyz, 3 * m->N);
  drelease(c, m->d_xyz4, m->N);
  drelease(c, m->d_conn, m->E * m->npe);
  drelease(c, m->d_adj_ptr, m->N + 1);
  drelease(c, m->d_adj, m->nadj);
  drelease(c, m->d_diag_slot, m->N);
  drelease(c, m->d_inc_ptr, m->N + 1);
  drelease(c, m->d_inc_elem, m->ninc);
  drelease(c, m->d_inc_loc, m->ninc);
  if (m->d_inc_map) dfree(c, m->d_inc_map, (size_t)m->ninc * m->npe * m->map_bytes);
  m->d_inc_map = nullptr;
  if (m->d_rec) dfree(c, m->d_rec, (size_t)m->rec_count * m->rec_bytes);
  m->d_rec = nullptr;
  drelease(c, m->d_desc, m->desc_count);
  if (m->d_geo) dfree(c, m->d_geo, m->geo_bytes);
  if (m->d_tab) dfree(c, m->d_tab, m->tab_bytes);
  drelease(c, m->d_valK, 9 * m->nadj);
  drelease(c, m->d_valM, 3 * m->nadj);
  drelease(c, m->d_lumped, 3 * m->N);
  drelease(c, m->d_valK32, 9 * m->nadj);
  drelease(c, m->d_diagK, 3 * m->N);
  drelease(c, m->d_gslot, (size_t)m->E * m->npe * m->npe);
  drelease(c, m->d_gcol, 2 * (size_t)m->E * m->npe);
  drelease(c, m->d_mptr, (size_t)m->nmulti + 1);
  drelease(c, m->d_mref, 2 * (size_t)m->nmulti);
  if (m->d_ho) dfree(c, m->d_ho, m->ho_bytes);
  drelease(c, m->d_ho_perm, (size_t)m->npe);
  for (auto& kv : m->amg) {
    amg_free(c, kv.second.first);
    dfree(c, kv.second.second, 3 * (size_t)m->N);
  }
  m->amg.clear();
}

extern "C" {

int llfem_mesh_create(llfem_ctx* ctx, const double* xyz, int64_t N, int32_t elem_type, int32_t npe,
                      const int64_t* conn, int64_t E, int32_t pdim, int32_t one_based,
                      llfem_mesh** out) {
  if (!ctx || !xyz || !conn || !out) return fail(LLFEM_EINVAL, "llfem_mesh_create: NULL argument");
  *out = nullptr;
  if (N <= 0 || E <= 0 || npe <= 0) return fail(LLFEM_EINVAL, "llfem_mesh_create: empty mesh");
  if (pdim != 3) return fail(LLFEM_EUNSUP, "llfem_mesh_create: only pdim = 3 (:Solid) is supported");
  if (etype_dim(elem_type) != 3)
    return fail(LLFEM_EUNSUP, "llfem_mesh_create: unsupported volume element type " +
                                  std::to_string(elem_type));
  if (N >= ((int64_t)1 << 31) / 3) return fail(LLFEM_EUNSUP, "llfem_mesh_create: too many nodes");
  // connectivity: validate on the host and convert to SoA int32
  std::vector<int32_t> h_conn((size_t)E * npe);
  const int64_t base = one_based ? 1 : 0;
  for (int64_t e = 0; e < E; ++e)
    for (int a = 0; a < npe; ++a) {
      int64_t v = conn[e * npe + a] - base;
      if (v < 0 || v >= N)
        return fail(LLFEM_EINVAL, "build_csc_pattern: node tag " + std::to_string(v + base) +
                                      " is outside the trial-space node range");
      h_conn[(size_t)a * E + e] = (int32_t)v;
    }
  LLFEM_HIP(hipSetDevice(ctx->device));
  llfem_mesh* m = new llfem_mesh();
  m->ctx = ctx;
  m->N = N; m->E = E; m->npe = npe; m->etype = elem_type; m->pdim = pdim;
  m->h_conn0.resize(npe);
  for (int a = 0; a < npe; ++a) m->h_conn0[a] = h_conn[(size_t)a * E];
  int s = LLFEM_OK;
  double* d_aos = nullptr;
  do {
    if ((s = dalloc(ctx, &m->d_xyz, 3 * N))) break;
    if ((s = dalloc(ctx, &m->d_xyz4, N))) break;
    if ((s = dalloc(ctx, &m->d_conn, E * npe))) break;
    if ((s = dalloc(ctx, &d_aos, 3 * N))) break;
    if (hipMemcpyAsync(d_aos, xyz, sizeof(double) * 3 * N, hipMemcpyHostToDevice, ctx->stream) ||
        hipMemcpyAsync(m->d_conn, h_conn.data(), sizeof(int32_t) * E * npe, hipMemcpyHostToDevice,
                       ctx->stream)) {
      s = fail(LLFEM_EHIP, "llfem_mesh_create: upload failed");
      break;
    }
    k_soa_coords<<<(unsigned)((N + 255) / 256), 256, 0, ctx->stream>>>(N, d_aos, m->d_xyz, m->d_xyz4);
    if ((s = build_pattern(m))) break;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
      s = fail(LLFEM_EHIP, "llfem_mesh_create: device error");
      break;
    }
  } while (0);
  if (d_aos) dfree(ctx, d_aos, sizeof(double) * 3 * N);
  if (s != LLFEM_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    mesh_free(m);
    delete m;
    return s;
  }
  *out = m;
  return LLFEM_OK;
}

int llfem_mesh_destroy(llfem_mesh* m) {
  if (!m) return LLFEM_OK;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  mesh_free(m);
  delete m;
  return LLFEM_OK;
}

int llfem_mesh_pattern(llfem_mesh* m, int32_t op, int32_t one_based, int64_t* nnz, int64_t* colptr,
                       int64_t* rowval) {
  if (!m || !nnz) return fail(LLFEM_EINVAL, "llfem_mesh_pattern: NULL argument");
  const int64_t N = m->N;
  const int pd = m->pdim;
  const int blk = (op == LLFEM_OP_MASS) ? pd : pd * pd;   // entries per node pair
  *nnz = m->nadj * blk;
  if (!colptr && !rowval) return LLFEM_OK;
  const int64_t base = one_based ? 1 : 0;
  const int64_t nt = N < 65536 ? 1 : std::max<int64_t>(1, std::min<int64_t>(8, std::thread::hardware_concurrency()));
  auto on_threads = [&](int64_t n, auto&& f) {   // f(lo, hi) over contiguous ranges of [0, n)
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t) {
      const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      if (t + 1 == nt) f(lo, hi); else th.emplace_back(f, lo, hi);
    }
    for (auto& x : th) x.join();
  };
  std::unique_ptr<int32_t[]> adj;   // left uninitialised: the copy overwrites it
  if (rowval) {
    adj.reset(new int32_t[m->nadj]);
    // first touch on the host threads (the runtime's copy would otherwise
    // fault the pages in one at a time)
    uint8_t* p = reinterpret_cast<uint8_t*>(adj.get());
    const int64_t pages = (int64_t)((sizeof(int32_t) * m->nadj + 4095) / 4096);
    on_threads(pages, [p](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) p[i * 4096] = 0;
    });
    LLFEM_HIP(hipMemcpy(adj.get(), m->d_adj, sizeof(int32_t) * m->nadj, hipMemcpyDeviceToHost));
  }
  const auto& ap = m->h_adj_ptr;
  // columns of disjoint node ranges fill disjoint parts of colptr / rowval:
  // host threads take contiguous node ranges
  auto fill = [&](int64_t j0, int64_t j1) {
    for (int64_t j = j0; j < j1; ++j) {
      const int64_t R = ap[j + 1] - ap[j];
      const int64_t p0 = ap[j] * blk;
      for (int c = 0; c < pd; ++c) {
        const int64_t pc = p0 + c * R * (blk / pd);
        if (colptr) colptr[j * pd + c] = pc + base;
        if (rowval) {
          if (op == LLFEM_OP_MASS) {
            for (int64_t s = 0; s < R; ++s) rowval[pc + s] = (int64_t)adj[ap[j] + s] * pd + c + base;
          } else {
            for (int64_t s = 0; s < R; ++s)
              for (int d = 0; d < pd; ++d)
                rowval[pc + s * pd + d] = (int64_t)adj[ap[j] + s] * pd + d + base;
          }
        }
      }
    }
  };
  on_threads(N, fill);
  if (colptr) colptr[N * pd] = *nnz + base;
  return LLFEM_OK;
}

}  // extern "C"
