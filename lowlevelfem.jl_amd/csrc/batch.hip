// Batched assembly of many independent problems, and the launch-shape
// autotuner.
//
// Reference: each problem is one stiffnessMatrix / massMatrix call
// (src/linear.jl:70-87, 1843-2004); a caller with many problems loops over
// them.  Here all (mesh, operator) pairs share ONE persistent launch: every
// pair's panel launches become segments of one global list of wavefront work
// items, and resident wavefronts take items from a device queue (one 32-bit
// counter), so uneven problem sizes balance across the device and the launch
// overhead is paid once.  Every item runs the per-mesh wavefront body on the
// per-mesh records, so results do not depend on which wavefront takes which
// item (deterministic), and equal the per-mesh launch's; an affine mesh whose
// two operators use different rules shares one set of Jacobians (evaluated at
// the first rule's first point), which can move values by rounding only.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "common.h"
#include "panel.h"

namespace llfem {

namespace {

struct BatchSeg {
  PanelParams P;      // the panel launch this segment runs
  int64_t item0;      // first global work item of the segment
  int32_t span;       // launch waves per work item (the last item may have fewer)
  int32_t kind;       // kernel variant (batch_kind)
};

struct GeoSeg {
  int64_t e0;         // first global element of the segment
  int64_t E, N;
  int32_t npe, q, affine, pad;
  const double* xyz;
  const int32_t* conn;
  const double* dN;
  const double* w;
  double* geo;
};

int npe_index(int npe) {
  switch (npe) {
    case 4: return 0;
    case 6: return 1;
    case 8: return 2;
    case 10: return 3;
    default: return -1;
  }
}

constexpr int batch_kind(int npe_idx, int op, bool affine) { return (npe_idx * 3 + op) * 2 + (affine ? 1 : 0); }

template <int NPE, int OP, bool AFFINE>
__device__ __forceinline__ void run_item(const PanelParams& P, int64_t w0, int nw, double* wl) {
  panel_wave<NPE, uint8_t, OP, AFFINE, true, double>(P, w0, nw, wl);
}

// Persistent panel kernel.  Each wavefront repeatedly takes the next work item
// (`span` consecutive waves of one segment, run as one stream) from the
// queue; the item index is
// made wave-uniform with readfirstlane, so the segment search and the
// parameter loads are scalar.  Every wavefront exits once the queue passes
// `total`.
__global__ __launch_bounds__(256) void k_panel_batch(const BatchSeg* __restrict__ segs, int nseg,
                                                     uint32_t total, uint32_t* __restrict__ queue,
                                                     int lds_per_wave) {
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* wl = lds + (size_t)wave * lds_per_wave;
  for (;;) {
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(queue, 1u);
    const uint32_t item = __builtin_amdgcn_readfirstlane(got);
    if (item >= total) break;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segs[mid].item0 <= (int64_t)item) lo = mid; else hi = mid - 1;
    }
    const PanelParams P = segs[lo].P;
    const int span = segs[lo].span;
    const int64_t w = ((int64_t)item - segs[lo].item0) * span;
    const int nw = (int)min((int64_t)span, P.waves - w);
    switch (segs[lo].kind) {
      case batch_kind(0, 0, false): run_item<4, 0, false>(P, w, nw, wl); break;
      case batch_kind(0, 0, true): run_item<4, 0, true>(P, w, nw, wl); break;
      case batch_kind(0, 1, false): run_item<4, 1, false>(P, w, nw, wl); break;
      case batch_kind(0, 1, true): run_item<4, 1, true>(P, w, nw, wl); break;
      case batch_kind(1, 0, false): run_item<6, 0, false>(P, w, nw, wl); break;
      case batch_kind(1, 0, true): run_item<6, 0, true>(P, w, nw, wl); break;
      case batch_kind(1, 1, false): run_item<6, 1, false>(P, w, nw, wl); break;
      case batch_kind(1, 1, true): run_item<6, 1, true>(P, w, nw, wl); break;
      case batch_kind(2, 0, false): run_item<8, 0, false>(P, w, nw, wl); break;
      case batch_kind(2, 0, true): run_item<8, 0, true>(P, w, nw, wl); break;
      case batch_kind(2, 1, false): run_item<8, 1, false>(P, w, nw, wl); break;
      case batch_kind(2, 1, true): run_item<8, 1, true>(P, w, nw, wl); break;
      case batch_kind(3, 0, false): run_item<10, 0, false>(P, w, nw, wl); break;
      case batch_kind(3, 0, true): run_item<10, 0, true>(P, w, nw, wl); break;
      case batch_kind(3, 1, false): run_item<10, 1, false>(P, w, nw, wl); break;
      case batch_kind(3, 1, true): run_item<10, 1, true>(P, w, nw, wl); break;
      case batch_kind(0, kOpKM, true): run_item<4, kOpKM, true>(P, w, nw, wl); break;
      case batch_kind(1, kOpKM, true): run_item<6, kOpKM, true>(P, w, nw, wl); break;
      case batch_kind(2, kOpKM, true): run_item<8, kOpKM, true>(P, w, nw, wl); break;
      case batch_kind(3, kOpKM, true): run_item<10, kOpKM, true>(P, w, nw, wl); break;
      default: break;
    }
  }
}

// Geometry records of every batched mesh in one launch (thread per element).
__global__ void k_geometry_batch(const GeoSeg* __restrict__ segs, int nseg, int64_t total) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].e0 <= t) lo = mid; else hi = mid - 1;
  }
  const GeoSeg& g = segs[lo];
  geometry_one(g.E, g.N, g.npe, g.q, g.affine, (const double*)gp(g.xyz), (const int32_t*)gp(g.conn),
               (const double*)gp(g.dN), (const double*)gp(g.w), (double*)gp(g.geo), t - g.e0);
}

__global__ void k_zero_u32(uint32_t* p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

// Context-owned device copy of a table, keyed by content.
int cached_table(llfem_ctx* ctx, uint64_t key, const std::vector<double>& host, const double** out) {
  auto it = ctx->tables.find(key);
  if (it == ctx->tables.end()) {
    DevTable t;
    t.bytes = sizeof(double) * std::max<size_t>(host.size(), 1);
    LLFEM_TRY(dmalloc(ctx, (void**)&t.ptr, t.bytes));
    LLFEM_HIP(hipMemcpyAsync(t.ptr, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                             ctx->stream));
    it = ctx->tables.emplace(key, t).first;
  }
  *out = it->second.ptr;
  return LLFEM_OK;
}

int ensure_geo(llfem_ctx* ctx, llfem_mesh* m, size_t bytes) {
  if (m->d_geo && m->geo_bytes >= bytes) return LLFEM_OK;
  if (m->d_geo) dfree(ctx, m->d_geo, m->geo_bytes);
  m->d_geo = nullptr;
  m->geo_bytes = 0;
  LLFEM_TRY(dmalloc(ctx, (void**)&m->d_geo, bytes));
  m->geo_bytes = bytes;
  return LLFEM_OK;
}

int ensure_scratch(llfem_ctx* ctx, size_t bytes) {
  if (ctx->d_scratch && ctx->scratch_bytes >= bytes) return LLFEM_OK;
  if (ctx->d_scratch) dfree(ctx, ctx->d_scratch, ctx->scratch_bytes);
  ctx->d_scratch = nullptr;
  ctx->scratch_bytes = 0;
  LLFEM_TRY(dmalloc(ctx, &ctx->d_scratch, bytes));
  ctx->scratch_bytes = bytes;
  return LLFEM_OK;
}

uint64_t rule_key(const llfem_op_desc* d, int npe) {
  uint64_t key = hash_bytes(d->dN_ref, sizeof(double) * (size_t)d->q * npe * 3);
  return hash_bytes(d->weights, sizeof(double) * (size_t)d->q, key);
}

int check_desc(const llfem_op_desc* d) {
  if (!d) return fail(LLFEM_EINVAL, "llfem_assemble_batch: NULL descriptor");
  if (d->op != LLFEM_OP_STIFFNESS && d->op != LLFEM_OP_MASS)
    return fail(LLFEM_EINVAL, "llfem_assemble_batch: unknown operator");
  if (d->q <= 0 || d->q > 1024 || !d->weights || !d->dN_ref)
    return fail(LLFEM_EINVAL, "llfem_assemble_batch: bad quadrature tables");
  if (d->op == LLFEM_OP_MASS && !d->N_ref)
    return fail(LLFEM_EINVAL, "llfem_assemble_batch: mass needs N_ref");
  if (d->element_scale)
    return fail(LLFEM_EUNSUP, "llfem_assemble_batch: element_scale is for llfem_assemble only");
  return LLFEM_OK;
}

struct Item {
  llfem_mesh* m;
  const llfem_op_desc* d;
  uint64_t rkey;
  bool batched;
  bool affine;
  const double* rule;   // device [dN | w]
  const double* tab;    // device operator table
  int mass_pair = -1;   // stiffness item: the mass item of the same mesh fused into its launch
  bool fused = false;   // mass item assembled inside its stiffness item's launch
};

}  // namespace

}  // namespace llfem

using namespace llfem;

extern "C" int llfem_assemble_batch(int32_t n, llfem_mesh* const* meshes, const llfem_op_desc* descs,
                                    llfem_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  if (n < 0 || (n > 0 && (!meshes || !descs))) return fail(LLFEM_EINVAL, "llfem_assemble_batch: bad arguments");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return LLFEM_OK;
  llfem_ctx* ctx = meshes[0] ? meshes[0]->ctx : nullptr;
  for (int i = 0; i < n; ++i) {
    if (!meshes[i]) return fail(LLFEM_EINVAL, "llfem_assemble_batch: NULL mesh");
    if (meshes[i]->ctx != ctx) return fail(LLFEM_EINVAL, "llfem_assemble_batch: meshes of different contexts");
    LLFEM_TRY(check_desc(&descs[i]));
  }
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // LLFEM_BATCH_TRACE=1: wall time of the host phases on stderr
  const bool trace = std::getenv("LLFEM_BATCH_TRACE") != nullptr;
  auto tmark = [&](const char* what) {
    if (!trace) return;
    static thread_local auto last = std::chrono::steady_clock::now();
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[batch] %s %.3f ms\n", what, std::chrono::duration<double, std::milli>(now - last).count());
    last = now;
  };
  tmark("enter");

  // ---- classify; rule tables and affine checks ---------------------------------
  std::vector<Item> items(n);
  std::vector<std::pair<llfem_mesh*, uint64_t>> geo_owner;   // mesh -> rule of its records
  std::vector<int> pending;                                  // items whose rule needs a check
  for (int i = 0; i < n; ++i) {
    Item& it = items[i];
    it.m = meshes[i];
    it.d = &descs[i];
    const llfem_op_desc* d = it.d;
    llfem_mesh* m = it.m;
    const bool ho = (m->etype == 92 || m->etype == 93) && d->op == LLFEM_OP_STIFFNESS;
    it.batched = d->D == nullptr && d->precision == LLFEM_FP64 && npe_index(m->npe) >= 0 &&
                 m->map_bytes == 1 && !ho && !m->launches.empty() && m->pdim == 3;
    it.rkey = rule_key(d, m->npe);
    if (!it.batched) continue;
    std::vector<double> rule(d->dN_ref, d->dN_ref + (size_t)d->q * m->npe * 3);
    rule.insert(rule.end(), d->weights, d->weights + d->q);
    LLFEM_TRY(cached_table(ctx, it.rkey ^ 0x9e3779b97f4a7c15ull, rule, &it.rule));
    if (!m->rules[it.rkey].checked &&
        std::none_of(pending.begin(), pending.end(),
                     [&](int k) { return items[k].m == m && items[k].rkey == it.rkey; }))
      pending.push_back(i);
  }
  if (!pending.empty()) {
    int* d_flags = nullptr;
    LLFEM_TRY(dalloc(ctx, &d_flags, pending.size()));
    LLFEM_HIP(hipMemsetAsync(d_flags, 0, sizeof(int) * pending.size(), st));
    for (size_t k = 0; k < pending.size(); ++k) {
      const Item& it = items[pending[k]];
      k_geom_check<<<(unsigned)((it.m->E + 255) / 256), 256, 0, st>>>(it.m->E, it.m->N, it.m->npe, it.d->q,
                                                                      it.m->d_xyz, it.m->d_conn, it.rule,
                                                                      d_flags + k);
    }
    std::vector<int> hf(pending.size());
    LLFEM_HIP(hipMemcpyAsync(hf.data(), d_flags, sizeof(int) * hf.size(), hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipStreamSynchronize(st));
    drelease(ctx, d_flags, pending.size());
    for (size_t k = 0; k < pending.size(); ++k) {
      RuleState& rs = items[pending[k]].m->rules[items[pending[k]].rkey];
      rs.checked = true;
      rs.affine = !(hf[k] & 1);
      rs.singular = (hf[k] & 2) != 0;
    }
  }
  for (Item& it : items) {
    if (!it.batched) continue;
    const RuleState& rs = it.m->rules[it.rkey];
    if (rs.singular)
      return fail(LLFEM_EGEOM, "compute_element_geometry!: singular geometry at integration point.");
    it.affine = rs.affine;
    // geometry records are per mesh: the first item of a mesh defines them;
    // another rule can share them only if both are affine (one record per
    // element, independent of the points)
    auto go = std::find_if(geo_owner.begin(), geo_owner.end(),
                           [&](const std::pair<llfem_mesh*, uint64_t>& x) { return x.first == it.m; });
    if (go == geo_owner.end()) {
      geo_owner.push_back({it.m, it.rkey});
    } else if (go->second != it.rkey && !(it.affine && it.m->rules[go->second].affine)) {
      it.batched = false;
      continue;
    }
    // operator table, shared by all meshes with the same (rule, operator)
    uint64_t tkey = hash_bytes(&it.d->op, sizeof(int32_t), it.rkey);
    tkey = hash_bytes(&it.affine, sizeof(bool), tkey);
    tkey = hash_bytes(&it.m->npe, sizeof(int32_t), tkey);
    if (it.d->op == LLFEM_OP_MASS)
      tkey = hash_bytes(it.d->N_ref, sizeof(double) * (size_t)it.d->q * it.m->npe, tkey);
    const auto found = ctx->tables.find(tkey);
    if (found != ctx->tables.end()) it.tab = found->second.ptr;
    else LLFEM_TRY(cached_table(ctx, tkey, operator_table(it.d, it.m->npe, it.affine), &it.tab));
  }

  tmark("classify + tables");
  // ---- geometry segments and panel segments ------------------------------------
  std::vector<GeoSeg> gsegs;
  std::vector<BatchSeg> bsegs;
  int64_t etot = 0, wtot = 0;
  size_t max_lds = 0;
  double alg_bytes = 0.0;
  for (auto& go : geo_owner) {
    llfem_mesh* m = go.first;
    const Item* it = nullptr;
    for (const Item& x : items)
      if (x.batched && x.m == m) { it = &x; break; }
    const int np = it->affine ? 1 : it->d->q;
    LLFEM_TRY(ensure_geo(ctx, m, sizeof(double) * kGeoStride * (size_t)m->E * np));
    GeoSeg g{};
    g.e0 = etot;
    g.E = m->E;
    g.N = m->N;
    g.npe = m->npe;
    g.q = it->d->q;
    g.affine = it->affine ? 1 : 0;
    g.xyz = m->d_xyz;
    g.conn = m->d_conn;
    g.dN = it->rule;
    g.w = it->rule + (size_t)it->d->q * m->npe * 3;
    g.geo = m->d_geo;
    gsegs.push_back(g);
    etot += m->E;
  }
  // stiffness and mass of the same affine mesh run as one fused pass
  for (size_t i = 0; i < items.size(); ++i) {
    Item& k = items[i];
    if (!k.batched || !k.affine || k.d->op != LLFEM_OP_STIFFNESS || k.mass_pair >= 0) continue;
    for (size_t j = 0; j < items.size(); ++j) {
      Item& mi = items[j];
      if (j == i || !mi.batched || !mi.affine || mi.fused || mi.d->op != LLFEM_OP_MASS || mi.m != k.m) continue;
      k.mass_pair = (int)j;
      mi.fused = true;
      break;
    }
  }
  for (Item& it : items) {
    if (!it.batched || it.fused) continue;
    llfem_mesh* m = it.m;
    const llfem_op_desc* d = it.d;
    const int op = d->op;
    const Item* mp = it.mass_pair >= 0 ? &items[it.mass_pair] : nullptr;
    if (mp && !m->d_valM) LLFEM_TRY(dalloc(ctx, &m->d_valM, (size_t)m->nadj * 3));
    double** outp = op == LLFEM_OP_STIFFNESS ? &m->d_valK : &m->d_valM;
    const int blk = op == LLFEM_OP_STIFFNESS ? 9 : 3;
    if (!*outp) LLFEM_TRY(dalloc(ctx, outp, (size_t)m->nadj * blk));
    const int npw = 64 / m->npe;
    const int blkl = mp ? 10 : (op == LLFEM_OP_STIFFNESS ? 9 : 1);
    for (const PanelLaunch& L : m->launches) {
      BatchSeg s{};
      PanelParams& P = s.P;
      P.rec = static_cast<const uint8_t*>(m->d_rec) + L.rec_base * m->rec_bytes;
      P.rec_bytes = m->rec_bytes;
      P.desc = m->d_desc + L.desc_base;
      P.waves = L.waves;
      P.rounds = L.rounds;
      P.cnt = L.cnt;
      P.max_r = L.max_r;
      P.geo = m->d_geo;
      P.E = m->E;
      P.q = d->q;
      P.tab = it.tab;
      P.tabf = reinterpret_cast<const float*>(it.tab);   // FP32 variants are not batched
      if (op == LLFEM_OP_STIFFNESS) {
        // elastic_constants (src/general.jl:171-175)
        P.mu = d->E / (2.0 * (1.0 + d->nu));
        P.lam = 2.0 * P.mu * d->nu / (1.0 - 2.0 * d->nu);
      }
      P.rho = d->rho;
      P.out = *outp;
      if (mp) {
        P.rho = mp->d->rho;
        P.tab2 = mp->tab;
        P.out2 = m->d_valM;
      }
      P.lds_per_wave = npw * L.max_r * blkl + 9;
      max_lds = std::max<size_t>(max_lds, (size_t)P.lds_per_wave);
      s.kind = batch_kind(npe_index(m->npe), mp ? kOpKM : op, it.affine);
      bsegs.push_back(s);
      wtot += L.waves;
    }
    alg_bytes += 24.0 * m->N + 4.0 * m->npe * m->E + 8.0 * m->nadj * blk;
    if (mp) alg_bytes += 24.0 * m->N + 4.0 * m->npe * m->E + 8.0 * m->nadj * 3;
  }
  // one wavefront per block: each wave takes work items from the queue on its
  // own, and a CU refills a finished wave's slot at once (4-wave blocks: the
  // persistent kernel 10.05 -> 9.25 ms on the 256-problem batch; 2-wave
  // blocks 10.5 ms).  The kernel's registers allow two waves per SIMD.
  const size_t wave_bytes = sizeof(double) * max_lds;
  if (!bsegs.empty() && wave_bytes > 160 * 1024)
    return fail(LLFEM_EUNSUP, "llfem_assemble_batch: node panel exceeds LDS");
  int wpb = 1;
  if (const char* env = std::getenv("LLFEM_BATCH_WPB")) wpb = std::max(1, std::min(4, std::atoi(env)));   // A/B
  while (wpb > 1 && wpb * wave_bytes > 160 * 1024) --wpb;
  const size_t lds_bytes = wave_bytes * wpb;
  int per_cu = std::max<int>(1, (int)((160 * 1024) / std::max<size_t>(lds_bytes, 1)));
  {   // blocks resident per CU: LDS, and the kernel's registers
    (void)hipFuncSetAttribute((const void*)k_panel_batch, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_panel_batch, 64 * wpb, lds_bytes) ==
            hipSuccess && nb > 0)
      per_cu = std::min(per_cu, nb);
  }
  // work items: about 8 per resident wavefront, each `span` launch waves
  const int64_t resident = (int64_t)std::max(1, ctx->num_cus) * per_cu * wpb;
  const int64_t span = std::max<int64_t>(1, wtot / (8 * resident));
  int64_t items_total = 0;
  for (BatchSeg& s : bsegs) {
    s.span = (int32_t)std::min<int64_t>(span, std::max<int64_t>(1, s.P.waves));
    s.item0 = items_total;
    items_total += (s.P.waves + s.span - 1) / s.span;
  }
  if (items_total >= ((int64_t)1 << 31)) return fail(LLFEM_EUNSUP, "llfem_assemble_batch: too many work items");

  tmark("segments");
  float ms_all = 0.f, ms_main = 0.f;
  if (!bsegs.empty()) {
    const size_t gbytes = sizeof(GeoSeg) * gsegs.size(), bbytes = sizeof(BatchSeg) * bsegs.size();
    const size_t boff = (gbytes + 255) / 256 * 256, qoff = boff + (bbytes + 255) / 256 * 256;
    LLFEM_TRY(ensure_scratch(ctx, qoff + 256));
    char* sc = static_cast<char*>(ctx->d_scratch);
    GeoSeg* d_gsegs = reinterpret_cast<GeoSeg*>(sc);
    BatchSeg* d_bsegs = reinterpret_cast<BatchSeg*>(sc + boff);
    uint32_t* d_queue = reinterpret_cast<uint32_t*>(sc + qoff);
    LLFEM_HIP(hipMemcpyAsync(d_gsegs, gsegs.data(), gbytes, hipMemcpyHostToDevice, st));
    LLFEM_HIP(hipMemcpyAsync(d_bsegs, bsegs.data(), bbytes, hipMemcpyHostToDevice, st));
    const int64_t want = (items_total + wpb - 1) / wpb;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->num_cus * per_cu, want));

    LLFEM_HIP(hipEventRecord(ctx->ev[0], st));
    k_geometry_batch<<<(unsigned)((etot + 255) / 256), 256, 0, st>>>(d_gsegs, (int)gsegs.size(), etot);
    k_zero_u32<<<1, 64, 0, st>>>(d_queue);
    LLFEM_HIP(hipEventRecord(ctx->ev[1], st));
    hipLaunchKernelGGL(k_panel_batch, dim3(blocks), dim3(64 * wpb), lds_bytes, st, d_bsegs,
                       (int)bsegs.size(), (uint32_t)items_total, d_queue, (int)max_lds);
    LLFEM_HIP(hipGetLastError());
    LLFEM_HIP(hipEventRecord(ctx->ev[2], st));
    for (Item& it : items) {
      if (!it.batched) continue;
      llfem_mesh* m = it.m;
      if (it.d->op == LLFEM_OP_STIFFNESS) {
        m->have_K = true;
        m->solver_ready = false;
      } else {
        m->have_M = true;
        m->have_lumped = it.d->lumped != 0;
        if (it.d->lumped) {
          if (!m->d_lumped) LLFEM_TRY(dalloc(ctx, &m->d_lumped, 3 * m->N));
          k_lump<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, m->d_adj_ptr, m->d_valM, m->d_lumped);
        }
      }
    }
    LLFEM_HIP(hipEventRecord(ctx->ev[3], st));
    tmark("upload + launches");
    LLFEM_HIP(hipEventSynchronize(ctx->ev[3]));
    tmark("wait for the device");
    LLFEM_HIP(hipEventElapsedTime(&ms_all, ctx->ev[0], ctx->ev[3]));
    LLFEM_HIP(hipEventElapsedTime(&ms_main, ctx->ev[1], ctx->ev[2]));
  }
  double kernel_ms = ms_all, main_ms = ms_main;
  int64_t nnz = 0, elements = 0;
  for (Item& it : items) {
    const int blk = it.d->op == LLFEM_OP_STIFFNESS ? 9 : 3;
    if (it.batched) {
      nnz += it.m->nadj * blk;
      elements += it.m->E;
      continue;
    }
    llfem_stats s{};
    LLFEM_TRY(llfem_assemble(it.m, it.d, nullptr, &s));
    kernel_ms += s.kernel_ms;
    alg_bytes += s.alg_bytes;
    nnz += s.nnz;
    elements += s.elements;
  }
  if (stats) {
    stats->kernel_ms = kernel_ms;
    stats->main_kernel_ms = main_ms;
    stats->nnz = nnz;
    stats->elements = elements;
    stats->alg_bytes = alg_bytes;
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return LLFEM_OK;
}

extern "C" int llfem_mesh_values(llfem_mesh* m, int32_t op, int32_t lumped, double* nzval_out) {
  if (!m || !nzval_out) return fail(LLFEM_EINVAL, "llfem_mesh_values: NULL argument");
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  const double* src = nullptr;
  size_t count = 0;
  if (op == LLFEM_OP_STIFFNESS) {
    if (!m->have_K) return fail(LLFEM_EINVAL, "llfem_mesh_values: no stiffness assembled on this mesh");
    src = m->d_valK;
    count = (size_t)m->nadj * 9;
  } else if (op == LLFEM_OP_MASS) {
    if (!m->have_M) return fail(LLFEM_EINVAL, "llfem_mesh_values: no mass assembled on this mesh");
    if (lumped) {
      if (!m->have_lumped || !m->d_lumped)
        return fail(LLFEM_EINVAL, "llfem_mesh_values: the last mass assembly was not lumped");
      src = m->d_lumped;
      count = (size_t)m->N * 3;
    } else {
      src = m->d_valM;
      count = (size_t)m->nadj * 3;
    }
  } else {
    return fail(LLFEM_EINVAL, "llfem_mesh_values: unknown operator");
  }
  return copy_to_host(ctx, nzval_out, src, count);
}

extern "C" int llfem_autotune(llfem_mesh* m, const llfem_op_desc* d, llfem_tune_result* out) {
  if (!m || !d) return fail(LLFEM_EINVAL, "llfem_autotune: NULL argument");
  if (m->launches.empty())
    return fail(LLFEM_EUNSUP, "llfem_autotune: this element type has no panel launches to tune");
  llfem_ctx* ctx = m->ctx;
  const int et = m->etype;
  const Tune saved = ctx->tune.count(et) ? ctx->tune[et] : Tune{};
  auto time_shape = [&](const Tune& t, double* ms) -> int {
    ctx->tune[et] = t;
    llfem_stats s{};
    LLFEM_TRY(llfem_assemble(m, d, nullptr, &s));   // warm-up
    double best = 1e30;
    for (int r = 0; r < 3; ++r) {
      LLFEM_TRY(llfem_assemble(m, d, nullptr, &s));
      best = std::min(best, s.main_kernel_ms);
    }
    *ms = best;
    return LLFEM_OK;
  };
  const int mults[] = {1, 2, 4, 8, 16, 32};
  const int wpbs[] = {0, 1, 2, 4, 8};
  Tune best{kPanelMult, 0, 1e30};
  double default_ms = 0.0;
  int cand = 0;
  for (int mult : mults) {
    ctx->tune[et] = Tune{mult, 0, 0.0};
    int rc = rebuild_records(m);
    if (rc != LLFEM_OK) { ctx->tune[et] = saved; return rc; }
    for (int wpb : wpbs) {
      double ms = 0.0;
      rc = time_shape(Tune{mult, wpb, 0.0}, &ms);
      if (rc != LLFEM_OK) { ctx->tune[et] = saved; rebuild_records(m); return rc; }
      ++cand;
      if (mult == kPanelMult && wpb == 0) default_ms = ms;
      if (ms < best.ms) best = Tune{mult, wpb, ms};
    }
  }
  // the best shape, then the class launches' stream count x round kernels
  // (the device time of an assembly spans every stream's launches)
  ctx->tune[et] = best;
  LLFEM_TRY(rebuild_records(m));
  for (int streams = 1; streams <= 3; ++streams) {
    for (int rounds = 1; rounds >= 0; --rounds) {
      Tune t = best;
      t.streams = streams;
      t.rounds = rounds;
      double ms = 0.0;
      const int rc = time_shape(t, &ms);
      if (rc != LLFEM_OK) { ctx->tune[et] = saved; rebuild_records(m); return rc; }
      ++cand;
      if (ms < best.ms) {
        t.ms = ms;
        best = t;
      }
    }
  }
  ctx->tune[et] = best;
  if (out) {
    out->elem_type = et;
    out->mult = best.mult;
    out->waves_per_block = best.wpb;
    out->candidates = cand;
    out->best_ms = best.ms;
    out->default_ms = default_ms;
    out->streams = best.streams > 0 ? best.streams : 2;
    out->round_kernels = best.rounds == 0 ? 0 : 1;
  }
  return LLFEM_OK;
}
