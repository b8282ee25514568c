// Assemble-and-solve server: a device-resident request loop with continuous
// batching.
//
// Reference: each request is the reference's script
//   K = stiffnessMatrix(problem)                      (src/linear.jl:70-87)
//   f = loadVector(problem, [load("face", fy = t)])   (src/linear.jl:3088-3434)
//   u = solveDisplacement(K, f, support = [...])      (src/linear.jl:3711-3743)
// run for one structured box.  The server keeps what does not change between
// requests on the device: the mesh with its CSC pattern and element->nnz map
// (keyed by mesh content), the per-node face-load weights and the
// constrained-DoF masks.  Per request it assembles K (its own E, nu), builds
// f and adds the system to the batched mixed-precision PCG engine.
//
// Continuous batching: a worker thread alternates admission and a chunk of
// PCG iterations over ALL admitted systems (one launch per kernel covers
// them all).  Requests that arrive while others run are admitted at the next
// chunk boundary when a slot (max_active) and device memory allow; systems
// leave the batch as they converge.  Admission is by the context's HBM
// budget: a request whose allocations do not fit is rolled back and retried
// after running requests release memory (unused cached meshes are evicted
// first).  A system's arithmetic does not depend on the other systems in the
// batch, so a request's result is the same alone and in any mix.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

namespace llfem {

namespace {

using Clock = std::chrono::steady_clock;

// LLFEM_SERVE_TRACE=1: per-request admission phases and per-chunk times on stderr.
bool trace_on() {
  static const bool on = [] {
    const char* e = std::getenv("LLFEM_SERVE_TRACE");
    return e && e[0] == '1';
  }();
  return on;
}

double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

// trace time stamp: ms since the first traced event of the process
double trace_t() {
  static const Clock::time_point t0 = Clock::now();
  return ms_since(t0);
}

__global__ void k_load_from_weights(int64_t N, const double* __restrict__ s, double tx, double ty, double tz,
                                    double* __restrict__ f) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double w = s[j];
  f[3 * j] = tx * w;
  f[3 * j + 1] = ty * w;
  f[3 * j + 2] = tz * w;
}

__global__ void k_mask_from_list(int64_t nfix, const int64_t* __restrict__ idx, int64_t n, uint8_t* __restrict__ mask,
                                 int* __restrict__ err) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfix) return;
  const int64_t i = idx[k];
  if (i < 0 || i >= n) { atomicOr(err, 1); return; }
  mask[i] = 1;
}

// AMG hierarchy of one (mesh, support): built from the stiffness of a fixed
// material (E = 1, nu = 0.3), so it depends on the mesh and support only and
// every request on them gets the same preconditioner.
struct AmgEntry {
  AmgHier* h = nullptr;
  AmgJob* job = nullptr;    // host build in flight
  bool failed = false;      // build failed: requests fall back to Jacobi
};

struct CachedMesh {
  llfem_mesh* m = nullptr;
  int users = 0;
  Clock::time_point last;
  std::map<uint64_t, std::pair<double*, int64_t>> weights;   // face key -> per-node load weights
  std::map<uint64_t, uint8_t*> masks;                        // support key -> constrained-DoF mask
  std::map<uint64_t, AmgEntry> amg;                          // support key -> AMG hierarchy
};

void free_cached(llfem_ctx* ctx, CachedMesh& c) {
  for (auto& a : c.amg) {
    amg_job_drop(a.second.job);
    amg_free(ctx, a.second.h);
  }
  c.amg.clear();
  for (auto& w : c.weights) dfree(ctx, w.second.first, sizeof(double) * (size_t)w.second.second);
  for (auto& mk : c.masks) dfree(ctx, mk.second, 3 * (size_t)c.m->N);
  llfem_mesh_destroy(c.m);
}

struct Req {
  int64_t id = 0;
  llfem_request in{};
  Clock::time_point t_submit, t_admit, t_done;
  bool done = false;
  int rc = LLFEM_OK;
  std::string err;
  // device state while active
  uint64_t mesh_key = 0;
  CachedMesh* cm = nullptr;
  bool cache_hit = false;
  double* d_K = nullptr;
  double* d_f = nullptr;
  double* d_u = nullptr;
  SolveSys* sys = nullptr;
  BatchSolver* eng = nullptr;
  SysSpec spec;              // the system (kept while it waits for its AMG hierarchy)
  uint64_t support_key = 0;
  bool waiting = false;      // admitted, its AMG hierarchy still building
  int64_t cm_nodes = 0;      // mesh nodes (trace)
  int64_t iterations = 0, updates = 0, rollbacks = 0, restarts = 0;
  double relres = 0.0;
};

}  // namespace

}  // namespace llfem

struct llfem_server {
  llfem_ctx* ctx = nullptr;
  int32_t max_active = 64;
  int32_t chunk = 8;         // iterations between host checks (AMG systems update every ~10)
  std::unique_ptr<llfem::BatchSolver> eng32, eng64;   // FP32 / FP64 Krylov vectors
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<llfem::Req>> pending;
  std::vector<std::shared_ptr<llfem::Req>> active;
  std::map<int64_t, std::shared_ptr<llfem::Req>> all;
  std::map<uint64_t, llfem::CachedMesh> meshes;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;
};

namespace llfem {

namespace {

uint64_t mesh_hash(const llfem_request& r) {
  if (r.mesh_key) return r.mesh_key;
  uint64_t h = hash_bytes(&r.elem_type, sizeof(r.elem_type));
  h = hash_bytes(&r.npe, sizeof(r.npe), h);
  h = hash_bytes(&r.N, sizeof(r.N), h);
  h = hash_bytes(&r.E, sizeof(r.E), h);
  h = hash_bytes(r.xyz, sizeof(double) * 3 * (size_t)r.N, h);
  return hash_bytes(r.conn, sizeof(int64_t) * (size_t)r.E * r.npe, h);
}

uint64_t face_hash(const llfem_request& r) {
  uint64_t h = hash_bytes(&r.face_type, sizeof(r.face_type));
  h = hash_bytes(r.face_conn, sizeof(int64_t) * (size_t)r.F * r.nf, h);
  h = hash_bytes(r.face_weights, sizeof(double) * r.face_q, h);
  return hash_bytes(r.face_N_ref, sizeof(double) * (size_t)r.face_q * r.nf, h);
}

uint64_t support_hash(const llfem_request& r) {
  uint64_t h = hash_bytes(&r.nfixed, sizeof(r.nfixed));
  return hash_bytes(r.fixed, sizeof(int64_t) * (size_t)r.nfixed, h);
}

void free_request_buffers(llfem_ctx* ctx, Req* q) {
  const int64_t n = 3 * (q->cm && q->cm->m ? q->cm->m->N : 0);
  if (q->d_K) dfree(ctx, q->d_K, sizeof(double) * 9 * (size_t)q->cm->m->nadj);
  if (q->d_f) dfree(ctx, q->d_f, sizeof(double) * (size_t)n);
  if (q->d_u) dfree(ctx, q->d_u, sizeof(double) * (size_t)n);
  q->d_K = q->d_f = q->d_u = nullptr;
}

void release_mesh_use(Req* q) {
  if (q->cm) {
    q->cm->users--;
    q->cm->last = Clock::now();
  }
}

// Evict cached meshes that no active request uses.  Returns true if any was freed.
bool evict_unused(llfem_server* S) {
  bool any = false;
  for (auto it = S->meshes.begin(); it != S->meshes.end();) {
    CachedMesh& c = it->second;
    if (c.users == 0) {
      free_cached(S->ctx, c);
      it = S->meshes.erase(it);
      any = true;
    } else {
      ++it;
    }
  }
  return any;
}

// Set up one request on the device.  On failure everything allocated here is
// released again and the error code is returned.
int admit(llfem_server* S, Req* q) {
  llfem_ctx* ctx = S->ctx;
  hipStream_t st = ctx->stream;
  const llfem_request& r = q->in;
  const Clock::time_point t0 = Clock::now();
  double t_mesh = 0, t_wm = 0, t_k = 0, t_amg = 0;
  q->mesh_key = mesh_hash(r);
  auto it = S->meshes.find(q->mesh_key);
  if (it == S->meshes.end()) {
    llfem_mesh* m = nullptr;
    LLFEM_TRY(llfem_mesh_create(ctx, r.xyz, r.N, r.elem_type, r.npe, r.conn, r.E, 3, 0, &m));
    CachedMesh c;
    c.m = m;
    it = S->meshes.emplace(q->mesh_key, std::move(c)).first;
    q->cache_hit = false;
  } else {
    q->cache_hit = true;
  }
  CachedMesh* cm = &it->second;
  t_mesh = ms_since(t0);
  q->cm = cm;
  q->cm_nodes = cm->m->N;
  cm->users++;
  cm->last = Clock::now();
  llfem_mesh* m = cm->m;
  const int64_t n = 3 * m->N;
  int rc = LLFEM_OK;
  do {
    // per-node face-load weights (the load for traction (1,0,0), x components)
    const uint64_t fk = face_hash(r);
    auto wi = cm->weights.find(fk);
    if (wi == cm->weights.end()) {
      std::vector<double> f1((size_t)n);
      const double t1[3] = {1.0, 0.0, 0.0};
      if ((rc = llfem_load_traction(m, r.face_type, r.nf, r.face_conn, r.F, 0, r.face_q, r.face_weights,
                                    r.face_N_ref, r.face_dN_ref, t1, f1.data())))
        break;
      std::vector<double> w((size_t)m->N);
      for (int64_t j = 0; j < m->N; ++j) w[j] = f1[3 * j];
      double* d_w = nullptr;
      if ((rc = dalloc(ctx, &d_w, (size_t)m->N))) break;
      if (hipMemcpyAsync(d_w, w.data(), sizeof(double) * m->N, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        dfree(ctx, d_w, sizeof(double) * m->N);
        rc = fail(LLFEM_EHIP, "server: load upload failed");
        break;
      }
      wi = cm->weights.emplace(fk, std::make_pair(d_w, m->N)).first;
    }
    // constrained-DoF mask
    const uint64_t sk = support_hash(r);
    auto mi = cm->masks.find(sk);
    if (mi == cm->masks.end()) {
      uint8_t* d_mask = nullptr;
      int64_t* d_idx = nullptr;
      int* d_err = nullptr;
      if ((rc = dalloc(ctx, &d_mask, (size_t)n))) break;
      if ((rc = dalloc(ctx, &d_idx, (size_t)r.nfixed + 1)) || (rc = dalloc(ctx, &d_err, 1))) {
        drelease(ctx, d_mask, (size_t)n);
        drelease(ctx, d_idx, (size_t)r.nfixed + 1);
        break;
      }
      int herr = 0;
      hipError_t e = hipMemsetAsync(d_mask, 0, (size_t)n, st);
      if (e == hipSuccess) e = hipMemsetAsync(d_err, 0, sizeof(int), st);
      if (e == hipSuccess && r.nfixed > 0)
        e = hipMemcpyAsync(d_idx, r.fixed, sizeof(int64_t) * r.nfixed, hipMemcpyHostToDevice, st);
      if (e == hipSuccess && r.nfixed > 0)
        k_mask_from_list<<<(unsigned)((r.nfixed + 255) / 256), 256, 0, st>>>(r.nfixed, d_idx, n, d_mask, d_err);
      if (e == hipSuccess) e = hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      drelease(ctx, d_idx, (size_t)r.nfixed + 1);
      drelease(ctx, d_err, 1);
      if (e != hipSuccess || herr) {
        drelease(ctx, d_mask, (size_t)n);
        rc = fail(e != hipSuccess ? LLFEM_EHIP : LLFEM_EINVAL, "server: constrained DoF out of range");
        break;
      }
      mi = cm->masks.emplace(sk, d_mask).first;
    }
    t_wm = ms_since(t0);
    // K for this request's material, f = t (x) weights, u
    if ((rc = dalloc(ctx, &q->d_K, 9 * (size_t)m->nadj))) break;
    if ((rc = dalloc(ctx, &q->d_f, (size_t)n))) break;
    if ((rc = dalloc(ctx, &q->d_u, (size_t)n))) break;
    llfem_op_desc d = r.stiffness;
    d.op = LLFEM_OP_STIFFNESS;
    if ((rc = assemble_impl(m, &d, nullptr, nullptr, q->d_K))) break;
    k_load_from_weights<<<(unsigned)((m->N + 255) / 256), 256, 0, st>>>(m->N, wi->second.first, r.traction[0],
                                                                       r.traction[1], r.traction[2], q->d_f);
    SysSpec sp;
    sp.N = m->N;
    sp.nadj = m->nadj;
    sp.adj_ptr = m->d_adj_ptr;
    sp.adj = m->d_adj;
    sp.dslot = m->d_diag_slot;
    sp.val64 = q->d_K;
    sp.fixed = mi->second;
    sp.f = q->d_f;
    sp.ufix = nullptr;
    sp.reltol = r.reltol;
    sp.maxiter = r.maxiter;
    q->eng = (r.method == LLFEM_PCG_F32OP_RELIABLE) ? S->eng64.get() : S->eng32.get();
    q->support_key = sk;
    t_k = ms_since(t0);
    if (r.method == LLFEM_PCG_AMG_RELIABLE) {
      AmgEntry& ae = cm->amg[sk];
      if (!ae.h && !ae.job && !ae.failed) {
        double* d_Kc = nullptr;
        if ((rc = dalloc(ctx, &d_Kc, 9 * (size_t)m->nadj))) break;
        llfem_op_desc dc = r.stiffness;
        dc.op = LLFEM_OP_STIFFNESS;
        dc.D = nullptr;
        dc.E = 1.0;
        dc.nu = 0.3;
        dc.precision = LLFEM_FP64;
        int rb = assemble_impl(m, &dc, nullptr, nullptr, d_Kc);
        if (rb == LLFEM_OK) {
          SysSpec sc = sp;
          sc.val64 = d_Kc;
          rb = amg_job_start(ctx, sc, m->d_xyz, &ae.job);
        }
        (void)hipStreamSynchronize(st);
        dfree(ctx, d_Kc, sizeof(double) * 9 * (size_t)m->nadj);
        if (rb == LLFEM_ENOMEM) {
          rc = rb;
          break;
        }
        if (rb != LLFEM_OK) ae.failed = true;
      }
      sp.amg = ae.h;
      t_amg = ms_since(t0);
      if (ae.job) {
        q->spec = sp;
        q->waiting = true;
        q->t_admit = Clock::now();
        if (trace_on())
          std::fprintf(stderr, "[serve %9.1f] admit #%lld N=%lld mesh %.1f wm %.1f K %.1f amg %.1f -> waits\n",
                       trace_t(), (long long)q->id, (long long)m->N, t_mesh, t_wm - t_mesh, t_k - t_wm, t_amg - t_k);
        return LLFEM_OK;
      }
    }
    q->spec = sp;
    if ((rc = q->eng->add(sp, &q->sys))) break;
    if (trace_on())
      std::fprintf(stderr, "[serve %9.1f] admit #%lld N=%lld mesh %.1f wm %.1f K %.1f amg %.1f add %.1f ms\n",
                   trace_t(), (long long)q->id, (long long)m->N, t_mesh, t_wm - t_mesh, t_k - t_wm, t_amg > 0 ? t_amg - t_k : 0.0,
                   ms_since(t0) - (t_amg > 0 ? t_amg : t_k));
  } while (0);
  if (rc != LLFEM_OK) {
    (void)hipStreamSynchronize(st);
    free_request_buffers(ctx, q);
    q->sys = nullptr;
    release_mesh_use(q);
    q->cm = nullptr;
    return rc;
  }
  q->t_admit = Clock::now();
  return LLFEM_OK;
}

// A converged (or failed) system: write u to the caller, release the request.
void complete(llfem_server* S, Req* q) {
  llfem_ctx* ctx = S->ctx;
  int rc = LLFEM_OK;
  const int status = sys_status(q->sys);
  rc = q->eng->finish(q->sys, q->d_u);
  if (rc == LLFEM_OK) {
    const int64_t n = 3 * q->cm->m->N;
    if (hipMemcpyAsync(q->in.u_out, q->d_u, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(LLFEM_EHIP, "server: result copy failed");
  }
  q->iterations = sys_iterations(q->sys);
  q->updates = sys_updates(q->sys);
  q->rollbacks = sys_rollbacks(q->sys);
  q->restarts = sys_refreshes(q->sys);
  q->relres = sys_relres(q->sys);
  if (rc == LLFEM_OK && status == -2)
    rc = fail(LLFEM_ENOCONV, "cg: the FP32 correction was not finite in three successive updates");
  else if (rc == LLFEM_OK && status != 1)
    rc = fail(LLFEM_ENOCONV, "cg: maxiter reached (relative residual " + std::to_string(q->relres) + ")");
  if (rc != LLFEM_OK) q->err = llfem_last_error();
  q->eng->remove(q->sys);
  q->sys = nullptr;
  free_request_buffers(ctx, q);
  release_mesh_use(q);
  q->t_done = Clock::now();
  if (trace_on())
    std::fprintf(stderr, "[serve %9.1f] done #%lld N=%lld %lld it, %lld updates\n", trace_t(), (long long)q->id,
                 (long long)q->cm_nodes, (long long)q->iterations, (long long)q->updates);
  std::lock_guard<std::mutex> lk(S->mu);
  q->rc = rc;
  q->done = true;
  S->cv.notify_all();
}

void fail_request(llfem_server* S, Req* q, int rc) {
  std::lock_guard<std::mutex> lk(S->mu);
  q->rc = rc;
  q->err = llfem_last_error();
  q->t_done = Clock::now();
  q->done = true;
  S->cv.notify_all();
}

void worker_loop(llfem_server* S) {
  (void)hipSetDevice(S->ctx->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->active.empty(); });
      if (S->stop && S->pending.empty() && S->active.empty()) return;
    }
    // admission, first come first served while slots and memory allow
    for (;;) {
      std::shared_ptr<Req> q;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->pending.empty() || (int32_t)S->active.size() >= S->max_active) break;
        q = S->pending.front();
      }
      int rc = admit(S, q.get());
      if (rc == LLFEM_ENOMEM) {
        if (!S->active.empty()) break;             // retry once running requests release memory
        if (evict_unused(S)) rc = admit(S, q.get());
      }
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->pending.pop_front();
        if (rc == LLFEM_OK) S->active.push_back(q);
      }
      if (rc != LLFEM_OK) {
        fail_request(S, q.get(), rc);
        continue;
      }
      if (q->sys && sys_status(q->sys) != 0) {    // b = 0: converged at admission
        complete(S, q.get());
        std::lock_guard<std::mutex> lk(S->mu);
        S->active.erase(std::find(S->active.begin(), S->active.end(), q));
      }
    }
    // requests waiting for an AMG hierarchy whose host build has finished
    bool waiting = false;
    {
      std::vector<std::shared_ptr<Req>> act;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        act = S->active;
      }
      for (auto& q : act) {
        if (!q->waiting) continue;
        AmgEntry& ae = q->cm->amg[q->support_key];
        if (ae.job && amg_job_ready(ae.job)) {
          AmgJob* j = ae.job;
          ae.job = nullptr;
          if (amg_job_finish(S->ctx, j, q->spec, &ae.h) != LLFEM_OK) {
            ae.h = nullptr;
            ae.failed = true;
          }
        }
        if (ae.job) {
          waiting = true;
          continue;
        }
        q->spec.amg = ae.h;
        const Clock::time_point ta = Clock::now();
        const int rc = q->eng->add(q->spec, &q->sys);
        if (trace_on())
          std::fprintf(stderr, "[serve %9.1f] start #%lld after hierarchy: add %.1f ms\n", trace_t(), (long long)q->id,
                       ms_since(ta));
        if (rc == LLFEM_ENOMEM && (S->eng32->any_running() || S->eng64->any_running())) {
          waiting = true;          // retry once running systems release memory
          continue;
        }
        q->waiting = false;
        if (rc != LLFEM_OK) {
          q->sys = nullptr;
          free_request_buffers(S->ctx, q.get());
          release_mesh_use(q.get());
          {
            std::lock_guard<std::mutex> lk(S->mu);
            S->active.erase(std::find(S->active.begin(), S->active.end(), q));
          }
          fail_request(S, q.get(), rc);
        } else if (sys_status(q->sys) != 0) {
          complete(S, q.get());
          std::lock_guard<std::mutex> lk(S->mu);
          S->active.erase(std::find(S->active.begin(), S->active.end(), q));
        }
      }
    }
    if (S->active.empty()) continue;
    if (waiting && !S->eng32->any_running() && !S->eng64->any_running()) {
      std::this_thread::sleep_for(std::chrono::microseconds(200));
      continue;
    }
    // one chunk of iterations over every admitted system
    std::vector<SolveSys*> fin;
    int rc = LLFEM_OK;
    const Clock::time_point tc = Clock::now();
    for (BatchSolver* e : {S->eng32.get(), S->eng64.get()})
      if (rc == LLFEM_OK && e->any_running()) rc = e->run(S->chunk, &fin);
    if (trace_on())
      std::fprintf(stderr, "[serve %9.1f] chunk %.1f ms, %zu active, %zu finished\n", trace_t(), ms_since(tc),
                   S->active.size(),
                   fin.size());
    std::vector<std::shared_ptr<Req>> done;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (auto it = S->active.begin(); it != S->active.end();) {
        Req* q = it->get();
        const bool finished = q->sys && (std::find(fin.begin(), fin.end(), q->sys) != fin.end() || rc != LLFEM_OK);
        if (finished) {
          done.push_back(*it);
          it = S->active.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (auto& q : done) {
      if (rc != LLFEM_OK) {
        q->eng->remove(q->sys);
        q->sys = nullptr;
        free_request_buffers(S->ctx, q.get());
        release_mesh_use(q.get());
        fail_request(S, q.get(), rc);
      } else {
        complete(S, q.get());
      }
    }
  }
}

int check_request(const llfem_request* r) {
  if (!r) return fail(LLFEM_EINVAL, "llfem_server_submit: NULL request");
  if (!r->xyz || !r->conn || r->N <= 0 || r->E <= 0 || r->npe <= 0)
    return fail(LLFEM_EINVAL, "llfem_server_submit: bad mesh");
  if (!r->face_conn || r->F <= 0 || r->nf <= 0 || r->face_q <= 0 || !r->face_weights || !r->face_N_ref ||
      !r->face_dN_ref)
    return fail(LLFEM_EINVAL, "llfem_server_submit: bad face load");
  if (r->nfixed > 0 && !r->fixed) return fail(LLFEM_EINVAL, "llfem_server_submit: NULL fixed list");
  if (!r->u_out) return fail(LLFEM_EINVAL, "llfem_server_submit: NULL u_out");
  if (r->method != LLFEM_PCG_MIXED_RELIABLE && r->method != LLFEM_PCG_F32OP_RELIABLE &&
      r->method != LLFEM_PCG_AMG_RELIABLE)
    return fail(LLFEM_EUNSUP, "llfem_server_submit: method must be MIXED_RELIABLE, F32OP_RELIABLE or AMG_RELIABLE");
  if (!(r->reltol > 0) || r->maxiter <= 0) return fail(LLFEM_EINVAL, "llfem_server_submit: bad tolerance");
  if (!r->stiffness.weights || !r->stiffness.dN_ref || r->stiffness.q <= 0)
    return fail(LLFEM_EINVAL, "llfem_server_submit: bad quadrature tables");
  if (r->stiffness.element_scale)
    return fail(LLFEM_EUNSUP, "llfem_server_submit: element_scale is for llfem_assemble only");
  return LLFEM_OK;
}

}  // namespace

}  // namespace llfem

using namespace llfem;

extern "C" int llfem_server_create(int device, size_t hbm_budget, int32_t max_active, llfem_server** out) {
  if (!out) return fail(LLFEM_EINVAL, "llfem_server_create: NULL output");
  llfem_ctx* ctx = nullptr;
  LLFEM_TRY(llfem_ctx_create(device, hbm_budget, &ctx));
  llfem_server* S = new llfem_server();
  S->ctx = ctx;
  S->max_active = max_active > 0 ? max_active : 64;
  S->eng32.reset(new BatchSolver(ctx, false));
  S->eng64.reset(new BatchSolver(ctx, true));
  S->worker = std::thread(worker_loop, S);
  *out = S;
  return LLFEM_OK;
}

extern "C" int llfem_server_submit(llfem_server* S, const llfem_request* r, int64_t* id) {
  if (!S) return fail(LLFEM_EINVAL, "llfem_server_submit: NULL server");
  LLFEM_TRY(check_request(r));
  auto q = std::make_shared<Req>();
  q->in = *r;
  q->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(LLFEM_EINVAL, "llfem_server_submit: server is shutting down");
    q->id = S->next_id++;
    S->all[q->id] = q;
    S->pending.push_back(q);
  }
  S->cv.notify_all();
  if (id) *id = q->id;
  return LLFEM_OK;
}

extern "C" int llfem_server_wait(llfem_server* S, int64_t id, llfem_request_stats* stats) {
  if (!S) return fail(LLFEM_EINVAL, "llfem_server_wait: NULL server");
  std::shared_ptr<Req> q;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(LLFEM_EINVAL, "llfem_server_wait: unknown request id");
    q = it->second;
    S->cv.wait(lk, [&] { return q->done; });
    S->all.erase(id);
  }
  if (stats) {
    auto ms = [](Clock::time_point a, Clock::time_point b) {
      return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const bool admitted = q->t_admit.time_since_epoch().count() != 0;
    stats->latency_ms = ms(q->t_submit, q->t_done);
    stats->queue_ms = admitted ? ms(q->t_submit, q->t_admit) : stats->latency_ms;
    stats->service_ms = admitted ? ms(q->t_admit, q->t_done) : 0.0;
    stats->iterations = q->iterations;
    stats->updates = q->updates;
    stats->rollbacks = q->rollbacks;
    stats->restarts = q->restarts;
    stats->relres = q->relres;
    stats->mesh_cache_hit = q->cache_hit ? 1 : 0;
    stats->status = q->rc;
  }
  if (q->rc != LLFEM_OK) return fail(q->rc, q->err);
  return LLFEM_OK;
}

extern "C" int llfem_server_destroy(llfem_server* S) {
  if (!S) return LLFEM_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  S->eng32.reset();
  S->eng64.reset();
  for (auto& kv : S->meshes) free_cached(S->ctx, kv.second);
  S->meshes.clear();
  llfem_ctx_destroy(S->ctx);
  delete S;
  return LLFEM_OK;
}
