// Internal definitions shared by the libllfem translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "llfem.h"

namespace llfem {

void set_error(const std::string& msg);

#define LLFEM_HIP(expr)                                                            \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      ::llfem::set_error(std::string("HIP error: ") + hipGetErrorString(_e) +      \
                         " at " + __FILE__ + ":" + std::to_string(__LINE__) +     \
                         " (" #expr ")");                                          \
      return LLFEM_EHIP;                                                           \
    }                                                                              \
  } while (0)

#define LLFEM_TRY(expr)          \
  do {                           \
    int _s = (expr);             \
    if (_s != LLFEM_OK) return _s; \
  } while (0)

// Pointers read from device-memory tables are generic to the compiler, which
// then emits flat loads and stores (counted against both vmcnt and lgkmcnt,
// so every wait becomes a full drain).  Every such pointer here addresses
// device memory: casting it to the global address space restores global_*
// instructions.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gp(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

// One panel-kernel launch.  Nodes of similar incidence count are grouped
// (padded with null steps to `cnt`), laid out in processing order as
// per-step records so a wavefront streams `rounds` nodes per lane group.
struct PanelLaunch {
  int64_t rec_base;     // first record of the launch
  int64_t desc_base;    // first node descriptor of the launch
  int64_t waves;        // wavefronts in the launch
  int32_t cnt;          // steps (incidences, padded) per node
  int32_t rounds;       // nodes per lane group (per wavefront stream)
  int32_t max_r;        // LDS panel rows per lane group
  int32_t nodes;        // real nodes in the launch
};

// Per-rule geometry state: affine detection is a property of the mesh and
// the rule's points, computed once.
struct RuleState {
  bool checked = false;
  bool affine = false;
  bool singular = false;
};

// Tuned panel parameters of one element type (llfem_autotune): launch width
// multiplier of the record planner (0 = kPanelMult), waves per block (0 = the
// most resident waves), streams the node-class launches are spread over (0 =
// default two) and whether the step-unrolled round kernels run where they
// exist (-1 = default yes, 0 = no, 1 = yes).
constexpr int kPanelMult = 4;   // default launch-width multiplier of the record planner

struct Tune {
  int32_t mult = 0;
  int32_t wpb = 0;
  double ms = 0.0;
  int32_t streams = 0;
  int32_t rounds = -1;
};

struct AmgHier;

// A device copy of a reference table, shared by every mesh of a context.
struct DevTable {
  double* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace llfem

struct llfem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // panel launches of one assembly may run on the side streams too (default
  // priority), forked from and joined to `stream`
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[2] = {nullptr, nullptr};
  int num_cus = 0;
  std::map<int, llfem::Tune> tune;                 // by gmsh element type
  std::map<uint64_t, llfem::DevTable> tables;      // by content hash
  void* d_scratch = nullptr;                       // batch descriptors + queue counter
  size_t scratch_bytes = 0;
  double* d_sink = nullptr;                        // 64 doubles: target of padding-lane panel stores
  // pinned staging for large device-to-host copies (copy_to_host)
  void* h_stage[2] = {nullptr, nullptr};
  size_t stage_bytes = 0;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

struct llfem_mesh {
  llfem_ctx* ctx = nullptr;
  int64_t N = 0, E = 0;
  int32_t npe = 0, etype = 0, pdim = 3;
  // coordinates SoA x[N] y[N] z[N] and node-major (x, y, z, 0); connectivity SoA [npe][E]
  double* d_xyz = nullptr;
  double4* d_xyz4 = nullptr;
  int32_t* d_conn = nullptr;
  // node adjacency (sorted) = the structural pattern at node level
  int64_t nadj = 0;
  int64_t* d_adj_ptr = nullptr;   // [N+1]
  int32_t* d_adj = nullptr;       // [nadj]
  int32_t* d_diag_slot = nullptr; // [N] position of j in its own row list
  // incidences sorted by (node, element)
  int64_t ninc = 0;
  int32_t* d_inc_ptr = nullptr;   // [N+1]
  int32_t* d_inc_elem = nullptr;  // [ninc]
  uint8_t* d_inc_loc = nullptr;   // [ninc]
  void* d_inc_map = nullptr;      // [ninc][npe] slot of each element node in the panel
  int32_t map_bytes = 1;
  // processing-order step records + node descriptors (see PanelLaunch)
  void* d_rec = nullptr;
  int64_t rec_count = 0;
  int32_t rec_bytes = 16;
  int2* d_desc = nullptr;         // {adj offset, R} per node slot
  int64_t desc_count = 0;
  std::vector<llfem::PanelLaunch> launches;
  int32_t max_r = 0;
  // host copies kept for pattern export
  std::vector<int64_t> h_adj_ptr;
  // per-assembly scratch
  double* d_geo = nullptr;
  size_t geo_bytes = 0;
  double* d_tab = nullptr;
  size_t tab_bytes = 0;
  std::map<uint64_t, llfem::RuleState> rules;
  // device-resident results
  double* d_valK = nullptr;
  double* d_valM = nullptr;
  double* d_lumped = nullptr;
  bool have_K = false;
  bool have_M = false;
  bool have_lumped = false;
  // resident operators assembled per material group (element_scale): which
  // elements have contributed since the last non-accumulating call [op]
  bool partial[2] = {false, false};
  std::vector<uint8_t> covered[2];
  // solver state cached per assembled K
  float* d_valK32 = nullptr;
  double* d_diagK = nullptr;
  bool solver_ready = false;
  // high-order (Gram) assembly tables, built once per mesh.  Positions are
  // the Gram row order: local node d_ho_perm[pos] (element 0's nodes in
  // ascending global id).  An element block (column pos, row pos) that is the
  // only contribution to its CSC block is written straight into the column's
  // panel; the others are staged in the order of their CSC block and, within
  // it, ascending element, and summed by the gather.
  std::vector<int32_t> h_conn0;   // nodes of element 0 (0-based)
  int32_t* d_ho_perm = nullptr;   // [npe] local node of each row position
  int32_t* d_gslot = nullptr;     // [E*npe*npe] (e, col pos, row pos): >= 0: offset 3 s of the block in its
                                  // column panel (slot s); < 0: -(k+1) for staged block k
  int64_t* d_gcol = nullptr;      // [E*npe][2] (e, col pos): 9 a0 and 3 R of the column node's panel
  int64_t nstaged = 0;            // staged element blocks
  int64_t nmulti = 0;             // CSC blocks with more than one contribution
  int32_t* d_mptr = nullptr;      // [nmulti+1] their staged blocks d_mptr[m] .. d_mptr[m+1]
  int64_t* d_mref = nullptr;      // [nmulti][2] value offset 9 a0 + 3 s and panel stride 3 R of each
  void* d_ho = nullptr;           // staged blocks, general material, gradient table by position
  size_t ho_bytes = 0;
  // AMG hierarchies kept per support (key: hash of the constrained-DoF list),
  // each with the device mask it was built for
  std::map<uint64_t, std::pair<llfem::AmgHier*, uint8_t*>> amg;
};

namespace llfem {

int dmalloc(llfem_ctx* ctx, void** p, size_t bytes);
// n doubles from device memory to pageable host memory, after the work
// already queued on the context's stream; returns when dst holds them.
int copy_to_host(llfem_ctx* ctx, double* dst, const double* src, size_t n);
void dfree(llfem_ctx* ctx, void* p, size_t bytes);

template <typename T>
int dalloc(llfem_ctx* ctx, T** p, size_t count) {
  return dmalloc(ctx, reinterpret_cast<void**>(p), count * sizeof(T));
}
template <typename T>
void drelease(llfem_ctx* ctx, T*& p, size_t count) {
  if (p) dfree(ctx, p, count * sizeof(T));
  p = nullptr;
}

int build_pattern(llfem_mesh* m);
int rebuild_records(llfem_mesh* m);

// Element-type facts the kernels need.
inline int etype_dim(int et) {
  switch (et) {
    case 4: case 5: case 6: case 11: case 12: case 92: case 93: return 3;
    case 2: case 3: case 9: case 10: case 36: case 37: return 2;
    case 1: case 8: case 26: case 27: return 1;
    case 15: return 0;
    default: return -1;
  }
}

int assemble_impl(llfem_mesh* m, const llfem_op_desc* d, double* nzval_out, llfem_stats* stats,
                  double* dev_out);
std::vector<double> operator_table(const llfem_op_desc* d, int npe, bool affine);
int voigt(int i, int j);

int assemble_highorder(llfem_mesh* m, const llfem_op_desc* d, bool affine, const double* d_dN,
                       const double* d_w, double lam, double mu, const double* C81, double* out);

uint64_t hash_bytes(const void* p, size_t n, uint64_t h = 1469598103934665603ull);

// ---- batched mixed-precision PCG engine (engine.hip) --------------------------
// One linear system in node-panel form; all pointers are device memory owned
// by the caller and must stay valid while the system is in the engine.
struct SysSpec {
  int64_t N = 0;                    // nodes (3 DoFs each)
  int64_t nadj = 0;                 // node blocks of K
  const int64_t* adj_ptr = nullptr;
  const int32_t* adj = nullptr;
  const int32_t* dslot = nullptr;   // diagonal slot of each node
  const double* val64 = nullptr;    // FP64 panels of K
  const uint8_t* fixed = nullptr;   // [3N] constrained-DoF mask
  const double* f = nullptr;        // [3N] load
  const double* ufix = nullptr;     // [3N] prescribed values (NULL = 0)
  double reltol = 1e-10;
  int64_t maxiter = 100000;
  const struct AmgHier* amg = nullptr;   // AMG preconditioner (FP32 vectors only); NULL = Jacobi
};

struct SolveSys;

class BatchSolver {
 public:
  static constexpr double kDelta = 0.1;      // reliable-update trigger (residual reduction)
  static constexpr double kDeltaAmg = 0.1;   // the same for AMG systems
  static constexpr int kMaxRunAmg = 40;      // AMG systems: FP64 update after at most this many iterations
  static constexpr double kDriftAmg = 0.25;  // AMG update: relative drift beyond which z is recomputed
  // f64_vectors: FP64 Krylov vectors over the FP32 operator (else FP32 vectors)
  BatchSolver(llfem_ctx* c, bool f64_vectors);
  ~BatchSolver();
  static size_t bytes_needed(int64_t N, int64_t nadj, bool f64_vectors);
  int add(const SysSpec& spec, SolveSys** out);               // set up and start a system
  int run(int iters, std::vector<SolveSys*>* finished);       // iterations on all running systems
  int finish(SolveSys* s, double* u);                          // write u (device), true residual
  void remove(SolveSys* s);
  bool any_running() const;

 private:
  int update_many(const std::vector<SolveSys*>& ss, bool init);
  int host_slots(size_t n);
  int rebuild_tables();
  void release(SolveSys* s);
  int launch_chunk(int iters);       // the iterations of one chunk, as kernel launches
  void drop_graph();
  llfem_ctx* ctx_;
  // the chunk captured as a HIP graph, valid while the live set is unchanged
  hipGraphExec_t chunk_exec_ = nullptr;
  int chunk_iters_ = 0;
  bool f64v_;
  std::vector<SolveSys*> sys_, live_;
  bool dirty_ = true;
  void* tables_ = nullptr;
  size_t tables_bytes_ = 0;
  void* d_sys_ = nullptr;
  struct AmgBatch* amgb_ = nullptr;
  int32_t* d_sb_ = nullptr;
  int32_t* d_vb_ = nullptr;
  double* d_states_ = nullptr;
  double* hslot_ = nullptr;          // pinned readback slots of the updates
  size_t hslot_n_ = 0;
  int nsys_ = 0;
  int64_t nsb_ = 0, nvb_ = 0;
};

// ---- aggregation AMG preconditioner (amg.hip) ------------------------------------
struct AmgInput;   // pattern, strong couplings, coordinates, support (host copies)
struct AmgHost;    // host-built hierarchy
struct AmgHier;    // device hierarchy of one mesh + support
struct AmgSys;     // numeric levels of one system
struct AmgBatch;   // V-cycle launch tables over several systems
int amg_gather(llfem_ctx* ctx, const SysSpec& sp, const double* d_xyz, AmgInput* in);
int amg_build_host(const AmgInput& in, AmgHost* h);
int amg_upload(llfem_ctx* ctx, const AmgHost& H, const SysSpec& sp, AmgHier** out);
void amg_free(llfem_ctx* ctx, AmgHier* h);
int64_t amg_levels(const AmgHier* h);
size_t amg_bytes(const AmgHier* h);
double amg_host_ms(const AmgHier* h);
int amg_sys_create(llfem_ctx* ctx, const AmgHier* h, const float* val32, float* r, float* z, float* q,
                   double* rzpart, AmgSys** out);
void amg_sys_free(llfem_ctx* ctx, AmgSys* s);
int amg_apply_solo(llfem_ctx* ctx, AmgSys* s);
int amg_batch_build(llfem_ctx* ctx, const std::vector<AmgSys*>& sys, const std::vector<const double*>& states,
                    AmgBatch** b);
int amg_batch_apply(llfem_ctx* ctx, AmgBatch* b);
void amg_batch_free(llfem_ctx* ctx, AmgBatch* b);
// One-call hierarchy build for a system (gather + host build + upload).
int amg_build(llfem_ctx* ctx, const SysSpec& sp, const double* d_xyz, AmgHier** out);
// The same with the host build on a thread of its own.
struct AmgJob;
int amg_job_start(llfem_ctx* ctx, const SysSpec& sp, const double* d_xyz, AmgJob** job);
bool amg_job_ready(AmgJob* job);
int amg_job_finish(llfem_ctx* ctx, AmgJob* job, const SysSpec& sp, AmgHier** out);   // frees the job
void amg_job_drop(AmgJob* job);                                                      // waits, frees

int64_t sys_iterations(const SolveSys* s);
int64_t sys_updates(const SolveSys* s);
int64_t sys_rollbacks(const SolveSys* s);   // updates that kept the previous solution
int64_t sys_refreshes(const SolveSys* s);   // AMG updates that recomputed z
int sys_status(const SolveSys* s);      // 0 running, 1 converged, -1 maxiter, -2 non-finite FP32 correction
double sys_relres(const SolveSys* s);

// Size of the per-device workspace array of llfem_solve_csc_noalloc, and its
// set-up: looks up the array's device address on `device` (current), once
// per device; called by llfem_ctx_create.  noalloc_scratch gives that
// address and the lock that serialises its users on the device (false
// before set-up).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
int noalloc_scratch_setup(int device);
bool noalloc_scratch(int device, void** ptr, std::mutex** lock);

}  // namespace llfem
