// Context, error reporting and device-memory accounting.
#include <mutex>

#include "common.h"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace llfem {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int dmalloc(llfem_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return LLFEM_OK;
  if (ctx->budget && ctx->in_use + bytes > ctx->budget)
    return fail(LLFEM_ENOMEM, "HBM budget exceeded: need " + std::to_string(bytes) +
                                  " more bytes, " + std::to_string(ctx->in_use) + " of " +
                                  std::to_string(ctx->budget) + " in use");
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(LLFEM_ENOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                  ") failed: " + hipGetErrorString(e));
  }
  ctx->in_use += bytes;
  return LLFEM_OK;
}

void dfree(llfem_ctx* ctx, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  ctx->in_use -= bytes;
}

// Large copies go through two pinned staging buffers: the DMA engine fills
// one while host threads copy the other out to the caller's pageable memory
// (HIP's own pageable path stages serially through one buffer).
int copy_to_host(llfem_ctx* ctx, double* dst, const double* src, size_t n) {
  constexpr size_t kStage = size_t(64) << 20;   // bytes per staging buffer
  const size_t bytes = n * sizeof(double);
  hipStream_t st = ctx->stream;
  if (bytes <= 2 * kStage) {
    LLFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipStreamSynchronize(st));
    return LLFEM_OK;
  }
  if (!ctx->h_stage[0]) {
    for (int b = 0; b < 2; ++b) {
      LLFEM_HIP(hipHostMalloc(&ctx->h_stage[b], kStage, hipHostMallocDefault));
      LLFEM_HIP(hipEventCreateWithFlags(&ctx->stage_ev[b], hipEventDisableTiming));
    }
    ctx->stage_bytes = kStage;
  }
  const size_t chunk = kStage / sizeof(double);
  const size_t nchunks = (n + chunk - 1) / chunk;
  auto issue = [&](size_t i) -> int {
    const size_t off = i * chunk, len = std::min(chunk, n - off);
    LLFEM_HIP(hipMemcpyAsync(ctx->h_stage[i & 1], src + off, len * sizeof(double), hipMemcpyDeviceToHost, st));
    LLFEM_HIP(hipEventRecord(ctx->stage_ev[i & 1], st));
    return LLFEM_OK;
  };
  LLFEM_TRY(issue(0));
  if (nchunks > 1) LLFEM_TRY(issue(1));
  const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  for (size_t i = 0; i < nchunks; ++i) {
    LLFEM_HIP(hipEventSynchronize(ctx->stage_ev[i & 1]));
    const size_t off = i * chunk, len = std::min(chunk, n - off);
    const double* from = static_cast<const double*>(ctx->h_stage[i & 1]);
    std::vector<std::thread> th;
    const size_t per = (len + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
      const size_t lo = t * per, hi = std::min(len, lo + per);
      if (lo < hi) th.emplace_back([=] { std::memcpy(dst + off + lo, from + lo, (hi - lo) * sizeof(double)); });
    }
    for (auto& x : th) x.join();
    if (i + 2 < nchunks) LLFEM_TRY(issue(i + 2));
  }
  return LLFEM_OK;
}

// Content hash for cache keys (not cryptographic): four independent 64-bit
// multiply-rotate lanes over 32-byte strides, byte-wise FNV-1a on the tail,
// then a final avalanche.  Several GB/s, so hashing a request's mesh arrays
// costs milliseconds.
uint64_t hash_bytes(const void* p, size_t n, uint64_t h) {
  constexpr uint64_t kA = 0x9E3779B97F4A7C15ull, kB = 0xBF58476D1CE4E5B9ull, kC = 0x94D049BB133111EBull;
  auto rotl = [](uint64_t x, int r) { return (x << r) | (x >> (64 - r)); };
  const unsigned char* c = static_cast<const unsigned char*>(p);
  uint64_t lane[4] = {h ^ kA, h + kB, rotl(h, 21) ^ kC, h - kA};
  size_t i = 0;
  for (; i + 32 <= n; i += 32)
    for (int k = 0; k < 4; ++k) {
      uint64_t w;
      std::memcpy(&w, c + i + 8 * k, 8);
      lane[k] = rotl(lane[k] + w * kB, 31) * kA;
    }
  uint64_t r = h ^ ((uint64_t)n * kC);
  for (int k = 0; k < 4; ++k) r = (r ^ rotl(lane[k], 17 + 8 * k)) * kB;
  for (; i < n; ++i) r = (r ^ c[i]) * 1099511628211ull;
  r ^= r >> 31;
  r *= kC;
  r ^= r >> 29;
  return r;
}

}  // namespace llfem

// ---- workspace of the allocation-free solve ------------------------------
// One array per GPU, part of this library's device image: the runtime places
// it when the image is loaded on a device and it stays until the process
// ends.  In host code `scratch` names a host-side stand-in, never the device
// memory: host code uses only the address hipGetSymbolAddress returns for
// the current device, kept per device in g_scratch.
static __device__ __align__(256) unsigned char scratch[llfem::kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct ScratchSlot {
  std::mutex call;          // one allocation-free solve at a time on the device
  void* ptr = nullptr;      // device address of `scratch` on this device
};
ScratchSlot g_scratch[kMaxDevices];
std::mutex g_scratch_setup;
}  // namespace

namespace llfem {

int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(LLFEM_EUNSUP, "device ordinal " + std::to_string(device) + " above the supported maximum");
  std::lock_guard<std::mutex> lk(g_scratch_setup);
  if (g_scratch[device].ptr) return LLFEM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(LLFEM_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch[device].ptr = p;
  return LLFEM_OK;
}

bool noalloc_scratch(int device, void** ptr, std::mutex** lock) {
  if (device < 0 || device >= kMaxDevices) return false;
  {
    std::lock_guard<std::mutex> lk(g_scratch_setup);
    if (!g_scratch[device].ptr) return false;
    *ptr = g_scratch[device].ptr;
  }
  *lock = &g_scratch[device].call;
  return true;
}

}  // namespace llfem

using namespace llfem;

extern "C" {

const char* llfem_last_error(void) { return g_error.c_str(); }

int llfem_ctx_create(int device, size_t hbm_budget_bytes, llfem_ctx** out) {
  if (!out) return fail(LLFEM_EINVAL, "llfem_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  LLFEM_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(LLFEM_EINVAL, "llfem_ctx_create: device " + std::to_string(device) +
                                  " out of range (" + std::to_string(ndev) + " devices)");
  LLFEM_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  LLFEM_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(LLFEM_EUNSUP, std::string("libllfem is built for gfx950; device is ") +
                                  prop.gcnArchName);
  // device address of the allocation-free solve's workspace, once per device
  LLFEM_TRY(llfem::noalloc_scratch_setup(device));
  llfem_ctx* c = new llfem_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  c->num_cus = prop.multiProcessorCount;
  // default-priority stream (never a priority stream)
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(LLFEM_EHIP, "hipStreamCreate failed");
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  // side streams for concurrent panel launches, created here so the first
  // assembly does not pay for them
  for (int k = 0; k < 2; ++k) {
    if (hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev[k], hipEventDisableTiming) != hipSuccess) {
      llfem_ctx_destroy(c);
      return fail(LLFEM_EHIP, "hipStreamCreate failed");
    }
  }
  if (hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess) {
    llfem_ctx_destroy(c);
    return fail(LLFEM_EHIP, "hipEventCreate failed");
  }
  *out = c;
  return LLFEM_OK;
}

int llfem_ctx_destroy(llfem_ctx* ctx) {
  if (!ctx) return LLFEM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& kv : ctx->tables) llfem::dfree(ctx, kv.second.ptr, kv.second.bytes);
  ctx->tables.clear();
  if (ctx->d_scratch) llfem::dfree(ctx, ctx->d_scratch, ctx->scratch_bytes);
  if (ctx->d_sink) llfem::dfree(ctx, ctx->d_sink, 64 * sizeof(double));
  for (int k = 0; k < 2; ++k) {
    if (ctx->side[k]) {
      (void)hipStreamSynchronize(ctx->side[k]);
      (void)hipStreamDestroy(ctx->side[k]);
    }
    if (ctx->join_ev[k]) (void)hipEventDestroy(ctx->join_ev[k]);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  for (int b = 0; b < 2; ++b) {
    if (ctx->h_stage[b]) (void)hipHostFree(ctx->h_stage[b]);
    if (ctx->stage_ev[b]) (void)hipEventDestroy(ctx->stage_ev[b]);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LLFEM_OK;
}

int llfem_ctx_synchronize(llfem_ctx* ctx) {
  if (!ctx) return fail(LLFEM_EINVAL, "llfem_ctx_synchronize: NULL context");
  LLFEM_HIP(hipStreamSynchronize(ctx->stream));
  return LLFEM_OK;
}

size_t llfem_ctx_bytes_in_use(const llfem_ctx* ctx) { return ctx ? ctx->in_use : 0; }

}  // extern "C"
