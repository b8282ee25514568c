// Traction load vector on a set of 2D face elements.
//
// Reference: loadVector(problem, [load(name; fx, fy, fz)]) (src/linear.jl:3088-3434):
// per face element and Gauss point, Ja = |t1 x t2| from the face Jacobian
// columns (src/linear.jl:3378-3384), f1 += H1' * t * Ja * w, then
// fp[nn2] += f1 in element order (src/linear.jl:3408-3413).
//
// Device version: one thread per face element computes the scalar nodal
// integrals s_a = sum_q N_a Ja w (the traction is uniform over the face);
// a second pass gathers them per node in ascending face order (face-node
// incidences prepared on the host), so the summation order matches the
// reference and no atomics are used.
#include <algorithm>
#include <numeric>

#include "common.h"

namespace {

__global__ void k_face_integrals(int64_t F, int nf, int q, int64_t N, const double* __restrict__ xyz,
                                 const int32_t* __restrict__ fconn, const double* __restrict__ w,
                                 const double* __restrict__ Nref, const double* __restrict__ dN,
                                 double* __restrict__ s) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= F) return;
  for (int a = 0; a < nf; ++a) s[e * nf + a] = 0.0;
  for (int p = 0; p < q; ++p) {
    double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
    for (int a = 0; a < nf; ++a) {
      const int32_t v = fconn[e * nf + a];
      const double x[3] = {xyz[v], xyz[N + v], xyz[2 * N + v]};
      const double g0 = dN[((size_t)p * nf + a) * 3], g1 = dN[((size_t)p * nf + a) * 3 + 1];
      for (int i = 0; i < 3; ++i) {
        t1[i] += x[i] * g0;
        t2[i] += x[i] * g1;
      }
    }
    const double xy = t1[0] * t2[1] - t1[1] * t2[0];
    const double yz = t1[1] * t2[2] - t1[2] * t2[1];
    const double zx = t1[2] * t2[0] - t1[0] * t2[2];
    const double Ja = sqrt(xy * xy + yz * yz + zx * zx);
    for (int a = 0; a < nf; ++a) s[e * nf + a] += Nref[(size_t)p * nf + a] * Ja * w[p];
  }
}

__global__ void k_face_gather(int64_t nnodes, const int32_t* __restrict__ nodes,
                              const int32_t* __restrict__ ptr, const int32_t* __restrict__ slots,
                              const double* __restrict__ s, double t0, double t1, double t2,
                              double* __restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  double acc = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) acc += s[slots[k]];
  const int64_t v = nodes[i];
  f[3 * v] = t0 * acc;
  f[3 * v + 1] = t1 * acc;
  f[3 * v + 2] = t2 * acc;
}

// Physical coordinates of the face Gauss points: thread per (face, point).
__global__ void k_face_points(int64_t F, int nf, int q, int64_t N, const double* __restrict__ xyz,
                              const int32_t* __restrict__ fconn, const double* __restrict__ Nref,
                              double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F * q) return;
  const int64_t e = t / q;
  const int p = (int)(t - e * q);
  double x[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < nf; ++a) {
    const int32_t v = fconn[e * nf + a];
    const double h = Nref[(size_t)p * nf + a];
    x[0] += h * xyz[v];
    x[1] += h * xyz[N + v];
    x[2] += h * xyz[2 * N + v];
  }
  for (int i = 0; i < 3; ++i) out[t * 3 + i] = x[i];
}

// Per face element and node a: s3[a][c] = sum_q N_a t_c(q) Ja w_q, the
// traction given per Gauss point (t [F][q][3]).
__global__ void k_face_integrals_gp(int64_t F, int nf, int q, int64_t N, const double* __restrict__ xyz,
                                    const int32_t* __restrict__ fconn, const double* __restrict__ w,
                                    const double* __restrict__ Nref, const double* __restrict__ dN,
                                    const double* __restrict__ tgp, double* __restrict__ s3) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= F) return;
  for (int a = 0; a < 3 * nf; ++a) s3[e * nf * 3 + a] = 0.0;
  for (int p = 0; p < q; ++p) {
    double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
    for (int a = 0; a < nf; ++a) {
      const int32_t v = fconn[e * nf + a];
      const double x[3] = {xyz[v], xyz[N + v], xyz[2 * N + v]};
      const double g0 = dN[((size_t)p * nf + a) * 3], g1 = dN[((size_t)p * nf + a) * 3 + 1];
      for (int i = 0; i < 3; ++i) {
        t1[i] += x[i] * g0;
        t2[i] += x[i] * g1;
      }
    }
    const double xy = t1[0] * t2[1] - t1[1] * t2[0];
    const double yz = t1[1] * t2[2] - t1[2] * t2[1];
    const double zx = t1[2] * t2[0] - t1[0] * t2[2];
    const double Ja = sqrt(xy * xy + yz * yz + zx * zx);
    const double* tq = tgp + (e * q + p) * 3;
    for (int a = 0; a < nf; ++a) {
      const double c = Nref[(size_t)p * nf + a] * Ja * w[p];
      for (int i = 0; i < 3; ++i) s3[(e * nf + a) * 3 + i] += c * tq[i];
    }
  }
}

__global__ void k_face_gather3(int64_t nnodes, const int32_t* __restrict__ nodes,
                               const int32_t* __restrict__ ptr, const int32_t* __restrict__ slots,
                               const double* __restrict__ s3, double* __restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = ptr[i]; k < ptr[i + 1]; ++k)
    for (int c = 0; c < 3; ++c) acc[c] += s3[(size_t)slots[k] * 3 + c];
  const int64_t v = nodes[i];
  for (int c = 0; c < 3; ++c) f[3 * v + c] = acc[c];
}

}  // namespace

using namespace llfem;

namespace {

// Host side of a face set: 0-based connectivity and the face-node incidences
// in ascending face order (nodes, ptr into slots).
struct FacePrep {
  std::vector<int32_t> fc, slots, nodes, ptr;
};

int prepare_faces(const llfem_mesh* m, int32_t face_type, int32_t nf, const int64_t* face_conn, int64_t F,
                  int32_t one_based, int32_t q, FacePrep* P) {
  if (etype_dim(face_type) != 2)
    return fail(LLFEM_EUNSUP, "loadVector: face element type " + std::to_string(face_type) +
                                  " is not a 2D element");
  if (nf <= 0 || q <= 0 || F < 0) return fail(LLFEM_EINVAL, "loadVector: bad face-set sizes");
  const int64_t base = one_based ? 1 : 0;
  P->fc.resize((size_t)F * nf);
  for (int64_t i = 0; i < F * nf; ++i) {
    const int64_t v = face_conn[i] - base;
    if (v < 0 || v >= m->N) return fail(LLFEM_EINVAL, "loadVector: face node out of range");
    P->fc[i] = (int32_t)v;
  }
  P->slots.resize((size_t)F * nf);
  std::iota(P->slots.begin(), P->slots.end(), 0);
  std::stable_sort(P->slots.begin(), P->slots.end(), [&](int32_t x, int32_t y) { return P->fc[x] < P->fc[y]; });
  P->nodes.clear();
  P->ptr.assign(1, 0);
  for (size_t k = 0; k < P->slots.size(); ++k)
    if (k == 0 || P->fc[P->slots[k]] != P->fc[P->slots[k - 1]]) {
      if (k) P->ptr.push_back((int32_t)k);
      P->nodes.push_back(P->fc[P->slots[k]]);
    }
  P->ptr.push_back((int32_t)P->slots.size());
  return LLFEM_OK;
}

}  // namespace

extern "C" int llfem_load_traction(llfem_mesh* m, int32_t face_type, int32_t nf,
                                   const int64_t* face_conn, int64_t F, int32_t one_based, int32_t q,
                                   const double* weights, const double* N_ref, const double* dN_ref,
                                   const double traction[3], double* f_out) {
  if (!m || !face_conn || !weights || !N_ref || !dN_ref || !traction || !f_out)
    return fail(LLFEM_EINVAL, "llfem_load_traction: NULL argument");
  if (etype_dim(face_type) != 2)
    return fail(LLFEM_EUNSUP, "loadVector: face element type " + std::to_string(face_type) +
                                  " is not a 2D element");
  if (nf <= 0 || q <= 0 || F < 0) return fail(LLFEM_EINVAL, "llfem_load_traction: bad sizes");
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N = m->N;
  const int64_t base = one_based ? 1 : 0;
  // face-node incidences in ascending face order (host; face sets are 2D)
  std::vector<int32_t> fc((size_t)F * nf);
  for (int64_t i = 0; i < F * nf; ++i) {
    int64_t v = face_conn[i] - base;
    if (v < 0 || v >= N) return fail(LLFEM_EINVAL, "loadVector: face node out of range");
    fc[i] = (int32_t)v;
  }
  std::vector<int32_t> slot_order((size_t)F * nf);
  std::iota(slot_order.begin(), slot_order.end(), 0);
  std::stable_sort(slot_order.begin(), slot_order.end(),
                   [&](int32_t x, int32_t y) { return fc[x] < fc[y]; });
  std::vector<int32_t> nodes, ptr{0};
  for (size_t k = 0; k < slot_order.size(); ++k) {
    if (k == 0 || fc[slot_order[k]] != fc[slot_order[k - 1]]) {
      if (k) ptr.push_back((int32_t)k);
      nodes.push_back(fc[slot_order[k]]);
    }
  }
  ptr.push_back((int32_t)slot_order.size());
  const int64_t nn = (int64_t)nodes.size();

  int32_t *d_fc = nullptr, *d_nodes = nullptr, *d_ptr = nullptr, *d_slots = nullptr;
  double *d_s = nullptr, *d_tab = nullptr, *d_f = nullptr;
  const size_t tab_n = (size_t)q + (size_t)q * nf + (size_t)q * nf * 3;
  int s = LLFEM_OK;
  do {
    if ((s = dalloc(ctx, &d_fc, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_nodes, (size_t)nn))) break;
    if ((s = dalloc(ctx, &d_ptr, (size_t)nn + 1))) break;
    if ((s = dalloc(ctx, &d_slots, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_s, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_tab, tab_n))) break;
    if ((s = dalloc(ctx, &d_f, 3 * (size_t)N))) break;
    std::vector<double> tab(tab_n);
    std::copy(weights, weights + q, tab.begin());
    std::copy(N_ref, N_ref + (size_t)q * nf, tab.begin() + q);
    std::copy(dN_ref, dN_ref + (size_t)q * nf * 3, tab.begin() + q + (size_t)q * nf);
    hipError_t e = hipSuccess;
    e = e ? e : hipMemcpyAsync(d_fc, fc.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_nodes, nodes.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_ptr, ptr.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_slots, slot_order.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab_n, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemsetAsync(d_f, 0, sizeof(double) * 3 * N, st);
    if (e != hipSuccess) { s = fail(LLFEM_EHIP, std::string("loadVector upload: ") + hipGetErrorString(e)); break; }
    if (F > 0) {
      k_face_integrals<<<(unsigned)((F + 127) / 128), 128, 0, st>>>(
          F, nf, q, N, m->d_xyz, d_fc, d_tab, d_tab + q, d_tab + q + (size_t)q * nf, d_s);
      k_face_gather<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(nn, d_nodes, d_ptr, d_slots, d_s,
                                                                  traction[0], traction[1],
                                                                  traction[2], d_f);
    }
    e = hipMemcpyAsync(f_out, d_f, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) s = fail(LLFEM_EHIP, std::string("loadVector: ") + hipGetErrorString(e));
  } while (0);
  drelease(ctx, d_fc, (size_t)F * nf);
  drelease(ctx, d_nodes, (size_t)nn);
  drelease(ctx, d_ptr, (size_t)nn + 1);
  drelease(ctx, d_slots, (size_t)F * nf);
  drelease(ctx, d_s, (size_t)F * nf);
  drelease(ctx, d_tab, tab_n);
  drelease(ctx, d_f, 3 * (size_t)N);
  return s;
}

extern "C" int llfem_face_points(llfem_mesh* m, int32_t face_type, int32_t nf, const int64_t* face_conn, int64_t F,
                                 int32_t one_based, int32_t q, const double* N_ref, double* xyz_out) {
  if (!m || !face_conn || !N_ref || !xyz_out) return fail(LLFEM_EINVAL, "llfem_face_points: NULL argument");
  FacePrep P;
  LLFEM_TRY(prepare_faces(m, face_type, nf, face_conn, F, one_based, q, &P));
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int32_t* d_fc = nullptr;
  double *d_N = nullptr, *d_out = nullptr;
  const size_t npts = (size_t)F * q;
  int s = LLFEM_OK;
  do {
    if ((s = dalloc(ctx, &d_fc, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_N, (size_t)q * nf))) break;
    if ((s = dalloc(ctx, &d_out, 3 * npts))) break;
    hipError_t e = hipMemcpyAsync(d_fc, P.fc.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_N, N_ref, sizeof(double) * q * nf, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && npts > 0)
      k_face_points<<<(unsigned)((npts + 255) / 256), 256, 0, st>>>(F, nf, q, m->N, m->d_xyz, d_fc, d_N, d_out);
    e = e ? e : hipGetLastError();
    e = e ? e : hipMemcpyAsync(xyz_out, d_out, sizeof(double) * 3 * npts, hipMemcpyDeviceToHost, st);
    e = e ? e : hipStreamSynchronize(st);
    if (e != hipSuccess) s = fail(LLFEM_EHIP, std::string("llfem_face_points: ") + hipGetErrorString(e));
  } while (0);
  drelease(ctx, d_fc, (size_t)F * nf);
  drelease(ctx, d_N, (size_t)q * nf);
  drelease(ctx, d_out, 3 * npts);
  return s;
}

extern "C" int llfem_load_traction_gp(llfem_mesh* m, int32_t face_type, int32_t nf, const int64_t* face_conn,
                                      int64_t F, int32_t one_based, int32_t q, const double* weights,
                                      const double* N_ref, const double* dN_ref, const double* t_gp,
                                      double* f_out) {
  if (!m || !face_conn || !weights || !N_ref || !dN_ref || !t_gp || !f_out)
    return fail(LLFEM_EINVAL, "llfem_load_traction_gp: NULL argument");
  FacePrep P;
  LLFEM_TRY(prepare_faces(m, face_type, nf, face_conn, F, one_based, q, &P));
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N = m->N, nn = (int64_t)P.nodes.size();
  const size_t tab_n = (size_t)q + (size_t)q * nf + (size_t)q * nf * 3;
  int32_t *d_fc = nullptr, *d_nodes = nullptr, *d_ptr = nullptr, *d_slots = nullptr;
  double *d_s3 = nullptr, *d_tab = nullptr, *d_t = nullptr, *d_f = nullptr;
  int s = LLFEM_OK;
  do {
    if ((s = dalloc(ctx, &d_fc, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_nodes, (size_t)nn))) break;
    if ((s = dalloc(ctx, &d_ptr, (size_t)nn + 1))) break;
    if ((s = dalloc(ctx, &d_slots, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_s3, 3 * (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_tab, tab_n))) break;
    if ((s = dalloc(ctx, &d_t, 3 * (size_t)F * q))) break;
    if ((s = dalloc(ctx, &d_f, 3 * (size_t)N))) break;
    std::vector<double> tab(tab_n);
    std::copy(weights, weights + q, tab.begin());
    std::copy(N_ref, N_ref + (size_t)q * nf, tab.begin() + q);
    std::copy(dN_ref, dN_ref + (size_t)q * nf * 3, tab.begin() + q + (size_t)q * nf);
    hipError_t e = hipSuccess;
    e = e ? e : hipMemcpyAsync(d_fc, P.fc.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_nodes, P.nodes.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_ptr, P.ptr.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_slots, P.slots.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab_n, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_t, t_gp, sizeof(double) * 3 * F * q, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemsetAsync(d_f, 0, sizeof(double) * 3 * N, st);
    if (e != hipSuccess) { s = fail(LLFEM_EHIP, std::string("loadVector upload: ") + hipGetErrorString(e)); break; }
    if (F > 0) {
      k_face_integrals_gp<<<(unsigned)((F + 127) / 128), 128, 0, st>>>(
          F, nf, q, N, m->d_xyz, d_fc, d_tab, d_tab + q, d_tab + q + (size_t)q * nf, d_t, d_s3);
      k_face_gather3<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(nn, d_nodes, d_ptr, d_slots, d_s3, d_f);
    }
    e = hipGetLastError();
    e = e ? e : hipMemcpyAsync(f_out, d_f, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, st);
    e = e ? e : hipStreamSynchronize(st);
    if (e != hipSuccess) s = fail(LLFEM_EHIP, std::string("loadVector: ") + hipGetErrorString(e));
  } while (0);
  drelease(ctx, d_fc, (size_t)F * nf);
  drelease(ctx, d_nodes, (size_t)nn);
  drelease(ctx, d_ptr, (size_t)nn + 1);
  drelease(ctx, d_slots, (size_t)F * nf);
  drelease(ctx, d_s3, 3 * (size_t)F * nf);
  drelease(ctx, d_tab, tab_n);
  drelease(ctx, d_t, 3 * (size_t)F * q);
  drelease(ctx, d_f, 3 * (size_t)N);
  return s;
}
