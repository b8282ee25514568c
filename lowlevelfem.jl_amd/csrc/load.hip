// Load vectors: traction on a set of boundary elements (surface, curve or
// point elements), pressure on a set of 2D face elements, body force on the
// mesh's volume elements.
//
// Reference: loadVector(problem, [load(name; fx, fy, fz)]) (src/linear.jl:3088-3434):
// per boundary element and Gauss point, Ja = |t1 x t2| on a surface,
// |t1| on a curve and 1 at a point (src/linear.jl:3378-3388), from the
// element's Jacobian columns, f1 += H1' * t * Ja * w, then fp[nn2] += f1 in
// element order (src/linear.jl:3408-3413).
//
// Device version: one thread per face element computes the scalar nodal
// integrals s_a = sum_q N_a Ja w (the traction is uniform over the face);
// a second pass gathers them per node in ascending face order (face-node
// incidences prepared on the host), so the summation order matches the
// reference and no atomics are used.
//
// Body force (loadVector on a volume group, src/linear.jl:3054-3068, and
// ∫(Pu ⋅ g; Ω), src/multifield.jl:3496-3748): one thread per element
// integrates N_a g det J w, and one thread per node sums its elements'
// parts over the mesh's (node, element)-sorted incidence lists.  Pressure
// (src/linear.jl:3196-3224): per-face nodal unit normals, their per-node
// average times -p (normalVector + elementsToNodes), then the traction
// interpolated from the nodes and integrated like a traction.
#include <algorithm>
#include <numeric>

#include "common.h"

namespace {

// Measure of a boundary element of dimension fd at one point, from its
// Jacobian columns t1 = dx/dxi, t2 = dx/deta (src/linear.jl:3378-3388 for a
// 3D solid): |t1 x t2| on a surface, |t1| on a curve, 1 at a point.
__device__ inline double face_measure(int fd, const double t1[3], const double t2[3]) {
  if (fd == 2) {
    const double xy = t1[0] * t2[1] - t1[1] * t2[0];
    const double yz = t1[1] * t2[2] - t1[2] * t2[1];
    const double zx = t1[2] * t2[0] - t1[0] * t2[2];
    return sqrt(xy * xy + yz * yz + zx * zx);
  }
  if (fd == 1) return sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  return 1.0;
}

__global__ void k_face_integrals(int64_t F, int nf, int fd, int q, int64_t N, const double* __restrict__ xyz,
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
    const double Ja = face_measure(fd, t1, t2);
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
__global__ void k_face_integrals_gp(int64_t F, int nf, int fd, int q, int64_t N, const double* __restrict__ xyz,
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
    const double Ja = face_measure(fd, t1, t2);
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

// ---- volume loads ---------------------------------------------------------
// Per element e of the set (pos[e] >= 0) and node a:
// s3[e][a][c] = sum_q ((N_a(xi_q) g_c(q)) det J_q) w_q with J_q = sum_b x_b dN_b(xi_q)^T
// (signed det J, as gmsh's getJacobian gives it), g constant or given per
// Gauss point (gp[pos[e]][q][3]); elements outside the set get zeros.
// Connectivity is the mesh's SoA [npe][E].
__global__ void k_body_integrals(int64_t E, int npe, int q, int64_t N, const double* __restrict__ xyz,
                                 const int32_t* __restrict__ conn, const int32_t* __restrict__ pos,
                                 const double* __restrict__ w, const double* __restrict__ Nref,
                                 const double* __restrict__ dN, double g0, double g1, double g2,
                                 const double* __restrict__ gp, double* __restrict__ s3) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double* out = s3 + e * npe * 3;
  for (int a = 0; a < 3 * npe; ++a) out[a] = 0.0;
  const int32_t k = pos[e];
  if (k < 0) return;
  for (int p = 0; p < q; ++p) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int b = 0; b < npe; ++b) {
      const int32_t v = conn[(int64_t)b * E + e];
      const double x[3] = {xyz[v], xyz[N + v], xyz[2 * N + v]};
      const double* g = dN + ((size_t)p * npe + b) * 3;
      for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) J[i][d] += x[i] * g[d];
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double gq[3] = {gp ? gp[((int64_t)k * q + p) * 3] : g0, gp ? gp[((int64_t)k * q + p) * 3 + 1] : g1,
                          gp ? gp[((int64_t)k * q + p) * 3 + 2] : g2};
    for (int a = 0; a < npe; ++a) {
      const double h = Nref[(size_t)p * npe + a];
      for (int c = 0; c < 3; ++c) out[a * 3 + c] += ((h * gq[c]) * det) * w[p];
    }
  }
}

// f[node] = sum of s3 over the node's incidences in ascending element order
// (the mesh's (node, element)-sorted incidence lists).
__global__ void k_body_gather(int64_t N, int npe, const int32_t* __restrict__ inc_ptr,
                              const int32_t* __restrict__ inc_elem, const uint8_t* __restrict__ inc_loc,
                              const double* __restrict__ s3, double* __restrict__ f) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int32_t k = inc_ptr[v]; k < inc_ptr[v + 1]; ++k) {
    const double* s = s3 + ((int64_t)inc_elem[k] * npe + inc_loc[k]) * 3;
    for (int c = 0; c < 3; ++c) acc[c] += s[c];
  }
  for (int c = 0; c < 3; ++c) f[3 * v + c] = acc[c];
}

// Physical coordinates of the Gauss points of the elements elems[0..ne).
__global__ void k_volume_points(int64_t ne, int64_t E, int npe, int q, int64_t N, const double* __restrict__ xyz,
                                const int32_t* __restrict__ conn, const int32_t* __restrict__ elems,
                                const double* __restrict__ Nref, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * q) return;
  const int64_t k = t / q;
  const int p = (int)(t - k * q);
  const int64_t e = elems[k];
  double x[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < npe; ++a) {
    const int32_t v = conn[(int64_t)a * E + e];
    const double h = Nref[(size_t)p * npe + a];
    x[0] += h * xyz[v];
    x[1] += h * xyz[N + v];
    x[2] += h * xyz[2 * N + v];
  }
  for (int i = 0; i < 3; ++i) out[t * 3 + i] = x[i];
}

// ---- pressure ---------------------------------------------------------------
// Unit normal of face e at its own local node k: n = t1 x t2 / |t1 x t2| with
// t1, t2 = sum_a x_a dN_a,xi / dN_a,eta evaluated at node k (dNn [k][a][3]).
__global__ void k_face_node_normals(int64_t F, int nf, int64_t N, const double* __restrict__ xyz,
                                    const int32_t* __restrict__ fconn, const double* __restrict__ dNn,
                                    double* __restrict__ nrm) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F * nf) return;
  const int64_t e = t / nf;
  const int k = (int)(t - e * nf);
  double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
  for (int a = 0; a < nf; ++a) {
    const int32_t v = fconn[e * nf + a];
    const double x[3] = {xyz[v], xyz[N + v], xyz[2 * N + v]};
    const double g0 = dNn[((size_t)k * nf + a) * 3], g1 = dNn[((size_t)k * nf + a) * 3 + 1];
    for (int i = 0; i < 3; ++i) {
      t1[i] += g0 * x[i];
      t2[i] += g1 * x[i];
    }
  }
  const double nx = t1[1] * t2[2] - t1[2] * t2[1];
  const double ny = t1[2] * t2[0] - t1[0] * t2[2];
  const double nz = t1[0] * t2[1] - t1[1] * t2[0];
  const double len = sqrt(nx * nx + ny * ny + nz * nz);
  nrm[t * 3] = nx / len;
  nrm[t * 3 + 1] = ny / len;
  nrm[t * 3 + 2] = nz / len;
}

// Nodal traction of the pressure on face-set node i (node nodes[i]):
// t_c = (sum over its incidences, ascending face order, of (-n_c) p) * (1 / count),
// p = p_nodal[node] or the constant p.
__global__ void k_pressure_nodal(int64_t nnodes, const int32_t* __restrict__ nodes, const int32_t* __restrict__ ptr,
                                 const int32_t* __restrict__ slots, const double* __restrict__ nrm, double p,
                                 const double* __restrict__ p_nodal, double* __restrict__ tn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  const double pv = p_nodal ? p_nodal[nodes[i]] : p;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = ptr[i]; k < ptr[i + 1]; ++k)
    for (int c = 0; c < 3; ++c) acc[c] += (-nrm[(size_t)slots[k] * 3 + c]) * pv;
  const double factor = 1.0 / (double)(ptr[i + 1] - ptr[i]);
  for (int c = 0; c < 3; ++c) tn[i * 3 + c] = acc[c] * factor;
}

// Per face e and node a: s3[e][a][c] = sum_q ((N_a t_c(q)) Ja) w_q with the
// traction interpolated from the nodal values, t(q) = sum_b N_b tn[b].
__global__ void k_face_integrals_nodal(int64_t F, int nf, int q, int64_t N, const double* __restrict__ xyz,
                                       const int32_t* __restrict__ fconn, const double* __restrict__ w,
                                       const double* __restrict__ Nref, const double* __restrict__ dN,
                                       const int32_t* __restrict__ slot_node, const double* __restrict__ tn,
                                       double* __restrict__ s3) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= F) return;
  for (int a = 0; a < 3 * nf; ++a) s3[e * nf * 3 + a] = 0.0;
  for (int p = 0; p < q; ++p) {
    double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
    for (int a = 0; a < nf; ++a) {
      const int32_t v = fconn[e * nf + a];
      const double x[3] = {xyz[v], xyz[N + v], xyz[2 * N + v]};
      const double g0 = dN[((size_t)p * nf + a) * 3], g1 = dN[((size_t)p * nf + a) * 3 + 1];
      const double h = Nref[(size_t)p * nf + a];
      const double* tb = tn + (size_t)slot_node[e * nf + a] * 3;
      for (int i = 0; i < 3; ++i) {
        t1[i] += x[i] * g0;
        t2[i] += x[i] * g1;
        tq[i] += h * tb[i];
      }
    }
    const double xy = t1[0] * t2[1] - t1[1] * t2[0];
    const double yz = t1[1] * t2[2] - t1[2] * t2[1];
    const double zx = t1[2] * t2[0] - t1[0] * t2[2];
    const double Ja = sqrt(xy * xy + yz * yz + zx * zx);
    for (int a = 0; a < nf; ++a) {
      const double h = Nref[(size_t)p * nf + a];
      for (int i = 0; i < 3; ++i) s3[(e * nf + a) * 3 + i] += ((h * tq[i]) * Ja) * w[p];
    }
  }
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
  const int fd = etype_dim(face_type);
  if (fd < 0 || fd > 2)
    return fail(LLFEM_EUNSUP, "loadVector: boundary element type " + std::to_string(face_type) +
                                  " is not a surface, curve or point element");
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
  FacePrep P;
  LLFEM_TRY(prepare_faces(m, face_type, nf, face_conn, F, one_based, q, &P));
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N = m->N;
  const int fd = etype_dim(face_type);
  const std::vector<int32_t>& fc = P.fc;
  const std::vector<int32_t>& nodes = P.nodes;
  const std::vector<int32_t>& ptr = P.ptr;
  const std::vector<int32_t>& slot_order = P.slots;
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
          F, nf, fd, q, N, m->d_xyz, d_fc, d_tab, d_tab + q, d_tab + q + (size_t)q * nf, d_s);
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
          F, nf, etype_dim(face_type), q, N, m->d_xyz, d_fc, d_tab, d_tab + q, d_tab + q + (size_t)q * nf, d_t, d_s3);
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

namespace {

// Host side of an element set of the mesh: pos[e] = index of element e in
// elems (or -1), after range and duplicate checks; elems == NULL is every
// element in mesh order.
int prepare_elems(const llfem_mesh* m, const int64_t* elems, int64_t ne, const char* who,
                  std::vector<int32_t>* pos, std::vector<int32_t>* list) {
  if (ne < 0 || (elems == nullptr && ne != 0)) return fail(LLFEM_EINVAL, std::string(who) + ": bad element set");
  pos->assign((size_t)m->E, -1);
  list->clear();
  if (ne == 0 && elems == nullptr) {
    list->resize((size_t)m->E);
    for (int64_t e = 0; e < m->E; ++e) {
      (*pos)[e] = (int32_t)e;
      (*list)[e] = (int32_t)e;
    }
    return LLFEM_OK;
  }
  list->resize((size_t)ne);
  for (int64_t k = 0; k < ne; ++k) {
    const int64_t e = elems[k];
    if (e < 0 || e >= m->E)
      return fail(LLFEM_EINVAL, std::string(who) + ": element index " + std::to_string(e) + " out of range");
    if ((*pos)[e] >= 0)
      return fail(LLFEM_EINVAL, std::string(who) + ": element " + std::to_string(e) + " listed twice");
    (*pos)[e] = (int32_t)k;
    (*list)[k] = (int32_t)e;
  }
  return LLFEM_OK;
}

}  // namespace

extern "C" int llfem_load_body(llfem_mesh* m, int32_t q, const double* weights, const double* N_ref,
                               const double* dN_ref, const int64_t* elems, int64_t ne, const double g[3],
                               const double* g_gp, double* f_out) {
  if (!m || !weights || !N_ref || !dN_ref || !f_out || (!g && !g_gp))
    return fail(LLFEM_EINVAL, "llfem_load_body: NULL argument");
  if (q <= 0) return fail(LLFEM_EINVAL, "llfem_load_body: bad rule size");
  if (etype_dim(m->etype) != 3)
    return fail(LLFEM_EUNSUP, "llfem_load_body: the mesh's elements are not 3D");
  std::vector<int32_t> pos, list;
  LLFEM_TRY(prepare_elems(m, elems, ne, "llfem_load_body", &pos, &list));
  const int64_t nset = (int64_t)list.size();
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N = m->N, E = m->E;
  const int npe = m->npe;
  const size_t tab_n = (size_t)q + (size_t)q * npe + (size_t)q * npe * 3;
  const size_t ngp = g_gp ? 3 * (size_t)nset * q : 0;
  int32_t* d_pos = nullptr;
  double *d_tab = nullptr, *d_s3 = nullptr, *d_gp = nullptr, *d_f = nullptr;
  int s = LLFEM_OK;
  do {
    if ((s = dalloc(ctx, &d_pos, (size_t)E))) break;
    if ((s = dalloc(ctx, &d_tab, tab_n))) break;
    if ((s = dalloc(ctx, &d_s3, 3 * (size_t)E * npe))) break;
    if (ngp && (s = dalloc(ctx, &d_gp, ngp))) break;
    if ((s = dalloc(ctx, &d_f, 3 * (size_t)N))) break;
    std::vector<double> tab(tab_n);
    std::copy(weights, weights + q, tab.begin());
    std::copy(N_ref, N_ref + (size_t)q * npe, tab.begin() + q);
    std::copy(dN_ref, dN_ref + (size_t)q * npe * 3, tab.begin() + q + (size_t)q * npe);
    hipError_t e = hipMemcpyAsync(d_pos, pos.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab_n, hipMemcpyHostToDevice, st);
    if (ngp) e = e ? e : hipMemcpyAsync(d_gp, g_gp, sizeof(double) * ngp, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) { s = fail(LLFEM_EHIP, std::string("llfem_load_body upload: ") + hipGetErrorString(e)); break; }
    if (E > 0)
      k_body_integrals<<<(unsigned)((E + 127) / 128), 128, 0, st>>>(
          E, npe, q, N, m->d_xyz, m->d_conn, d_pos, d_tab, d_tab + q, d_tab + q + (size_t)q * npe,
          g ? g[0] : 0.0, g ? g[1] : 0.0, g ? g[2] : 0.0, d_gp, d_s3);
    if (N > 0)
      k_body_gather<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(N, npe, m->d_inc_ptr, m->d_inc_elem,
                                                                  m->d_inc_loc, d_s3, d_f);
    e = hipGetLastError();
    e = e ? e : hipMemcpyAsync(f_out, d_f, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, st);
    e = e ? e : hipStreamSynchronize(st);
    if (e != hipSuccess) s = fail(LLFEM_EHIP, std::string("llfem_load_body: ") + hipGetErrorString(e));
  } while (0);
  drelease(ctx, d_pos, (size_t)E);
  drelease(ctx, d_tab, tab_n);
  drelease(ctx, d_s3, 3 * (size_t)E * npe);
  drelease(ctx, d_gp, ngp);
  drelease(ctx, d_f, 3 * (size_t)N);
  return s;
}

extern "C" int llfem_volume_points(llfem_mesh* m, int32_t q, const double* N_ref, const int64_t* elems, int64_t ne,
                                   double* xyz_out) {
  if (!m || !N_ref || !xyz_out) return fail(LLFEM_EINVAL, "llfem_volume_points: NULL argument");
  if (q <= 0) return fail(LLFEM_EINVAL, "llfem_volume_points: bad rule size");
  std::vector<int32_t> pos, list;
  LLFEM_TRY(prepare_elems(m, elems, ne, "llfem_volume_points", &pos, &list));
  const int64_t nset = (int64_t)list.size();
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int npe = m->npe;
  const size_t npts = (size_t)nset * q;
  int32_t* d_list = nullptr;
  double *d_N = nullptr, *d_out = nullptr;
  int s = LLFEM_OK;
  do {
    if ((s = dalloc(ctx, &d_list, (size_t)nset))) break;
    if ((s = dalloc(ctx, &d_N, (size_t)q * npe))) break;
    if ((s = dalloc(ctx, &d_out, 3 * npts))) break;
    hipError_t e = hipMemcpyAsync(d_list, list.data(), sizeof(int32_t) * nset, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_N, N_ref, sizeof(double) * q * npe, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && npts > 0)
      k_volume_points<<<(unsigned)((npts + 255) / 256), 256, 0, st>>>(nset, m->E, npe, q, m->N, m->d_xyz,
                                                                       m->d_conn, d_list, d_N, d_out);
    e = e ? e : hipGetLastError();
    e = e ? e : hipMemcpyAsync(xyz_out, d_out, sizeof(double) * 3 * npts, hipMemcpyDeviceToHost, st);
    e = e ? e : hipStreamSynchronize(st);
    if (e != hipSuccess) s = fail(LLFEM_EHIP, std::string("llfem_volume_points: ") + hipGetErrorString(e));
  } while (0);
  drelease(ctx, d_list, (size_t)nset);
  drelease(ctx, d_N, (size_t)q * npe);
  drelease(ctx, d_out, 3 * npts);
  return s;
}

extern "C" int llfem_load_pressure(llfem_mesh* m, int32_t face_type, int32_t nf, const int64_t* face_conn,
                                   int64_t F, int32_t one_based, int32_t q, const double* weights,
                                   const double* N_ref, const double* dN_ref, const double* dN_nodes, double p,
                                   const double* p_nodal, double* f_out) {
  if (!m || !face_conn || !weights || !N_ref || !dN_ref || !dN_nodes || !f_out)
    return fail(LLFEM_EINVAL, "llfem_load_pressure: NULL argument");
  if (etype_dim(face_type) != 2)
    return fail(LLFEM_EINVAL, "loadVector: pressure can be given on a surface of a 3D solid.");
  FacePrep P;
  LLFEM_TRY(prepare_faces(m, face_type, nf, face_conn, F, one_based, q, &P));
  llfem_ctx* ctx = m->ctx;
  LLFEM_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N = m->N, nn = (int64_t)P.nodes.size();
  std::vector<int32_t> slot_node((size_t)F * nf);
  for (int64_t i = 0; i < nn; ++i)
    for (int32_t k = P.ptr[i]; k < P.ptr[i + 1]; ++k) slot_node[P.slots[k]] = (int32_t)i;
  const size_t tab_n = (size_t)q + (size_t)q * nf + (size_t)q * nf * 3 + (size_t)nf * nf * 3;
  int32_t *d_fc = nullptr, *d_nodes = nullptr, *d_ptr = nullptr, *d_slots = nullptr, *d_sn = nullptr;
  double *d_tab = nullptr, *d_nrm = nullptr, *d_tn = nullptr, *d_s3 = nullptr, *d_p = nullptr, *d_f = nullptr;
  int s = LLFEM_OK;
  do {
    if ((s = dalloc(ctx, &d_fc, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_nodes, (size_t)nn))) break;
    if ((s = dalloc(ctx, &d_ptr, (size_t)nn + 1))) break;
    if ((s = dalloc(ctx, &d_slots, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_sn, (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_tab, tab_n))) break;
    if ((s = dalloc(ctx, &d_nrm, 3 * (size_t)F * nf))) break;
    if ((s = dalloc(ctx, &d_tn, 3 * (size_t)nn))) break;
    if ((s = dalloc(ctx, &d_s3, 3 * (size_t)F * nf))) break;
    if (p_nodal && (s = dalloc(ctx, &d_p, (size_t)N))) break;
    if ((s = dalloc(ctx, &d_f, 3 * (size_t)N))) break;
    std::vector<double> tab(tab_n);
    std::copy(weights, weights + q, tab.begin());
    std::copy(N_ref, N_ref + (size_t)q * nf, tab.begin() + q);
    std::copy(dN_ref, dN_ref + (size_t)q * nf * 3, tab.begin() + q + (size_t)q * nf);
    std::copy(dN_nodes, dN_nodes + (size_t)nf * nf * 3, tab.begin() + q + (size_t)q * nf * 4);
    hipError_t e = hipSuccess;
    e = e ? e : hipMemcpyAsync(d_fc, P.fc.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_nodes, P.nodes.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_ptr, P.ptr.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_slots, P.slots.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_sn, slot_node.data(), sizeof(int32_t) * F * nf, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab_n, hipMemcpyHostToDevice, st);
    if (p_nodal) e = e ? e : hipMemcpyAsync(d_p, p_nodal, sizeof(double) * N, hipMemcpyHostToDevice, st);
    e = e ? e : hipMemsetAsync(d_f, 0, sizeof(double) * 3 * N, st);
    if (e != hipSuccess) { s = fail(LLFEM_EHIP, std::string("loadVector upload: ") + hipGetErrorString(e)); break; }
    if (F > 0) {
      const double* d_w = d_tab;
      const double* d_N = d_tab + q;
      const double* d_dN = d_tab + q + (size_t)q * nf;
      const double* d_dNn = d_tab + q + (size_t)q * nf * 4;
      k_face_node_normals<<<(unsigned)((F * nf + 255) / 256), 256, 0, st>>>(F, nf, N, m->d_xyz, d_fc, d_dNn, d_nrm);
      k_pressure_nodal<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(nn, d_nodes, d_ptr, d_slots, d_nrm, p, d_p,
                                                                     d_tn);
      k_face_integrals_nodal<<<(unsigned)((F + 127) / 128), 128, 0, st>>>(F, nf, q, N, m->d_xyz, d_fc, d_w, d_N,
                                                                          d_dN, d_sn, d_tn, d_s3);
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
  drelease(ctx, d_sn, (size_t)F * nf);
  drelease(ctx, d_tab, tab_n);
  drelease(ctx, d_nrm, 3 * (size_t)F * nf);
  drelease(ctx, d_tn, 3 * (size_t)nn);
  drelease(ctx, d_s3, 3 * (size_t)F * nf);
  drelease(ctx, d_p, (size_t)N);
  drelease(ctx, d_f, 3 * (size_t)N);
  return s;
}
