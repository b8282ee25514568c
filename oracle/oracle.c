/*
 * CPU oracle: a plain-C FP64 restatement of LowLevelFEM.jl's operator-assembly
 * hot path.  TEST INFRASTRUCTURE ONLY -- it is the checker the GPU library is
 * compared against (and the timed CPU baseline); no product path links it.
 *
 * Each function cites the reference code it restates (paths relative to the
 * LowLevelFEM.jl 1.13.31 source tree).  Indices are 0-based here; the
 * reference's DoF rule dof = pdim*(tag-1)+c (src/linear.jl:309-312) becomes
 * dof = pdim*node + c.
 *
 * Reference-element tables (quadrature weights, dN_ref in gmsh layout
 * [q][node][3], N_ref [q][node]) are inputs, exactly as the reference
 * receives them from gmsh.
 *
 * Pinning (see the tests): closed-form hex8 element matrices, the reference's
 * recorded golden values (sum|K| for the 30^3 hex8 cube, K[1:3,1] for 40^3,
 * sum(u) for the 30^3 cantilever), resultant/volume identities of the load
 * vector and mass, CSC pattern counts 9(3n+1)^3, and brute-force
 * dense assembly on tiny meshes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define API __attribute__((visibility("default")))

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* ---------------------------------------------------------------------------
 * build_csc_pattern (src/multifield.jl:2083-2250): per column node, collect
 * the row nodes of every incident element, sort + unique, then expand every
 * node pair to a pdim x pdim block; columns node-major / component-minor,
 * rows sorted.  Two-phase: colptr/rowval may be NULL to query nnz.
 * ------------------------------------------------------------------------- */
API int64_t oracle_csc_pattern(int64_t N, int npe, int64_t E, const int64_t* conn, int pdim,
                               int64_t* colptr, int64_t* rowval, int64_t* adj_ptr, int64_t* adj) {
  int64_t* cnt = (int64_t*)calloc((size_t)N + 1, sizeof(int64_t));
  for (int64_t e = 0; e < E; ++e)
    for (int b = 0; b < npe; ++b) cnt[conn[e * npe + b] + 1] += npe;
  for (int64_t j = 0; j < N; ++j) cnt[j + 1] += cnt[j];
  int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(cnt[N] ? cnt[N] : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
  for (int64_t j = 0; j < N; ++j) fill[j] = cnt[j];
  for (int64_t e = 0; e < E; ++e)
    for (int b = 0; b < npe; ++b) {
      int64_t col = conn[e * npe + b];
      for (int a = 0; a < npe; ++a) rows[fill[col]++] = conn[e * npe + a];
    }
  int64_t* R = (int64_t*)calloc((size_t)N, sizeof(int64_t));
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : total)
  for (int64_t j = 0; j < N; ++j) {
    int64_t lo = cnt[j], n = cnt[j + 1] - cnt[j];
    if (n == 0) continue;
    qsort(rows + lo, (size_t)n, sizeof(int64_t), cmp_i64);
    int64_t w = 0;
    for (int64_t r = 1; r < n; ++r)
      if (rows[lo + r] != rows[lo + w]) rows[lo + (++w)] = rows[lo + r];
    R[j] = w + 1;
    total += w + 1;
  }
  int64_t nnz = total * pdim * pdim;
  if (adj_ptr) {
    adj_ptr[0] = 0;
    for (int64_t j = 0; j < N; ++j) adj_ptr[j + 1] = adj_ptr[j] + R[j];
    if (adj)
      for (int64_t j = 0; j < N; ++j)
        memcpy(adj + adj_ptr[j], rows + cnt[j], sizeof(int64_t) * (size_t)R[j]);
  }
  if (colptr) {
    int64_t pos = 0;
    for (int64_t j = 0; j < N; ++j)
      for (int c = 0; c < pdim; ++c) {
        colptr[j * pdim + c] = pos;
        for (int64_t s = 0; s < R[j]; ++s)
          for (int d = 0; d < pdim; ++d) {
            if (rowval) rowval[pos] = rows[cnt[j] + s] * pdim + d;
            ++pos;
          }
      }
    colptr[N * pdim] = pos;
  }
  free(cnt); free(rows); free(fill); free(R);
  return nnz;
}

/* find_csc_position (src/multifield.jl:2260-2283): binary search of a row in
 * a sorted column range; -1 when missing (the reference raises). */
static int64_t find_csc_position(const int64_t* rowval, int64_t first, int64_t last, int64_t row) {
  int64_t lo = first, hi = last;
  while (lo <= hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    int64_t cur = rowval[mid];
    if (cur < row) lo = mid + 1;
    else if (cur > row) hi = mid - 1;
    else return mid;
  }
  return -1;
}

/* inv3x3_fast (src/linear.jl:89-114): cofactor inverse; J row-major here. */
static double inv3x3(const double* J, double* inv) {
  double c11 = (J[4] * J[8] - J[5] * J[7]);
  double c12 = -(J[3] * J[8] - J[5] * J[6]);
  double c13 = (J[3] * J[7] - J[4] * J[6]);
  double c21 = -(J[1] * J[8] - J[2] * J[7]);
  double c22 = (J[0] * J[8] - J[2] * J[6]);
  double c23 = -(J[0] * J[7] - J[1] * J[6]);
  double c31 = (J[1] * J[5] - J[2] * J[4]);
  double c32 = -(J[0] * J[5] - J[2] * J[3]);
  double c33 = (J[0] * J[4] - J[1] * J[3]);
  double det = J[0] * c11 + J[1] * c12 + J[2] * c13;
  double id = 1.0 / det;
  inv[0] = c11 * id; inv[1] = c21 * id; inv[2] = c31 * id;
  inv[3] = c12 * id; inv[4] = c22 * id; inv[5] = c32 * id;
  inv[6] = c13 * id; inv[7] = c23 * id; inv[8] = c33 * id;
  return det;
}

/* compute_element_geometry! (src/multifield.jl:1666-1743):
 * J[i][al] = sum_a x_i^a dN_ref[q][a][al]; returns det (signed). */
static double element_jacobian(const double* xyz, const int64_t* en, int npe, const double* dNq,
                               double* J) {
  for (int k = 0; k < 9; ++k) J[k] = 0.0;
  for (int a = 0; a < npe; ++a) {
    const double* x = xyz + 3 * en[a];
    for (int al = 0; al < 3; ++al) {
      double g = dNq[3 * a + al];
      J[0 * 3 + al] += x[0] * g;
      J[1 * 3 + al] += x[1] * g;
      J[2 * 3 + al] += x[2] * g;
    }
  }
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
         J[2] * (J[3] * J[7] - J[4] * J[6]);
}

/* constitutive_matrix(3, :Solid, E, nu) (src/linear.jl:154-165). */
API void oracle_hooke_D(double E, double nu, double* D) {
  double f = E / ((1 + nu) * (1 - 2 * nu));
  for (int k = 0; k < 36; ++k) D[k] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i * 6 + j] = f * (i == j ? 1 - nu : nu);
  for (int i = 3; i < 6; ++i) D[i * 6 + i] = f * (1 - 2 * nu) / 2;
}

/* Physical gradients of one GP: dh[a][i] = sum_al invJ^T[i][al] dN[a][al]. */
static void phys_grad(const double* invJ, const double* dNq, int npe, double* dh) {
  for (int a = 0; a < npe; ++a)
    for (int i = 0; i < 3; ++i)
      dh[3 * a + i] = invJ[0 * 3 + i] * dNq[3 * a + 0] + invJ[1 * 3 + i] * dNq[3 * a + 1] +
                      invJ[2 * 3 + i] * dNq[3 * a + 2];
}

/* B rows for one GP.  legacy=1: fill_B_matrix!(..., Val(3)) row order
 * [xx,yy,zz,xy,yz,zx] (src/linear.jl:137-151); legacy=0: build_B!(SymGradOp)
 * row order [xx,yy,zz,xy,xz,yz] (src/multifield.jl:607-661). B is 6 x 3n. */
static void fill_B(const double* dh, int npe, int legacy, double* B) {
  int n3 = 3 * npe;
  memset(B, 0, sizeof(double) * 6 * (size_t)n3);
  for (int a = 0; a < npe; ++a) {
    double dx = dh[3 * a], dy = dh[3 * a + 1], dz = dh[3 * a + 2];
    int cx = 3 * a, cy = 3 * a + 1, cz = 3 * a + 2;
    B[0 * n3 + cx] = dx;
    B[1 * n3 + cy] = dy;
    B[2 * n3 + cz] = dz;
    B[3 * n3 + cx] = dy; B[3 * n3 + cy] = dx;
    if (legacy) {
      B[4 * n3 + cy] = dz; B[4 * n3 + cz] = dy;   /* gamma_yz */
      B[5 * n3 + cx] = dz; B[5 * n3 + cz] = dx;   /* gamma_zx */
    } else {
      B[4 * n3 + cx] = dz; B[4 * n3 + cz] = dx;   /* gamma_xz */
      B[5 * n3 + cy] = dz; B[5 * n3 + cz] = dy;   /* gamma_yz */
    }
  }
}

/* Ke += s * B^T D B (3n x 3n, row-major). */
static void add_BtDB(const double* B, const double* D, int n3, double s, double* Ke, double* tmp) {
  for (int I = 0; I < 6; ++I)
    for (int c = 0; c < n3; ++c) {
      double acc = 0.0;
      for (int J = 0; J < 6; ++J) acc += D[I * 6 + J] * B[J * n3 + c];
      tmp[I * n3 + c] = acc;
    }
  for (int r = 0; r < n3; ++r)
    for (int c = 0; c < n3; ++c) {
      double acc = 0.0;
      for (int I = 0; I < 6; ++I) acc += B[I * n3 + r] * tmp[I * n3 + c];
      Ke[r * n3 + c] += s * acc;
    }
}

/* Element stiffness, both reference variants.
 * legacy=1 -> stiffnessMatrixSolid (src/linear.jl:272-307):
 *   K1 = sum_q B'DB * b * det_q * w_q with the signed det.
 * legacy=0 -> assemble_element_matrix_with_geometry! (src/multifield.jl:1857-2052):
 *   Ke += (|det_q| w_q) * Bs' C Bu.
 * Returns 0, or -1 on singular geometry (|det| == 0, src/multifield.jl:1737). */
static int element_stiffness(const double* xyz, const int64_t* en, int npe, int q, const double* w,
                             const double* dN, const double* D, int legacy, double* Ke,
                             double* work) {
  int n3 = 3 * npe;
  double* dh = work;
  double* B = dh + n3;
  double* tmp = B + 6 * n3;
  memset(Ke, 0, sizeof(double) * (size_t)n3 * n3);
  for (int k = 0; k < q; ++k) {
    double J[9], invJ[9];
    const double* dNq = dN + (size_t)k * npe * 3;
    double det = element_jacobian(xyz, en, npe, dNq, J);
    if (!legacy && !(fabs(det) > 0.0)) return -1;
    inv3x3(J, invJ);
    phys_grad(invJ, dNq, npe, dh);
    fill_B(dh, npe, legacy, B);
    add_BtDB(B, D, n3, (legacy ? det : fabs(det)) * w[k], Ke, tmp);
  }
  return 0;
}

/* Legacy triplets of stiffnessMatrixSolid (src/linear.jl:309-316): entry
 * (k,l) of K1 goes to I = nn2[l], J = nn2[k] (stored transposed), element
 * order, k fastest.  I/J/V have E*(3n)^2 slots. */
API int oracle_stiffness_ijv(int64_t N, const double* xyz, int npe, int64_t E, const int64_t* conn,
                             int q, const double* w, const double* dN, double Emod, double nu,
                             int64_t* I, int64_t* J, double* V, int nthreads) {
  (void)N;
  double D[36];
  oracle_hooke_D(Emod, nu, D);
  int n3 = 3 * npe;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double* Ke = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
    double* work = (double*)malloc(sizeof(double) * (size_t)(n3 + 12 * n3));
#pragma omp for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
      const int64_t* en = conn + e * npe;
      element_stiffness(xyz, en, npe, q, w, dN, D, 1, Ke, work);
      int64_t base = e * (int64_t)n3 * n3;
      for (int l = 0; l < n3; ++l)
        for (int k = 0; k < n3; ++k) {
          int64_t t = base + (int64_t)l * n3 + k;
          I[t] = 3 * en[l / 3] + l % 3;
          J[t] = 3 * en[k / 3] + k % 3;
          V[t] = Ke[k * n3 + l];
        }
    }
    free(Ke); free(work);
  }
  return 0;
}

/* SparseArrays.sparse(I,J,V,m,n) with `+` combine in input order, followed by
 * optional dropzeros! (call sites src/linear.jl:323-324).  Stable counting
 * sort by column then row; duplicates summed left to right.  Returns nnz;
 * outputs must hold ntrip entries (colptr n+1). */
API int64_t oracle_sparse(int64_t n, int64_t ntrip, const int64_t* I, const int64_t* J,
                          const double* V, int dropzeros, int64_t* colptr, int64_t* rowval,
                          double* nzval) {
  int64_t* ccnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t* rcnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t t = 0; t < ntrip; ++t) { rcnt[I[t] + 1]++; }
  for (int64_t i = 0; i < n; ++i) rcnt[i + 1] += rcnt[i];
  int64_t* byrow = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ntrip ? ntrip : 1));
  for (int64_t t = 0; t < ntrip; ++t) byrow[rcnt[I[t]]++] = t;     /* stable by row */
  for (int64_t t = 0; t < ntrip; ++t) ccnt[J[t] + 1]++;
  for (int64_t j = 0; j < n; ++j) ccnt[j + 1] += ccnt[j];
  int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ntrip ? ntrip : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t j = 0; j < n; ++j) fill[j] = ccnt[j];
  /* stable by (column, row): scan by row, place by column; duplicates of one
     (row,col) keep input order because byrow is stable in t. */
  for (int64_t u = 0; u < ntrip; ++u) { int64_t t = byrow[u]; order[fill[J[t]]++] = t; }
  int64_t pos = 0;
  for (int64_t j = 0; j < n; ++j) {
    colptr[j] = pos;
    int64_t u = ccnt[j], end = ccnt[j + 1];
    while (u < end) {
      int64_t t = order[u];
      int64_t row = I[t];
      double s = V[t];
      ++u;
      while (u < end && I[order[u]] == row) { s += V[order[u]]; ++u; }
      if (dropzeros && s == 0.0) continue;
      rowval[pos] = row;
      nzval[pos] = s;
      ++pos;
    }
  }
  colptr[n] = pos;
  free(ccnt); free(rcnt); free(byrow); free(order); free(fill);
  return pos;
}

/* Direct-CSC assembly: assemble_operator(...; assembly=:csc) with
 * assemble_csc_worker! (src/multifield.jl:2394-2499) and, for nthreads > 1,
 * per-worker nzval buffers + reduce_csc_buffers! (src/multifield.jl:2345-2380).
 * Element chunks are assigned cyclically (src/multifield.jl:2421-2424).
 * D is the 6x6 coefficient in SymGrad row order.  Returns 0, -1 on singular
 * geometry, -2 on a missing pattern row. */
API int oracle_stiffness_csc(int64_t N, const double* xyz, int npe, int64_t E, const int64_t* conn,
                             int q, const double* w, const double* dN, const double* D,
                             const int64_t* colptr, const int64_t* rowval, double* nzval,
                             int nthreads, int64_t chunk) {
  int n3 = 3 * npe;
  int64_t nnz = colptr[3 * N];
  int T = nthreads > 0 ? nthreads : 1;
  if (chunk <= 0) chunk = 4096;
  double** bufs = (double**)malloc(sizeof(double*) * (size_t)T);
  bufs[0] = nzval;
  memset(nzval, 0, sizeof(double) * (size_t)nnz);
  for (int t = 1; t < T; ++t) bufs[t] = (double*)calloc((size_t)nnz, sizeof(double));
  int status = 0;
#pragma omp parallel num_threads(T)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double* buf = bufs[t];
    double* Ke = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
    double* work = (double*)malloc(sizeof(double) * (size_t)(13 * n3));
    for (int64_t c0 = (int64_t)t * chunk; c0 < E; c0 += chunk * T) {
      int64_t c1 = c0 + chunk < E ? c0 + chunk : E;
      for (int64_t e = c0; e < c1; ++e) {
        const int64_t* en = conn + e * npe;
        if (element_stiffness(xyz, en, npe, q, w, dN, D, 0, Ke, work)) {
#pragma omp atomic write
          status = -1;
          continue;
        }
        for (int nb = 0; nb < npe; ++nb) {
          int64_t col = 3 * en[nb];
          int64_t first = colptr[col], last = colptr[col + 1] - 1, cnnz = last - first + 1;
          for (int na = 0; na < npe; ++na) {
            int64_t p = find_csc_position(rowval, first, last, 3 * en[na]);
            if (p < 0) {
#pragma omp atomic write
              status = -2;
              continue;
            }
            for (int cb = 0; cb < 3; ++cb)
              for (int ca = 0; ca < 3; ++ca)
                buf[p + cb * cnnz + ca] += Ke[(3 * na + ca) * n3 + 3 * nb + cb];
          }
        }
      }
    }
    free(Ke); free(work);
  }
  if (T > 1) {
#pragma omp parallel for schedule(static) num_threads(T)
    for (int64_t p = 0; p < nnz; ++p) {
      double s = nzval[p];
      for (int t = 1; t < T; ++t) s += bufs[t][p];
      nzval[p] = s;
    }
  }
  for (int t = 1; t < T; ++t) free(bufs[t]);
  free(bufs);
  return status;
}

/* Consistent mass element matrix of massMatrixSolid (src/linear.jl:1966-1972):
 * M1 = rho * sum_q H'H det_q w_q; only equal-component entries are non-zero.
 * Emits triplets like the stiffness (transposed, src/linear.jl:1983-1991). */
API int oracle_mass_ijv(int64_t N, const double* xyz, int npe, int64_t E, const int64_t* conn, int q,
                        const double* w, const double* Nref, const double* dN, double rho,
                        int64_t* I, int64_t* J, double* V) {
  (void)N;
  int n3 = 3 * npe;
#pragma omp parallel
  {
    double* M1 = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
      const int64_t* en = conn + e * npe;
      memset(M1, 0, sizeof(double) * (size_t)n3 * n3);
      for (int k = 0; k < q; ++k) {
        double Jm[9];
        double det = element_jacobian(xyz, en, npe, dN + (size_t)k * npe * 3, Jm);
        double s = det * w[k];
        const double* Nk = Nref + (size_t)k * npe;
        for (int a = 0; a < npe; ++a)
          for (int b = 0; b < npe; ++b)
            for (int c = 0; c < 3; ++c) M1[(3 * a + c) * n3 + 3 * b + c] += Nk[a] * Nk[b] * s;
      }
      int64_t base = e * (int64_t)n3 * n3;
      for (int l = 0; l < n3; ++l)
        for (int k = 0; k < n3; ++k) {
          int64_t t = base + (int64_t)l * n3 + k;
          I[t] = 3 * en[l / 3] + l % 3;
          J[t] = 3 * en[k / 3] + k % 3;
          V[t] = M1[k * n3 + l] * rho;
        }
    }
    free(M1);
  }
  return 0;
}

/* Traction part of loadVector (src/linear.jl:3246-3413) for a boundary
 * element set of dimension fd in 3D: per GP, Ja = |t1 x t2| for a surface
 * (fd 2, src/linear.jl:3380-3384), |t1| for a curve (fd 1, :3385-3386) and 1
 * for a point (fd 0, :3387-3388), from the element's Jacobian columns;
 * f1 += H1' * t * Ja * w; fp[nn2] += f1.  dN is the element table [q][nf][3]. */
API int oracle_traction(int64_t N, const double* xyz, int nf, int fd, int64_t F, const int64_t* fconn, int q,
                        const double* w, const double* Nref, const double* dN, const double* t,
                        double* f) {
  memset(f, 0, sizeof(double) * 3 * (size_t)N);
  for (int64_t e = 0; e < F; ++e) {
    const int64_t* en = fconn + e * nf;
    double f1[3 * 64];
    memset(f1, 0, sizeof(f1));
    for (int k = 0; k < q; ++k) {
      double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
      const double* dNq = dN + (size_t)k * nf * 3;
      for (int a = 0; a < nf; ++a) {
        const double* x = xyz + 3 * en[a];
        for (int i = 0; i < 3; ++i) {
          t1[i] += x[i] * dNq[3 * a];
          t2[i] += x[i] * dNq[3 * a + 1];
        }
      }
      double Ja = 1.0;
      if (fd == 2) {
        double xy = t1[0] * t2[1] - t1[1] * t2[0];
        double yz = t1[1] * t2[2] - t1[2] * t2[1];
        double zx = t1[2] * t2[0] - t1[0] * t2[2];
        Ja = sqrt(xy * xy + yz * yz + zx * zx);
      } else if (fd == 1) {
        Ja = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
      }
      for (int a = 0; a < nf; ++a)
        for (int c = 0; c < 3; ++c) f1[3 * a + c] += Nref[k * nf + a] * t[c] * Ja * w[k];
    }
    for (int a = 0; a < nf; ++a)
      for (int c = 0; c < 3; ++c) f[3 * en[a] + c] += f1[3 * a + c];
  }
  return 0;
}

/* y = K x for a symmetric K stored in CSC, evaluated as the CSR of K^T so
 * rows parallelise (K^T = K). */
static void spmv_sym(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* val,
                     const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) {
    double s = 0.0;
    for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) s += val[p] * x[rowval[p]];
    y[j] = s;
  }
}

static double dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* solveDisplacement(K, f; support, iterative=true, preconditioner=Jacobi)
 * (src/linear.jl:3711-3743): u_fixed from the supports, rhs = f - K[:,fixed]
 * u_fixed restricted to the free DoFs, then IterativeSolvers.cg semantics
 * (x0 = 0, stop when ||r|| <= reltol*||b||) on K_ff applied as masked K.
 * Returns the iteration count (negative if maxiter was hit). */
API int64_t oracle_pcg(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* val,
                       const double* f, const uint8_t* fixed, const double* ufix, double reltol,
                       int64_t maxiter, double* u, double* relres_out) {
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)calloc((size_t)n, sizeof(double));
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  double* dinv = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) u[i] = fixed[i] ? ufix[i] : 0.0;
  spmv_sym(n, colptr, rowval, val, u, q);                 /* f_kin = K u_fixed */
  for (int64_t j = 0; j < n; ++j) {
    double d = 0.0;
    for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k)
      if (rowval[k] == j) d = val[k];
    dinv[j] = (fixed[j] || d == 0.0) ? 0.0 : 1.0 / d;
    r[j] = fixed[j] ? 0.0 : f[j] - q[j];
  }
  double* x = (double*)calloc((size_t)n, sizeof(double));
  double bnorm = sqrt(dot(n, r, r));
  double tol = reltol * bnorm, rho = 1.0, res = bnorm;
  int64_t it = 0;
  while (res > tol && it < maxiter) {
    for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    double rho_prev = rho;
    rho = dot(n, r, z);
    double beta = it == 0 ? 0.0 : rho / rho_prev;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    spmv_sym(n, colptr, rowval, val, p, q);
    for (int64_t i = 0; i < n; ++i)
      if (fixed[i]) q[i] = 0.0;
    double alpha = rho / dot(n, p, q);
    for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    res = sqrt(dot(n, r, r));
    ++it;
  }
  for (int64_t i = 0; i < n; ++i)
    if (!fixed[i]) u[i] = x[i];
  if (relres_out) *relres_out = bnorm > 0 ? res / bnorm : 0.0;
  free(r); free(z); free(p); free(q); free(dinv); free(x);
  return res <= tol ? it : -it;
}

API int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
