// CPU check of the AMG hierarchy build (amg_build_host): no device calls.
// Reads a node-level pattern, strong-coupling flags, coordinates and a
// support mask (raw little-endian files in the directory given as argv[1]),
// builds the hierarchy and checks its invariants:
//   * every pseudo-node with a neighbour belongs to exactly one aggregate,
//     and the member lists agree with agg[];
//   * the tentative prolongator of each aggregate has orthonormal kept
//     columns and zero dropped ones, and its coarse mask marks the dropped;
//   * at level 0 the rigid-body modes of each aggregate (constrained rows
//     zeroed) lie in the span of its columns;
//   * every fine block whose column and row aggregates exist appears in
//     exactly one Galerkin contribution list, the one of that aggregate pair;
//   * coarse patterns are structurally symmetric with the diagonal slot right.
// Prints "ok <levels> <sizes...>" and exits 0, or prints the first failure.
#include "../lowlevelfem.jl_amd/csrc/amg.hip"

#include <cstdio>
#include <string>

template <class T>
static std::vector<T> rd(const std::string& f) {
  FILE* fp = fopen(f.c_str(), "rb");
  if (!fp) {
    fprintf(stderr, "missing %s\n", f.c_str());
    exit(2);
  }
  fseek(fp, 0, SEEK_END);
  const long n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  std::vector<T> v(n / sizeof(T));
  if (fread(v.data(), 1, n, fp) != (size_t)n) exit(2);
  fclose(fp);
  return v;
}

#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      printf("FAIL: " __VA_ARGS__);      \
      printf("\n");                      \
      return 1;                          \
    }                                    \
  } while (0)

int main(int argc, char** argv) {
  const std::string d = argc > 1 ? argv[1] : ".";
  llfem::AmgInput in;
  in.adj_ptr = rd<int64_t>(d + "/adj_ptr.bin");
  {
    const std::vector<int32_t> a = rd<int32_t>(d + "/adj.bin");
    in.adj.assign(a.begin(), a.end());
  }
  in.dslot = rd<int32_t>(d + "/dslot.bin");
  {
    const std::vector<uint8_t> a = rd<uint8_t>(d + "/strong.bin");
    in.strong.assign(a.begin(), a.end());
  }
  in.xyz = rd<double>(d + "/xyz.bin");
  in.mask = rd<uint8_t>(d + "/mask.bin");
  in.N = (int64_t)in.adj_ptr.size() - 1;
  llfem::AmgHost H;
  const int rc = llfem::amg_build_host(in, &H);
  CHECK(rc == 0, "amg_build_host returned %d: %s", rc, llfem_last_error());
  const int L = (int)H.lev.size() - 1;
  CHECK(L >= 1, "no coarse level");
  for (int l = 0; l < L; ++l) {
    const llfem::HostLevel& F = H.lev[l];
    const llfem::HostLevel& C = H.lev[l + 1];
    const std::vector<int64_t>& ap = l == 0 ? in.adj_ptr : F.adj_ptr;
    const int32_t* aj = l == 0 ? in.adj.data() : F.adj.data();
    // membership
    std::vector<int> seen(F.N, 0);
    for (int64_t a = 0; a < F.nagg; ++a)
      for (int32_t m = F.mem_ptr[a]; m < F.mem_ptr[a + 1]; ++m) {
        const int32_t j = F.mem[m];
        CHECK(F.agg[j] == a, "level %d: member %d of aggregate %ld has agg %d", l, j, (long)a, F.agg[j]);
        ++seen[j];
      }
    for (int64_t j = 0; j < F.N; ++j) {
      const bool isolated = ap[j + 1] - ap[j] <= 1;   // only itself
      CHECK(seen[j] == (F.agg[j] >= 0 ? 1 : 0), "level %d: pseudo-node %ld listed %d times", l, (long)j, seen[j]);
      CHECK(F.agg[j] >= 0 || isolated, "level %d: connected pseudo-node %ld has no aggregate", l, (long)j);
    }
    // tentative prolongator columns
    for (int64_t a = 0; a < F.nagg; ++a) {
      double G[6][6] = {};
      for (int32_t m = F.mem_ptr[a]; m < F.mem_ptr[a + 1]; ++m) {
        const float* Q = &F.Q[18 * (size_t)F.mem[m]];
        for (int u = 0; u < 6; ++u)
          for (int v = 0; v < 6; ++v)
            for (int c = 0; c < 3; ++c) G[u][v] += (double)Q[6 * c + u] * Q[6 * c + v];
      }
      for (int u = 0; u < 6; ++u) {
        const bool dropped = C.mask[6 * a + u] != 0;
        for (int v = 0; v < 6; ++v) {
          const double want = (u == v && !dropped) ? 1.0 : 0.0;
          CHECK(std::fabs(G[u][v] - (dropped && u == v ? 0.0 : want)) < 1e-5,
                "level %d aggregate %ld: Q^T Q[%d][%d] = %g", l, (long)a, u, v, G[u][v]);
        }
      }
      if (l == 0) {   // rigid-body modes of the aggregate lie in span(Q)
        double c0[3] = {0, 0, 0};
        const int64_t N = in.N;
        for (int k = 0; k < 3; ++k) {
          for (int64_t j = 0; j < N; ++j) c0[k] += in.xyz[(size_t)k * N + j];
          c0[k] /= N;
        }
        std::vector<double> B;   // rows of the aggregate, 6 modes
        std::vector<const float*> Qr;
        for (int32_t m = F.mem_ptr[a]; m < F.mem_ptr[a + 1]; ++m) {
          const int64_t j = F.mem[m];
          const double r[3] = {in.xyz[j] - c0[0], in.xyz[N + j] - c0[1], in.xyz[2 * N + j] - c0[2]};
          for (int c = 0; c < 3; ++c) {
            double row[6] = {0, 0, 0, 0, 0, 0};
            if (!in.mask[3 * j + c]) {
              row[c] = 1.0;
              if (c == 0) { row[4] = r[2]; row[5] = -r[1]; }
              if (c == 1) { row[3] = -r[2]; row[5] = r[0]; }
              if (c == 2) { row[3] = r[1]; row[4] = -r[0]; }
            }
            B.insert(B.end(), row, row + 6);
            Qr.push_back(&F.Q[18 * (size_t)j + 6 * c]);
          }
        }
        const size_t rows = Qr.size();
        for (int k = 0; k < 6; ++k) {
          double coef[6] = {0, 0, 0, 0, 0, 0}, nb = 0.0, nres = 0.0;
          for (size_t r = 0; r < rows; ++r)
            for (int u = 0; u < 6; ++u) coef[u] += Qr[r][u] * B[6 * r + k];
          for (size_t r = 0; r < rows; ++r) {
            double p = 0.0;
            for (int u = 0; u < 6; ++u) p += Qr[r][u] * coef[u];
            nres += (B[6 * r + k] - p) * (B[6 * r + k] - p);
            nb += B[6 * r + k] * B[6 * r + k];
          }
          CHECK(nres <= 1e-8 * std::max(nb, 1e-300), "aggregate %ld: mode %d outside span(Q) (%g of %g)", (long)a,
                k, nres, nb);
        }
      }
    }
    // Galerkin contribution lists
    const int64_t nadj = ap[F.N];
    std::vector<int> hit(nadj, 0);
    const int64_t npairs = (int64_t)F.pair_o.size() / 2;
    for (int64_t A = 0; A < F.nagg; ++A) {
      const int64_t b0 = C.adj_ptr[2 * A];
      const int64_t nb = (C.adj_ptr[2 * A + 1] - b0) / 2;
      for (int64_t k = 0; k < nb; ++k) {
        const int64_t p = b0 / 4 + k;   // pairs of A start at 4 * (pair index) in the coarse pattern
        CHECK(p < npairs && F.pair_o[2 * p] == b0 + 2 * k, "level %d: pair offset of aggregate %ld", l, (long)A);
        const int32_t B = C.adj[b0 + 2 * k] / 2;
        for (int64_t c = F.gptr[p]; c < F.gptr[p + 1]; ++c) {
          const int32_t g = F.gblk[c], j = F.gcol[c];
          CHECK(g >= ap[j] && g < ap[j + 1], "level %d: block %d not in column %d", l, g, j);
          CHECK(F.agg[j] == A && F.agg[aj[g]] == B, "level %d: block %d filed under the wrong pair", l, g);
          ++hit[g];
        }
      }
    }
    for (int64_t j = 0; j < F.N; ++j)
      for (int64_t g = ap[j]; g < ap[j + 1]; ++g) {
        const bool want = F.agg[j] >= 0 && F.agg[aj[g]] >= 0;
        CHECK(hit[g] == (want ? 1 : 0), "level %d: fine block %ld contributes %d times", l, (long)g, hit[g]);
      }
    // coarse pattern symmetry and diagonal slots
    for (int64_t J = 0; J < C.N; ++J) {
      CHECK(C.adj[C.adj_ptr[J] + C.dslot[J]] == J, "level %d: diagonal slot of %ld", l + 1, (long)J);
      for (int64_t g = C.adj_ptr[J]; g < C.adj_ptr[J + 1]; ++g) {
        const int32_t I = C.adj[g];
        bool back = false;
        for (int64_t h = C.adj_ptr[I]; h < C.adj_ptr[I + 1] && !back; ++h) back = C.adj[h] == J;
        CHECK(back, "level %d: coarse pattern not symmetric (%ld, %d)", l + 1, (long)J, I);
      }
    }
  }
  printf("ok %d", L + 1);
  for (const auto& lv : H.lev) printf(" %ld", (long)lv.N);
  printf("\n");
  return 0;
}
