// Host set-up of the smoothed-aggregation multigrid preconditioner (an addition: the
// reference solves its CG with point-Jacobi scaling only, static_implicit.py:314-321).
//
// Built once per model from a symmetric positive definite tangent (the free-DOF CSR
// at the start of a solve) and the rigid-body modes of the mesh:
//   * node groups (the free rows of a node) are aggregated greedily on the group graph
//     (roots whose whole neighbourhood is free, then attachment to the neighbouring
//     aggregate with the most links, then leftovers on their own);
//   * the tentative prolongator T holds, per aggregate, the orthonormal basis Q of the
//     six rigid-body modes restricted to its DOFs (Householder QR); the coarse level's
//     near-null space is the R factor, so each aggregate carries 6 coarse DOFs (a mode
//     that is not independent on an aggregate gets a zero column, and its coarse DOF an
//     identity row);
//   * P = (I - omega D^-1 A) T with D the node-block diagonal and omega = 4 / (3 rho),
//     rho = lambda_max(D^-1 A) by power iteration;
//   * the coarse operator is the Galerkin product P^T A P, in 6x6 blocks;
//   * levels are added until at most `max_coarse` DOFs remain; the coarsest operator is
//     inverted densely (Cholesky).
// Everything here is host code (std::thread); the GPU applies the hierarchy (pcg.hip).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/fea_amd.h"

namespace fea {
void set_error(const char *fmt, ...);
}

namespace {

constexpr int NM = 6;  // rigid-body modes = coarse DOFs per aggregate

template <typename F>
void parallel_for(int n, int threads, F f) {
  if (threads <= 1 || n < 1024) {
    f(0, 0, n);
    return;
  }
  std::vector<std::thread> ts;
  const int chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const int b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back(f, t, b, e);
  }
  for (auto &th : ts) th.join();
}

struct Csr {
  int n = 0;
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;
};

// one level of the hierarchy (host)
struct Level {
  Csr A;                          // operator of this level (scalar CSR)
  std::vector<int32_t> gstart;    // node groups: rows gstart[g] .. gstart[g+1]-1
  std::vector<int32_t> group_of;  // row -> group
  std::vector<double> B;          // near-null space, n x NM row-major
  // smoother: inverse of each group's diagonal block (nr x nr, stored 6x6 row-major)
  std::vector<double> dinv;
  double rho = 0.0;
  // prolongator to the next level: per row, blocks (aggregate, 6 values)
  std::vector<int32_t> prow, pagg;
  std::vector<double> pval;
  int nagg = 0;
};

struct Hierarchy {
  std::vector<Level> lv;
  std::vector<double> coarse_inv;  // dense inverse of the coarsest operator
  int threads = 1;
};

// inverse of the symmetric part of an nr x nr block (nr <= 6) by Gauss-Jordan with
// partial pivoting; a singular block falls back to the inverse of its |diagonal|
void block_inverse(const double *A, int nr, double *out) {
  double M[6][12];
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nr; ++j) {
      M[i][j] = 0.5 * (A[i * 6 + j] + A[j * 6 + i]);
      M[i][nr + j] = (i == j) ? 1.0 : 0.0;
    }
  bool ok = true;
  for (int c = 0; c < nr && ok; ++c) {
    int p = c;
    for (int r = c + 1; r < nr; ++r)
      if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
    if (!(std::fabs(M[p][c]) > 0.0)) {
      ok = false;
      break;
    }
    if (p != c)
      for (int k = 0; k < 2 * nr; ++k) std::swap(M[p][k], M[c][k]);
    const double piv = 1.0 / M[c][c];
    for (int k = 0; k < 2 * nr; ++k) M[c][k] *= piv;
    for (int r = 0; r < nr; ++r) {
      if (r == c) continue;
      const double f = M[r][c];
      if (f != 0.0)
        for (int k = 0; k < 2 * nr; ++k) M[r][k] -= f * M[c][k];
    }
  }
  std::fill(out, out + 36, 0.0);
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nr; ++j) out[i * 6 + j] = ok ? M[i][nr + j] : 0.0;
  if (!ok)
    for (int i = 0; i < nr; ++i) {
      const double d = std::fabs(A[i * 6 + i]);
      out[i * 6 + i] = d > 0.0 ? 1.0 / d : 1.0;
    }
}

void group_diag_inverses(Level &L) {
  const int ng = (int)L.gstart.size() - 1;
  L.dinv.assign((size_t)ng * 36, 0.0);
  for (int g = 0; g < ng; ++g) {
    const int r0 = L.gstart[g], nr = L.gstart[g + 1] - r0;
    double blk[36] = {0};
    for (int a = 0; a < nr; ++a)
      for (int k = L.A.rowptr[r0 + a]; k < L.A.rowptr[r0 + a + 1]; ++k) {
        const int c = L.A.col[k] - r0;
        if (c >= 0 && c < nr) blk[a * 6 + c] = L.A.val[k];
      }
    block_inverse(blk, nr, &L.dinv[(size_t)g * 36]);
  }
}

// y = D^-1 A x over the groups
void dinv_a(const Level &L, const std::vector<double> &x, std::vector<double> &y, int threads) {
  const int ng = (int)L.gstart.size() - 1;
  parallel_for(ng, threads, [&](int, int b, int e) {
    double t[6];
    for (int g = b; g < e; ++g) {
      const int r0 = L.gstart[g], nr = L.gstart[g + 1] - r0;
      for (int a = 0; a < nr; ++a) {
        double s = 0.0;
        for (int k = L.A.rowptr[r0 + a]; k < L.A.rowptr[r0 + a + 1]; ++k) s += L.A.val[k] * x[L.A.col[k]];
        t[a] = s;
      }
      const double *D = &L.dinv[(size_t)g * 36];
      for (int a = 0; a < nr; ++a) {
        double s = 0.0;
        for (int c = 0; c < nr; ++c) s += D[a * 6 + c] * t[c];
        y[r0 + a] = s;
      }
    }
  });
}

// lambda_max(D^-1 A) by power iteration from a fixed pseudo-random start
double power_rho(const Level &L, int iters, int threads) {
  const int n = L.A.n;
  std::vector<double> x(n), y(n);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < n; ++i) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    x[i] = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  double lam = 0.0;
  for (int it = 0; it < iters; ++it) {
    double nx = 0.0;
    for (int i = 0; i < n; ++i) nx += x[i] * x[i];
    nx = std::sqrt(nx);
    if (nx == 0.0) return 1.0;
    for (int i = 0; i < n; ++i) x[i] /= nx;
    dinv_a(L, x, y, threads);
    double ny = 0.0;
    for (int i = 0; i < n; ++i) ny += y[i] * y[i];
    lam = std::sqrt(ny);
    std::swap(x, y);
  }
  return lam;
}

// group adjacency (groups linked by a nonzero of A), each list sorted, self excluded
void group_graph(const Level &L, std::vector<int32_t> &gptr, std::vector<int32_t> &gadj) {
  const int ng = (int)L.gstart.size() - 1;
  std::vector<std::vector<int32_t>> adj(ng);
  for (int g = 0; g < ng; ++g) {
    std::vector<int32_t> &a = adj[g];
    for (int r = L.gstart[g]; r < L.gstart[g + 1]; ++r)
      for (int k = L.A.rowptr[r]; k < L.A.rowptr[r + 1]; ++k) {
        const int h = L.group_of[L.A.col[k]];
        if (h != g) a.push_back(h);
      }
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  gptr.assign(ng + 1, 0);
  for (int g = 0; g < ng; ++g) gptr[g + 1] = gptr[g] + (int)adj[g].size();
  gadj.resize(gptr[ng]);
  for (int g = 0; g < ng; ++g) std::copy(adj[g].begin(), adj[g].end(), gadj.begin() + gptr[g]);
}

int aggregate(const std::vector<int32_t> &gptr, const std::vector<int32_t> &gadj, std::vector<int32_t> &agg) {
  const int ng = (int)gptr.size() - 1;
  agg.assign(ng, -1);
  int na = 0;
  for (int g = 0; g < ng; ++g) {  // roots: the group and all its neighbours unassigned
    if (agg[g] >= 0) continue;
    bool free_nb = true;
    for (int k = gptr[g]; k < gptr[g + 1] && free_nb; ++k) free_nb = agg[gadj[k]] < 0;
    if (!free_nb) continue;
    agg[g] = na;
    for (int k = gptr[g]; k < gptr[g + 1]; ++k) agg[gadj[k]] = na;
    ++na;
  }
  std::vector<int32_t> pass1(agg);
  std::vector<int> count(na, 0);
  for (int g = 0; g < ng; ++g) {  // attach to the neighbouring aggregate with the most links
    if (pass1[g] >= 0) continue;
    int best = -1, bc = 0;
    std::vector<std::pair<int, int>> seen;
    for (int k = gptr[g]; k < gptr[g + 1]; ++k) {
      const int a = pass1[gadj[k]];
      if (a < 0) continue;
      bool found = false;
      for (auto &p : seen)
        if (p.first == a) {
          ++p.second;
          found = true;
        }
      if (!found) seen.push_back({a, 1});
    }
    for (auto &p : seen)
      if (p.second > bc || (p.second == bc && p.first < best)) {
        best = p.first;
        bc = p.second;
      }
    agg[g] = best;
  }
  for (int g = 0; g < ng; ++g)
    if (agg[g] < 0) agg[g] = na++;  // isolated leftovers
  (void)count;
  return na;
}

// thin Householder QR of an m x 6 matrix (row-major): Q m x 6, R 6 x 6; a column whose
// R diagonal is negligible against the largest gets a zero Q column and R row
void thin_qr(std::vector<double> &M, int m, double *Q, double *R) {
  std::vector<double> V((size_t)m * NM, 0.0);
  double beta[NM] = {0};
  const int kmax = std::min(m, NM);
  for (int k = 0; k < kmax; ++k) {
    double nrm = 0.0;
    for (int i = k; i < m; ++i) nrm += M[(size_t)i * NM + k] * M[(size_t)i * NM + k];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) {
      beta[k] = 0.0;
      continue;
    }
    const double a = M[(size_t)k * NM + k];
    const double alpha = a > 0 ? -nrm : nrm;
    for (int i = k; i < m; ++i) V[(size_t)i * NM + k] = M[(size_t)i * NM + k];
    V[(size_t)k * NM + k] -= alpha;
    double vv = 0.0;
    for (int i = k; i < m; ++i) vv += V[(size_t)i * NM + k] * V[(size_t)i * NM + k];
    beta[k] = vv > 0 ? 2.0 / vv : 0.0;
    for (int j = k; j < NM; ++j) {
      double s = 0.0;
      for (int i = k; i < m; ++i) s += V[(size_t)i * NM + k] * M[(size_t)i * NM + j];
      s *= beta[k];
      for (int i = k; i < m; ++i) M[(size_t)i * NM + j] -= s * V[(size_t)i * NM + k];
    }
  }
  for (int i = 0; i < NM; ++i)
    for (int j = 0; j < NM; ++j) R[i * NM + j] = (i < m && j >= i) ? M[(size_t)i * NM + j] : 0.0;
  // Q = H_0 ... H_{k-1} applied to the first 6 unit columns
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < NM; ++j) Q[(size_t)i * NM + j] = (i == j) ? 1.0 : 0.0;
  for (int k = kmax - 1; k >= 0; --k) {
    if (beta[k] == 0.0) continue;
    for (int j = 0; j < NM; ++j) {
      double s = 0.0;
      for (int i = k; i < m; ++i) s += V[(size_t)i * NM + k] * Q[(size_t)i * NM + j];
      s *= beta[k];
      for (int i = k; i < m; ++i) Q[(size_t)i * NM + j] -= s * V[(size_t)i * NM + k];
    }
  }
  double rmax = 0.0;
  for (int k = 0; k < NM; ++k) rmax = std::max(rmax, std::fabs(R[k * NM + k]));
  for (int k = 0; k < NM; ++k)
    if (!(std::fabs(R[k * NM + k]) > 1e-8 * rmax)) {
      for (int i = 0; i < m; ++i) Q[(size_t)i * NM + k] = 0.0;
      for (int j = 0; j < NM; ++j) R[k * NM + j] = 0.0;
    }
}

// builds L's prolongator and returns the next level
bool coarsen(Hierarchy &H, int l, Level &next) {
  Level &L = H.lv[l];
  const int n = L.A.n, ng = (int)L.gstart.size() - 1, T = H.threads;
  std::vector<int32_t> gptr, gadj, agg;
  group_graph(L, gptr, gadj);
  const int na = aggregate(gptr, gadj, agg);
  if (na >= ng) return false;  // no coarsening
  L.nagg = na;
  // rows of each aggregate
  std::vector<int32_t> aptr(na + 1, 0), arow;
  for (int g = 0; g < ng; ++g) aptr[agg[g] + 1] += L.gstart[g + 1] - L.gstart[g];
  for (int a = 0; a < na; ++a) aptr[a + 1] += aptr[a];
  arow.resize(n);
  {
    std::vector<int32_t> fill(aptr.begin(), aptr.end() - 1);
    for (int g = 0; g < ng; ++g)
      for (int r = L.gstart[g]; r < L.gstart[g + 1]; ++r) arow[fill[agg[g]]++] = r;
  }
  // tentative prolongator: one 6-block per row (its aggregate); coarse near-null space R
  std::vector<double> Tval((size_t)n * NM, 0.0);
  next.B.assign((size_t)na * NM * NM, 0.0);
  std::vector<char> zero_col((size_t)na * NM, 0);
  parallel_for(na, T, [&](int, int b, int e) {
    std::vector<double> M, Q;
    double R[NM * NM];
    for (int a = b; a < e; ++a) {
      const int m = aptr[a + 1] - aptr[a];
      M.assign((size_t)m * NM, 0.0);
      Q.assign((size_t)m * NM, 0.0);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < NM; ++j) M[(size_t)i * NM + j] = L.B[(size_t)arow[aptr[a] + i] * NM + j];
      thin_qr(M, m, Q.data(), R);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < NM; ++j) Tval[(size_t)arow[aptr[a] + i] * NM + j] = Q[(size_t)i * NM + j];
      for (int i = 0; i < NM; ++i) {
        bool z = true;
        for (int j = 0; j < NM; ++j) {
          next.B[((size_t)a * NM + i) * NM + j] = R[i * NM + j];
          z = z && R[i * NM + j] == 0.0;
        }
        zero_col[(size_t)a * NM + i] = z;
      }
    }
  });
  // smoothing: P = T - omega D^-1 A T, per group with a dense accumulator over aggregates
  group_diag_inverses(L);
  L.rho = power_rho(L, 20, T);
  const double omega = 4.0 / (3.0 * L.rho);
  std::vector<std::vector<int32_t>> prow_aggs(ng);
  std::vector<std::vector<double>> prow_vals(ng);  // per group: nr x (naggs x 6)
  parallel_for(ng, T, [&](int, int b, int e) {
    std::vector<int32_t> slot(na, -1);
    std::vector<int32_t> list;
    std::vector<double> acc;  // [row a][slot][6]
    for (int g = b; g < e; ++g) {
      const int r0 = L.gstart[g], nr = L.gstart[g + 1] - r0;
      list.clear();
      list.push_back(agg[g]);
      slot[agg[g]] = 0;
      for (int a = 0; a < nr; ++a)
        for (int k = L.A.rowptr[r0 + a]; k < L.A.rowptr[r0 + a + 1]; ++k) {
          const int J = agg[L.group_of[L.A.col[k]]];
          if (slot[J] < 0) {
            slot[J] = (int)list.size();
            list.push_back(J);
          }
        }
      std::sort(list.begin(), list.end());
      for (size_t s = 0; s < list.size(); ++s) slot[list[s]] = (int)s;
      const int ns = (int)list.size();
      acc.assign((size_t)nr * ns * NM, 0.0);
      for (int a = 0; a < nr; ++a)
        for (int k = L.A.rowptr[r0 + a]; k < L.A.rowptr[r0 + a + 1]; ++k) {
          const int j = L.A.col[k];
          const int s = slot[agg[L.group_of[j]]];
          const double v = L.A.val[k];
          for (int m = 0; m < NM; ++m) acc[((size_t)a * ns + s) * NM + m] += v * Tval[(size_t)j * NM + m];
        }
      std::vector<double> &pv = prow_vals[g];
      pv.assign((size_t)nr * ns * NM, 0.0);
      const double *D = &L.dinv[(size_t)g * 36];
      for (int a = 0; a < nr; ++a)
        for (int s = 0; s < ns; ++s)
          for (int m = 0; m < NM; ++m) {
            double w = 0.0;
            for (int c = 0; c < nr; ++c) w += D[a * 6 + c] * acc[((size_t)c * ns + s) * NM + m];
            double t = (list[s] == agg[g]) ? Tval[(size_t)(r0 + a) * NM + m] : 0.0;
            pv[((size_t)a * ns + s) * NM + m] = t - omega * w;
          }
      prow_aggs[g] = list;
      for (int J : list) slot[J] = -1;
    }
  });
  // flatten P rows (row-major blocks)
  L.prow.assign(n + 1, 0);
  for (int g = 0; g < ng; ++g)
    for (int r = L.gstart[g]; r < L.gstart[g + 1]; ++r) L.prow[r + 1] = (int32_t)prow_aggs[g].size();
  for (int r = 0; r < n; ++r) L.prow[r + 1] += L.prow[r];
  L.pagg.resize(L.prow[n]);
  L.pval.resize((size_t)L.prow[n] * NM);
  for (int g = 0; g < ng; ++g) {
    const int r0 = L.gstart[g], nr = L.gstart[g + 1] - r0, ns = (int)prow_aggs[g].size();
    for (int a = 0; a < nr; ++a) {
      const int k0 = L.prow[r0 + a];
      for (int s = 0; s < ns; ++s) {
        L.pagg[k0 + s] = prow_aggs[g][s];
        for (int m = 0; m < NM; ++m)
          L.pval[(size_t)(k0 + s) * NM + m] =
              zero_col[(size_t)prow_aggs[g][s] * NM + m] ? 0.0 : prow_vals[g][((size_t)a * ns + s) * NM + m];
      }
    }
  }
  prow_aggs.clear();
  prow_vals.clear();
  // V = A P (rows of this level, 6-blocks over aggregates)
  std::vector<int32_t> vptr(n + 1, 0);
  std::vector<std::vector<int32_t>> vagg(n);
  std::vector<std::vector<double>> vval(n);
  parallel_for(n, T, [&](int, int b, int e) {
    std::vector<int32_t> slot(na, -1);
    for (int i = b; i < e; ++i) {
      std::vector<int32_t> &lst = vagg[i];
      std::vector<double> &acc = vval[i];
      for (int k = L.A.rowptr[i]; k < L.A.rowptr[i + 1]; ++k) {
        const int j = L.A.col[k];
        const double v = L.A.val[k];
        for (int q = L.prow[j]; q < L.prow[j + 1]; ++q) {
          const int J = L.pagg[q];
          if (slot[J] < 0) {
            slot[J] = (int)lst.size();
            lst.push_back(J);
            acc.resize(acc.size() + NM, 0.0);
          }
          double *dst = &acc[(size_t)slot[J] * NM];
          const double *p = &L.pval[(size_t)q * NM];
          for (int m = 0; m < NM; ++m) dst[m] += v * p[m];
        }
      }
      for (int J : lst) slot[J] = -1;
    }
  });
  // transposed P: for each aggregate I, its (row, block) entries
  std::vector<int32_t> tptr(na + 1, 0), trow, tblk;
  for (int q = 0; q < L.prow[n]; ++q) tptr[L.pagg[q] + 1]++;
  for (int a = 0; a < na; ++a) tptr[a + 1] += tptr[a];
  trow.resize(tptr[na]);
  tblk.resize(tptr[na]);
  {
    std::vector<int32_t> fill(tptr.begin(), tptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int q = L.prow[i]; q < L.prow[i + 1]; ++q) {
        const int I = L.pagg[q];
        trow[fill[I]] = i;
        tblk[fill[I]++] = q;
      }
  }
  // A_c = P^T V in 6x6 blocks per coarse aggregate row
  std::vector<std::vector<int32_t>> crow_aggs(na);
  std::vector<std::vector<double>> crow_vals(na);
  parallel_for(na, T, [&](int, int b, int e) {
    std::vector<int32_t> slot(na, -1);
    for (int I = b; I < e; ++I) {
      std::vector<int32_t> &lst = crow_aggs[I];
      std::vector<double> &acc = crow_vals[I];
      for (int t = tptr[I]; t < tptr[I + 1]; ++t) {
        const int i = trow[t];
        const double *p = &L.pval[(size_t)tblk[t] * NM];
        for (size_t s = 0; s < vagg[i].size(); ++s) {
          const int J = vagg[i][s];
          if (slot[J] < 0) {
            slot[J] = (int)lst.size();
            lst.push_back(J);
            acc.resize(acc.size() + NM * NM, 0.0);
          }
          double *dst = &acc[(size_t)slot[J] * NM * NM];
          const double *v = &vval[i][s * NM];
          for (int m = 0; m < NM; ++m)
            for (int mm = 0; mm < NM; ++mm) dst[m * NM + mm] += p[m] * v[mm];
        }
      }
      // sort the row's blocks by aggregate
      std::vector<int32_t> order(lst.size());
      for (size_t s = 0; s < order.size(); ++s) order[s] = (int32_t)s;
      std::sort(order.begin(), order.end(), [&](int x, int y) { return lst[x] < lst[y]; });
      std::vector<int32_t> l2(lst.size());
      std::vector<double> a2(acc.size());
      for (size_t s = 0; s < order.size(); ++s) {
        l2[s] = lst[order[s]];
        std::memcpy(&a2[s * NM * NM], &acc[(size_t)order[s] * NM * NM], sizeof(double) * NM * NM);
      }
      for (int J : lst) slot[J] = -1;
      lst.swap(l2);
      acc.swap(a2);
    }
  });
  vagg.clear();
  vval.clear();
  // symmetrise, pad the coarse DOFs of zero columns with an identity row
  for (int I = 0; I < na; ++I) {
    for (size_t s = 0; s < crow_aggs[I].size(); ++s) {
      const int J = crow_aggs[I][s];
      if (J < I) continue;
      double *blk = &crow_vals[I][s * NM * NM];
      if (J == I) {
        for (int m = 0; m < NM; ++m)
          for (int mm = m + 1; mm < NM; ++mm) {
            const double avg = 0.5 * (blk[m * NM + mm] + blk[mm * NM + m]);
            blk[m * NM + mm] = blk[mm * NM + m] = avg;
          }
        for (int m = 0; m < NM; ++m)
          if (zero_col[(size_t)I * NM + m]) {
            for (int mm = 0; mm < NM; ++mm) blk[m * NM + mm] = blk[mm * NM + m] = 0.0;
            blk[m * NM + m] = 1.0;
          }
      } else {
        const auto &lj = crow_aggs[J];
        const size_t sj = std::lower_bound(lj.begin(), lj.end(), I) - lj.begin();
        double *bt = &crow_vals[J][sj * NM * NM];
        for (int m = 0; m < NM; ++m)
          for (int mm = 0; mm < NM; ++mm) {
            const double avg = 0.5 * (blk[m * NM + mm] + bt[mm * NM + m]);
            blk[m * NM + mm] = bt[mm * NM + m] = avg;
          }
      }
    }
  }
  // next level: scalar CSR over 6 DOFs per aggregate, groups = aggregates
  Csr &C = next.A;
  C.n = na * NM;
  C.rowptr.assign(C.n + 1, 0);
  for (int I = 0; I < na; ++I)
    for (int m = 0; m < NM; ++m) C.rowptr[I * NM + m + 1] = (int32_t)(crow_aggs[I].size() * NM);
  for (int r = 0; r < C.n; ++r) C.rowptr[r + 1] += C.rowptr[r];
  C.col.resize(C.rowptr[C.n]);
  C.val.resize(C.rowptr[C.n]);
  for (int I = 0; I < na; ++I)
    for (int m = 0; m < NM; ++m) {
      int k = C.rowptr[I * NM + m];
      for (size_t s = 0; s < crow_aggs[I].size(); ++s)
        for (int mm = 0; mm < NM; ++mm) {
          C.col[k] = crow_aggs[I][s] * NM + mm;
          C.val[k++] = crow_vals[I][s * NM * NM + m * NM + mm];
        }
    }
  next.gstart.resize(na + 1);
  for (int I = 0; I <= na; ++I) next.gstart[I] = I * NM;
  next.group_of.resize(C.n);
  for (int r = 0; r < C.n; ++r) next.group_of[r] = r / NM;
  return true;
}

// rows 0..n-1 dealt round-robin to the threads (balances triangular loops)
template <class F>
void parallel_rows(int n, int threads, F f) {
  if (threads <= 1 || n < 64) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> ts;
  for (int t = 0; t < threads && t < n; ++t)
    ts.emplace_back([&, t] {
      for (int i = t; i < n; i += threads) f(i);
    });
  for (auto &th : ts) th.join();
}

// dense inverse of the coarsest operator (symmetric positive definite): a
// right-looking Cholesky A = L L^T in column panels of 64 (row-major lower
// triangle, every inner loop runs along a row), then W^T with W = L^-1 one
// column of W per task, then A^-1 = W^T W as dot products of rows of W^T.
bool dense_inverse(const Csr &A, std::vector<double> &inv, int threads) {
  const int n = A.n;
  const size_t N = (size_t)n;
  std::vector<double> M(N * N, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) M[i * N + A.col[k]] = A.val[k];
  constexpr int PB = 64;
  for (int J = 0; J < n; J += PB) {
    const int je = std::min(n, J + PB);
    for (int j = J; j < je; ++j) {  // the diagonal block
      double *rj = &M[j * N];
      double d = rj[j];
      for (int k = J; k < j; ++k) d -= rj[k] * rj[k];
      if (!(d > 0.0)) return false;
      rj[j] = std::sqrt(d);
      for (int i = j + 1; i < je; ++i) {
        double *ri = &M[i * N];
        double t = ri[j];
        for (int k = J; k < j; ++k) t -= ri[k] * rj[k];
        ri[j] = t / rj[j];
      }
    }
    const int rest = n - je;
    parallel_rows(rest, threads, [&](int r) {  // the panel below it, then the trailing update of row i
      const int i = je + r;
      double *ri = &M[i * N];
      for (int j = J; j < je; ++j) {
        const double *rj = &M[j * N];
        double t = ri[j];
        for (int k = J; k < j; ++k) t -= ri[k] * rj[k];
        ri[j] = t / rj[j];
      }
    });
    parallel_rows(rest, threads, [&](int r) {
      const int i = je + r;
      double *ri = &M[i * N];
      for (int j = je; j <= i; ++j) {
        const double *rj = &M[j * N];
        double t = 0.0;
        for (int k = J; k < je; ++k) t += ri[k] * rj[k];
        ri[j] -= t;
      }
    });
  }
  // W^T: row c holds column c of W = L^-1 (entries c..n-1), from L y = e_c
  std::vector<double> Wt(N * N, 0.0);
  parallel_rows(n, threads, [&](int c) {
    double *y = &Wt[c * N];
    for (int i = c; i < n; ++i) {
      const double *ri = &M[i * N];
      double t = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) t -= ri[k] * y[k];
      y[i] = t / ri[i];
    }
  });
  inv.assign(N * N, 0.0);
  parallel_rows(n, threads, [&](int a) {  // A^-1[a][b] = sum_{i >= max(a,b)} W[i][a] W[i][b]
    const double *wa = &Wt[a * N];
    for (int b = a; b < n; ++b) {
      const double *wb = &Wt[b * N];
      double t = 0.0;
      for (int i = b; i < n; ++i) t += wa[i] * wb[i];
      inv[a * N + b] = t;
    }
  });
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < a; ++b) inv[a * N + b] = inv[b * N + a];
  return true;
}

}  // namespace

struct fea_amg_host {
  Hierarchy H;
};

extern "C" int fea_amg_build(int n, const int32_t *rowptr, const int32_t *col, const double *val, int n_groups,
                             const int32_t *grp, const double *dof_xyz, const int32_t *dof_comp, int max_coarse,
                             int max_levels, int n_threads, fea_amg_host **out) {
  if (out == nullptr || n <= 0 || !rowptr || !col || !val || !grp || !dof_xyz || !dof_comp || n_groups <= 0 ||
      max_coarse < NM || max_levels < 2) {
    fea::set_error("fea_amg_build: bad arguments");
    return FEA_ERR_ARG;
  }
  if (grp[0] != 0 || grp[n_groups] != n) {
    fea::set_error("fea_amg_build: the row groups must cover rows 0 .. n-1");
    return FEA_ERR_ARG;
  }
  std::unique_ptr<fea_amg_host> h(new fea_amg_host());
  Hierarchy &H = h->H;
  H.threads = std::max(1, std::min(n_threads, 64));
  H.lv.emplace_back();
  Level &L0 = H.lv[0];
  L0.A.n = n;
  L0.A.rowptr.assign(rowptr, rowptr + n + 1);
  const int64_t nnz = rowptr[n];
  L0.A.col.assign(col, col + nnz);
  L0.A.val.assign(val, val + nnz);
  L0.gstart.assign(grp, grp + n_groups + 1);
  L0.group_of.resize(n);
  for (int g = 0; g < n_groups; ++g) {
    if (grp[g + 1] - grp[g] < 1 || grp[g + 1] - grp[g] > 3) {
      fea::set_error("fea_amg_build: row group %d has %d rows (1..3 expected)", g, grp[g + 1] - grp[g]);
      return FEA_ERR_ARG;
    }
    for (int r = grp[g]; r < grp[g + 1]; ++r) L0.group_of[r] = g;
  }
  // rigid-body modes: translations, then rotations about z, x, y (coordinates taken
  // relative to the centroid of all rows, which only changes the QR's basis)
  double c[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) c[d] += dof_xyz[3 * i + d] / n;
  L0.B.assign((size_t)n * NM, 0.0);
  for (int i = 0; i < n; ++i) {
    const int k = dof_comp[i];
    if (k < 0 || k > 2) {
      fea::set_error("fea_amg_build: dof_comp[%d] = %d", i, k);
      return FEA_ERR_ARG;
    }
    const double x = dof_xyz[3 * i] - c[0], y = dof_xyz[3 * i + 1] - c[1], z = dof_xyz[3 * i + 2] - c[2];
    double *b = &L0.B[(size_t)i * NM];
    b[k] = 1.0;
    b[3] = k == 0 ? -y : (k == 1 ? x : 0.0);
    b[4] = k == 1 ? -z : (k == 2 ? y : 0.0);
    b[5] = k == 2 ? -x : (k == 0 ? z : 0.0);
  }
  while ((int)H.lv.size() < max_levels && H.lv.back().A.n > max_coarse) {
    Level next;
    if (!coarsen(H, (int)H.lv.size() - 1, next)) break;
    H.lv.push_back(std::move(next));
  }
  Level &Lc = H.lv.back();
  if (H.lv.size() < 2) {
    fea::set_error("fea_amg_build: the operator did not coarsen");
    return FEA_ERR_STATE;
  }
  if (Lc.A.n > 8192) {
    fea::set_error("fea_amg_build: coarsest level has %d DOFs (more than 8192 for a dense inverse)", Lc.A.n);
    return FEA_ERR_STATE;
  }
  if (!dense_inverse(Lc.A, H.coarse_inv, H.threads)) {
    fea::set_error("fea_amg_build: the coarsest operator is not positive definite");
    return FEA_ERR_STATE;
  }
  for (size_t l = 1; l + 1 < H.lv.size(); ++l) {  // coarse-level smoothers (already set where coarsened)
    if (H.lv[l].dinv.empty()) group_diag_inverses(H.lv[l]);
  }
  *out = h.release();
  return FEA_OK;
}

extern "C" int fea_amg_num_levels(const fea_amg_host *h) { return h ? (int)h->H.lv.size() : 0; }

extern "C" int fea_amg_level_sizes(const fea_amg_host *h, int level, int64_t *sizes) {
  if (!h || !sizes || level < 0 || level >= (int)h->H.lv.size()) {
    fea::set_error("fea_amg_level_sizes: bad level");
    return FEA_ERR_ARG;
  }
  const Level &L = h->H.lv[level];
  const bool last = level + 1 == (int)h->H.lv.size();
  sizes[0] = L.A.n;                                    // DOFs
  sizes[1] = last ? 0 : L.nagg;                        // aggregates (next level's groups)
  sizes[2] = last ? 0 : (int64_t)L.pagg.size();        // prolongator blocks (6 values each)
  sizes[3] = (int64_t)L.A.val.size();                  // scalar nonzeros of this level's operator
  sizes[4] = (int64_t)L.gstart.size() - 1;             // row groups
  sizes[5] = last ? L.A.n : 0;                         // dense inverse order (coarsest only)
  return FEA_OK;
}

extern "C" int fea_amg_level_export(const fea_amg_host *h, int level, int32_t *prow, int32_t *pagg, float *pval,
                                    int32_t *pt_ptr, int32_t *pt_blk, int32_t *pt_row, int32_t *rowptr,
                                    int32_t *col, double *val, float *dinv36, double *rho, float *dense_inv) {
  if (!h || level < 0 || level >= (int)h->H.lv.size()) {
    fea::set_error("fea_amg_level_export: bad level");
    return FEA_ERR_ARG;
  }
  const Level &L = h->H.lv[level];
  const int n = L.A.n;
  const bool last = level + 1 == (int)h->H.lv.size();
  if (!last) {
    if (prow) std::copy(L.prow.begin(), L.prow.end(), prow);
    if (pagg) std::copy(L.pagg.begin(), L.pagg.end(), pagg);
    if (pval)
      for (size_t k = 0; k < L.pval.size(); ++k) pval[k] = (float)L.pval[k];
    if (pt_ptr || pt_blk || pt_row) {
      const int na = L.nagg;
      std::vector<int32_t> tptr(na + 1, 0);
      for (int32_t a : L.pagg) tptr[a + 1]++;
      for (int a = 0; a < na; ++a) tptr[a + 1] += tptr[a];
      if (pt_ptr) std::copy(tptr.begin(), tptr.end(), pt_ptr);
      std::vector<int32_t> fill(tptr.begin(), tptr.end() - 1);
      for (int i = 0; i < n; ++i)
        for (int q = L.prow[i]; q < L.prow[i + 1]; ++q) {
          const int t = fill[L.pagg[q]]++;
          if (pt_blk) pt_blk[t] = q;
          if (pt_row) pt_row[t] = i;
        }
    }
  }
  if (rowptr) std::copy(L.A.rowptr.begin(), L.A.rowptr.end(), rowptr);
  if (col) std::copy(L.A.col.begin(), L.A.col.end(), col);
  if (val) std::copy(L.A.val.begin(), L.A.val.end(), val);
  if (dinv36 && !L.dinv.empty())
    for (size_t k = 0; k < L.dinv.size(); ++k) dinv36[k] = (float)L.dinv[k];
  if (rho) *rho = L.rho;
  if (dense_inv && last)
    for (size_t k = 0; k < h->H.coarse_inv.size(); ++k) dense_inv[k] = (float)h->H.coarse_inv[k];
  return FEA_OK;
}

extern "C" void fea_amg_destroy(fea_amg_host *h) { delete h; }
