// The reference's Gauss-point-pair contact barrier on T3 facets (ContactSelf / Contact,
// contact.py:86-122, :300-711, :771-1207): pair search, energy, and the consistent
// gradient / Hessian staged into the CSR assembly.
//
// Per facet pair (p, q) and Gauss pair (g on p, G on q), current Gauss points y, unit
// normals n = N/|N| (N = x_xi x x_eta), reference-area weights w = |X_xi x X_eta| / 6:
//   dy = y_g - y_G, dn = n_p - n_q, D = dn.dy/2, M = n_p.n_q, L = |dy|
//   phi = ratio w_p w_q  c_g exp(k_g D)  S((s_f - M)/(s_f - e_f))  S((h_t - L)/(r_h h_t))
// with S the clamped quintic smoothstep; ratio = 1 for two surfaces, and for one surface
// S((-(n0_g - n0_G).dy0/|dy0| - ign_min)/(ign_max - ign_min)) from the initial geometry.
// Candidate pairs: facet centroids within `radius` at the current configuration, the same
// FMA-free centroid and distance arithmetic as the node-to-surface search.
//
// The gradient and Hessian are this library's own chain rule through (D, M, L) and the
// normal map (checked against central differences of the energy).  A pair's Hessian is
// built by one 64-lane workgroup: 9 lanes evaluate the Gauss pairs, the projections of
// every nodal DOF on dD, dM, dL are shared through LDS, and the 18 x 18 entries are
// spread over the lanes.  Every output has one writer: results are deterministic.
#include "common.h"

namespace fea {

namespace bar {

__device__ __forceinline__ double ad(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double mu(double a, double b) { return __dmul_rn(a, b); }

__constant__ double kNG[3][3] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},   // N_a(xi_g) [g][a]
                                 {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0},
                                 {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};
__constant__ double kDXI[3] = {-1.0, 1.0, 0.0};
__constant__ double kDETA[3] = {-1.0, 0.0, 1.0};

__device__ __forceinline__ double smooth(double t) {
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  return t * t * t * (6.0 * t * t - 15.0 * t + 10.0);
}
__device__ __forceinline__ void smooth_d(double t, double &d1, double &d2) {
  if (t > 0.0 && t < 1.0) {
    d1 = 30.0 * t * t * (t - 1.0) * (t - 1.0);
    d2 = 60.0 * t * (t - 1.0) * (2.0 * t - 1.0);
  } else {
    d1 = 0.0;
    d2 = 0.0;
  }
}

struct Facet {
  double x[3][3];   // node positions
  double y[3][3];   // Gauss points
  double u[3], v[3];
  double N[3], l, n[3];
};

__device__ __forceinline__ void load_facet(const int32_t *conn, int64_t f, const double *X, const double *U,
                                           Facet &F) {
  for (int a = 0; a < 3; ++a) {
    const int64_t nd = conn[3 * f + a];
    for (int i = 0; i < 3; ++i) F.x[a][i] = (U != nullptr) ? ad(X[3 * nd + i], U[3 * nd + i]) : X[3 * nd + i];
  }
  for (int g = 0; g < 3; ++g)
    for (int i = 0; i < 3; ++i) F.y[g][i] = kNG[g][0] * F.x[0][i] + kNG[g][1] * F.x[1][i] + kNG[g][2] * F.x[2][i];
  for (int i = 0; i < 3; ++i) {
    F.u[i] = F.x[1][i] - F.x[0][i];
    F.v[i] = F.x[2][i] - F.x[0][i];
  }
  F.N[0] = F.u[1] * F.v[2] - F.u[2] * F.v[1];
  F.N[1] = F.u[2] * F.v[0] - F.u[0] * F.v[2];
  F.N[2] = F.u[0] * F.v[1] - F.u[1] * F.v[0];
  F.l = sqrt(F.N[0] * F.N[0] + F.N[1] * F.N[1] + F.N[2] * F.N[2]);
  for (int i = 0; i < 3; ++i) F.n[i] = F.N[i] / F.l;
}

// centroid exactly as the grid build computes it: ((a + b) + c) / 3, no FMA
__device__ __forceinline__ void centroid(const int32_t *conn, int64_t f, const double *X, const double *U,
                                         double m[3]) {
  const int64_t a = conn[3 * f], b = conn[3 * f + 1], c = conn[3 * f + 2];
  for (int i = 0; i < 3; ++i)
    m[i] = ad(ad(ad(X[3 * a + i], U[3 * a + i]), ad(X[3 * b + i], U[3 * b + i])), ad(X[3 * c + i], U[3 * c + i])) / 3.0;
}
__device__ __forceinline__ double dist2(const double a[3], const double b[3]) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return ad(ad(mu(dx, dx), mu(dy, dy)), mu(dz, dz));
}
__device__ __forceinline__ bool is_candidate(const fea_barrier_params &P, const int32_t *c1, const int32_t *c2,
                                             int i, int j, const double *X, const double *U) {
  double a[3], b[3];
  centroid(c1, i, X, U, a);
  centroid(c2, j, X, U, b);
  return dist2(a, b) <= P.radius * P.radius;
}

// self-contact ratio of Gauss pair (g, G) from the initial geometry (X alone)
__device__ __forceinline__ double gp_ratio(const fea_barrier_params &P, const Facet &A0, const Facet &B0, int g,
                                           int G) {
  if (!P.self_contact) return 1.0;
  double dy[3], L2 = 0.0, T = 0.0;
  for (int i = 0; i < 3; ++i) {
    dy[i] = A0.y[g][i] - B0.y[G][i];
    L2 += dy[i] * dy[i];
  }
  const double L = sqrt(L2);
  for (int i = 0; i < 3; ++i) T -= (A0.n[i] - B0.n[i]) * dy[i] / L;
  return smooth((T - P.ign_min) / (P.ign_max - P.ign_min));
}

__device__ __forceinline__ double pair_energy(const fea_barrier_params &P, const Facet &A, const Facet &B,
                                              const Facet &A0, const Facet &B0) {
  const double w = (A0.l / 6.0) * (B0.l / 6.0);
  const double M = A.n[0] * B.n[0] + A.n[1] * B.n[1] + A.n[2] * B.n[2];
  const double f = smooth((P.s_f - M) / (P.s_f - P.e_f));
  double e = 0.0;
  if (f == 0.0) return 0.0;
  for (int g = 0; g < 3; ++g)
    for (int G = 0; G < 3; ++G) {
      double dy[3], D = 0.0, L2 = 0.0;
      for (int i = 0; i < 3; ++i) {
        dy[i] = A.y[g][i] - B.y[G][i];
        D += 0.5 * (A.n[i] - B.n[i]) * dy[i];
        L2 += dy[i] * dy[i];
      }
      const double h = smooth((P.h_t - sqrt(L2)) / (P.r_h * P.h_t));
      if (h == 0.0) continue;
      const double r = gp_ratio(P, A0, B0, g, G);
      e += r * w * P.c_g * exp(P.k_g * D) * f * h;
    }
  return e;
}

// ------------------------------------------------------------------ search
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ unsigned int cell_key(int64_t ix, int64_t iy, int64_t iz, unsigned int mask) {
  const uint64_t h = ((uint64_t)ix * 73856093ull) ^ ((uint64_t)iy * 19349663ull) ^ ((uint64_t)iz * 83492791ull);
  return (unsigned int)(h & mask);
}
__device__ __forceinline__ int64_t cell_of(double x, double o, double inv_h) {
  return (int64_t)floor((x - o) * inv_h);
}

// One wave per query facet i: lane c < 27 scans neighbour cell c (once per bucket); hits
// are target facets j (j > i for one surface) with centroid distance <= radius.  Count
// pass (offsets == nullptr) writes count[i]; fill pass writes the pairs at offsets[i] in
// (cell lane, bucket position) order, a deterministic order.
__global__ __launch_bounds__(kBlock) void k_bar_search(int nq, const int32_t *__restrict__ qconn,
                                                       const double *__restrict__ X, const double *__restrict__ U,
                                                       const double *__restrict__ cent,
                                                       const int32_t *__restrict__ bstart,
                                                       const int32_t *__restrict__ bitems, double ox, double oy,
                                                       double oz, double inv_h, unsigned int mask, double r2,
                                                       int self, const int32_t *__restrict__ offsets,
                                                       int32_t *__restrict__ count, int32_t *__restrict__ pi,
                                                       int32_t *__restrict__ pj) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * kWaves + wave;
  if (i >= nq) return;
  double c[3];
  centroid(qconn, i, X, U, c);
  const int64_t cx = cell_of(c[0], ox, inv_h), cy = cell_of(c[1], oy, inv_h), cz = cell_of(c[2], oz, inv_h);
  const bool active = lane < 27;
  unsigned int k = 0xffffffffu;
  if (active) k = cell_key(cx + (lane % 3) - 1, cy + (lane / 3) % 3 - 1, cz + lane / 9 - 1, mask);
  bool dup = false;
  for (int j = 0; j < 27; ++j) {
    const unsigned int kj = __shfl(k, j);
    dup = dup || (j < lane && kj == k);
  }
  int b0 = 0, b1 = 0;
  if (active && !dup) {
    b0 = bstart[k];
    b1 = bstart[k + 1];
  }
  auto hit = [&](int j) {
    if (self && j <= i) return false;
    const double *q = cent + 3 * (int64_t)j;
    return dist2(c, q) <= r2;
  };
  int cnt = 0;
  for (int it = b0; it < b1; ++it) cnt += hit(bitems[it]) ? 1 : 0;
  int incl = cnt;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  if (offsets == nullptr) {
    if (lane == 63) count[i] = incl;
    return;
  }
  int pos = offsets[i] + incl - cnt;
  for (int it = b0; it < b1; ++it) {
    const int j = bitems[it];
    if (hit(j)) {
      pi[pos] = i;
      pj[pos] = j;
      ++pos;
    }
  }
}

// flag[p] = 1 when some Gauss pair of pair p has a nonzero self-contact ratio (always 1
// for two surfaces): the pairs ContactSelf keeps (contact.py:336-340)
__global__ __launch_bounds__(kBlock) void k_bar_active(fea_barrier_params P, int np_, const int32_t *__restrict__ pi,
                                                       const int32_t *__restrict__ pj,
                                                       const int32_t *__restrict__ c1,
                                                       const int32_t *__restrict__ c2,
                                                       const double *__restrict__ X, int32_t *__restrict__ flag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np_) return;
  if (!P.self_contact) {
    flag[p] = 1;
    return;
  }
  Facet A0, B0;
  load_facet(c1, pi[p], X, nullptr, A0);
  load_facet(c2, pj[p], X, nullptr, B0);
  int any = 0;
  for (int g = 0; g < 3; ++g)
    for (int G = 0; G < 3; ++G) any |= gp_ratio(P, A0, B0, g, G) > 0.0 ? 1 : 0;
  flag[p] = any;
}

// count of candidate pairs with a nonzero ratio that are missing from the sorted key list
__global__ __launch_bounds__(kBlock) void k_bar_missing(fea_barrier_params P, int np_,
                                                        const int32_t *__restrict__ pi,
                                                        const int32_t *__restrict__ pj,
                                                        const int32_t *__restrict__ c1,
                                                        const int32_t *__restrict__ c2,
                                                        const double *__restrict__ X,
                                                        const int64_t *__restrict__ keys, int nkeys,
                                                        int64_t stride, int32_t *__restrict__ missing) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np_) return;
  const int64_t key = (int64_t)pi[p] * stride + pj[p];
  int lo = 0, hi = nkeys;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  if (lo < nkeys && keys[lo] == key) return;
  if (P.self_contact) {
    Facet A0, B0;
    load_facet(c1, pi[p], X, nullptr, A0);
    load_facet(c2, pj[p], X, nullptr, B0);
    bool any = false;
    for (int g = 0; g < 3; ++g)
      for (int G = 0; G < 3; ++G) any = any || gp_ratio(P, A0, B0, g, G) > 0.0;
    if (!any) return;
  }
  atomicAdd(missing, 1);
}

__global__ __launch_bounds__(kBlock) void k_bar_energy(fea_barrier_params P, int np_, const int32_t *__restrict__ pi,
                                                       const int32_t *__restrict__ pj,
                                                       const int32_t *__restrict__ c1,
                                                       const int32_t *__restrict__ c2,
                                                       const double *__restrict__ X, const double *__restrict__ U,
                                                       double *__restrict__ partial) {
  __shared__ double red[kBlock];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (p < np_) {
    Facet A, B, A0, B0;
    load_facet(c1, pi[p], X, U, A);
    load_facet(c2, pj[p], X, U, B);
    load_facet(c1, pi[p], X, nullptr, A0);
    load_facet(c2, pj[p], X, nullptr, B0);
    e = pair_energy(P, A, B, A0, B0);
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------ gradient / Hessian
// column x of a pair's 18 DOFs: side s = x / 9 (0: facet p, 1: facet q), node a = (x % 9) / 3,
// component al = x % 3
__global__ __launch_bounds__(64) void k_bar_assemble_wg(fea_barrier_params P, int np_,
                                                        const int32_t *__restrict__ pi,
                                                        const int32_t *__restrict__ pj,
                                                        const int32_t *__restrict__ c1,
                                                        const int32_t *__restrict__ c2,
                                                        const double *__restrict__ X, const double *__restrict__ U,
                                                        const int32_t *__restrict__ dst_k, void *__restrict__ kstage,
                                                        int kprec, const int32_t *__restrict__ dst_r,
                                                        double *__restrict__ rstage) {
  __shared__ double gpc[9][16];        // c_DD c_MM c_LL c_DM c_DL c_ML c_D c_M c_L dy[3] dhat[3] invL
  __shared__ double proj[9][18][3];    // pD pM pL
  __shared__ double dN[2][3][9], Jn[2][3][9];
  __shared__ double grad[18], bvec[2][3], cvec[2][3], B2[2][9];
  const int p = blockIdx.x;
  const int lane = threadIdx.x;
  if (p >= np_) return;
  const bool cand = is_candidate(P, c1, c2, pi[p], pj[p], X, U);
  const int64_t rbase = (int64_t)p * 6, kbase = (int64_t)p * 36;
  if (!cand) {
    // not a candidate at this configuration: its staged rows and blocks are zero
    for (int t = lane; t < 36 * 9; t += 64) {
      const int d = dst_k != nullptr ? dst_k[kbase + t / 9] : -1;
      if (d >= 0) {
        if (kprec == FEA_F64) ((double *)kstage)[(int64_t)d * 9 + t % 9] = 0.0;
        else ((float *)kstage)[(int64_t)d * 9 + t % 9] = 0.0f;
      }
    }
    if (lane < 18 && dst_r != nullptr) {
      const int d = dst_r[rbase + lane / 3];
      if (d >= 0) rstage[(int64_t)d * 3 + lane % 3] = 0.0;
    }
    return;
  }
  Facet S[2];
  load_facet(c1, pi[p], X, U, S[0]);
  load_facet(c2, pj[p], X, U, S[1]);
  const double M = S[0].n[0] * S[1].n[0] + S[0].n[1] * S[1].n[1] + S[0].n[2] * S[1].n[2];
  const double sf = P.s_f - P.e_f;
  const double tm = (P.s_f - M) / sf;
  const double f = smooth(tm);
  double f1, f2;
  smooth_d(tm, f1, f2);
  f1 = -f1 / sf;
  f2 = f2 / (sf * sf);
  const double hs = P.r_h * P.h_t;
  if (lane < 9) {
    const int g = lane / 3, G = lane % 3;
    Facet A0, B0;
    load_facet(c1, pi[p], X, nullptr, A0);
    load_facet(c2, pj[p], X, nullptr, B0);
    const double K = gp_ratio(P, A0, B0, g, G) * (A0.l / 6.0) * (B0.l / 6.0);
    double dy[3], D = 0.0, L2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      dy[i] = S[0].y[g][i] - S[1].y[G][i];
      D += 0.5 * (S[0].n[i] - S[1].n[i]) * dy[i];
      L2 += dy[i] * dy[i];
    }
    const double L = sqrt(L2);
    const double th = (P.h_t - L) / hs;
    const double h = smooth(th);
    double h1, h2;
    smooth_d(th, h1, h2);
    h1 = -h1 / hs;
    h2 = h2 / (hs * hs);
    const double gg = P.c_g * exp(P.k_g * D), g1 = P.k_g * gg, g2 = P.k_g * P.k_g * gg;
    double *c = gpc[lane];
    c[0] = K * g2 * f * h;
    c[1] = K * gg * f2 * h;
    c[2] = K * gg * f * h2;
    c[3] = K * g1 * f1 * h;
    c[4] = K * g1 * f * h1;
    c[5] = K * gg * f1 * h1;
    c[6] = K * g1 * f * h;
    c[7] = K * gg * f1 * h;
    c[8] = K * gg * f * h1;
    const double invL = L > 0.0 ? 1.0 / L : 0.0;
    for (int i = 0; i < 3; ++i) {
      c[9 + i] = dy[i];
      c[12 + i] = dy[i] * invL;
    }
    c[15] = invL;
  } else if (lane < 27) {
    // dN/dx and dn/dx columns: dN = dxi_a (e_al x v) + deta_a (u x e_al), dn = (I - n n^T) dN / l
    const int x = lane - 9;
    const int s = x / 9, a = (x % 9) / 3, al = x % 3;
    const Facet &F = S[s];
    double e[3] = {0.0, 0.0, 0.0};
    e[al] = 1.0;
    const double ev[3] = {e[1] * F.v[2] - e[2] * F.v[1], e[2] * F.v[0] - e[0] * F.v[2], e[0] * F.v[1] - e[1] * F.v[0]};
    const double ue[3] = {F.u[1] * e[2] - F.u[2] * e[1], F.u[2] * e[0] - F.u[0] * e[2], F.u[0] * e[1] - F.u[1] * e[0]};
    double d[3], nd = 0.0;
    for (int i = 0; i < 3; ++i) {
      d[i] = kDXI[a] * ev[i] + kDETA[a] * ue[i];
      nd += F.n[i] * d[i];
    }
    for (int i = 0; i < 3; ++i) {
      dN[s][i][x % 9] = d[i];
      Jn[s][i][x % 9] = (d[i] - F.n[i] * nd) / F.l;
    }
  }
  __syncthreads();
  for (int t = lane; t < 9 * 18; t += 64) {
    const int gp = t / 18, x = t % 18;
    const int g = gp / 3, G = gp % 3;
    const int s = x / 9, a = (x % 9) / 3, al = x % 3;
    const double sig = s == 0 ? kNG[g][a] : -kNG[G][a];
    const double *c = gpc[gp];
    double jdy = 0.0, pM = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double jn = s == 0 ? Jn[0][i][x % 9] : -Jn[1][i][x % 9];
      jdy += c[9 + i] * jn;
      pM += (s == 0 ? S[1].n[i] * Jn[0][i][x % 9] : S[0].n[i] * Jn[1][i][x % 9]);
    }
    const double dn = S[0].n[al] - S[1].n[al];
    proj[gp][x][0] = 0.5 * (dn * sig + jdy);
    proj[gp][x][1] = pM;
    proj[gp][x][2] = c[12 + al] * sig;
  }
  __syncthreads();
  if (lane < 18) {
    double acc = 0.0;
    for (int gp = 0; gp < 9; ++gp)
      acc += gpc[gp][6] * proj[gp][lane][0] + gpc[gp][7] * proj[gp][lane][1] + gpc[gp][8] * proj[gp][lane][2];
    grad[lane] = acc;
  } else if (lane < 24) {
    // total gradient in n_p (side 0) and n_q (side 1): sum of c_D (+-dy/2) + c_M n_other
    const int s = (lane - 18) / 3, i = (lane - 18) % 3;
    double acc = 0.0;
    for (int gp = 0; gp < 9; ++gp)
      acc += (s == 0 ? 0.5 : -0.5) * gpc[gp][6] * gpc[gp][9 + i] + gpc[gp][7] * S[1 - s].n[i];
    bvec[s][i] = acc;
  }
  __syncthreads();
  if (lane < 2) {
    // c = (I - n n^T) b / l ; B2_jk = [-(b_k n_j + b_j n_k + (b.n) d_jk) + 3 (b.n) n_j n_k] / l^2
    const Facet &F = S[lane];
    const double *b = bvec[lane];
    const double bn = b[0] * F.n[0] + b[1] * F.n[1] + b[2] * F.n[2];
    for (int i = 0; i < 3; ++i) cvec[lane][i] = (b[i] - F.n[i] * bn) / F.l;
    const double il2 = 1.0 / (F.l * F.l);
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        B2[lane][j * 3 + k] =
            (-(b[k] * F.n[j] + b[j] * F.n[k] + (j == k ? bn : 0.0)) + 3.0 * bn * F.n[j] * F.n[k]) * il2;
  }
  __syncthreads();
  double cMt = 0.0;
  for (int gp = 0; gp < 9; ++gp) cMt += gpc[gp][7];
  for (int t = lane; t < 18 * 18; t += 64) {
    const int x = t / 18, y = t % 18;
    const int sx = x / 9, ax = (x % 9) / 3, alx = x % 3;
    const int sy = y / 9, ay = (y % 9) / 3, aly = y % 3;
    double H = 0.0;
    for (int gp = 0; gp < 9; ++gp) {
      const int g = gp / 3, G = gp % 3;
      const double *c = gpc[gp];
      const double *px = proj[gp][x], *py = proj[gp][y];
      H += c[0] * px[0] * py[0] + c[1] * px[1] * py[1] + c[2] * px[2] * py[2] +
           c[3] * (px[0] * py[1] + px[1] * py[0]) + c[4] * (px[0] * py[2] + px[2] * py[0]) +
           c[5] * (px[1] * py[2] + px[2] * py[1]);
      const double sgx = sx == 0 ? kNG[g][ax] : -kNG[G][ax];
      const double sgy = sy == 0 ? kNG[g][ay] : -kNG[G][ay];
      const double jny = sy == 0 ? Jn[0][alx][y % 9] : -Jn[1][alx][y % 9];
      const double jnx = sx == 0 ? Jn[0][aly][x % 9] : -Jn[1][aly][x % 9];
      H += c[6] * 0.5 * (sgx * jny + sgy * jnx);                                    // c_D d2D
      H += c[8] * sgx * sgy * ((alx == aly ? 1.0 : 0.0) - c[12 + alx] * c[12 + aly]) * c[15];   // c_L d2L
    }
    if (sx != sy) {                                                                  // c_M d2M
      const int xp = sx == 0 ? x % 9 : y % 9, yq = sx == 0 ? y % 9 : x % 9;
      double dot = 0.0;
      for (int i = 0; i < 3; ++i) dot += Jn[0][i][xp] * Jn[1][i][yq];
      H += cMt * dot;
    } else {                                                                         // d2n terms
      const int s = sx, xx = x % 9, yy = y % 9;
      double q = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) q += dN[s][j][xx] * B2[s][j * 3 + k] * dN[s][k][yy];
      const double coef = kDXI[ax] * kDETA[ay] - kDETA[ax] * kDXI[ay];
      if (coef != 0.0 && alx != aly) {
        const int gam = 3 - alx - aly;                                              // the third index
        const double eps = ((alx + 1) % 3 == aly) ? 1.0 : -1.0;                     // eps_{gam alx aly}
        q += coef * eps * cvec[s][gam];
      }
      H += q;
    }
    if (dst_k != nullptr) {
      const int d = dst_k[kbase + (x / 3) * 6 + (y / 3)];
      if (d >= 0) {
        const int64_t o = (int64_t)d * 9 + alx * 3 + aly;
        if (kprec == FEA_F64) ((double *)kstage)[o] = H;
        else ((float *)kstage)[o] = (float)H;
      }
    }
  }
  if (lane < 18 && dst_r != nullptr) {
    const int d = dst_r[rbase + lane / 3];
    if (d >= 0) rstage[(int64_t)d * 3 + lane % 3] = grad[lane];
  }
}

// exclusive scan of n counts into out[0..n], one workgroup of 1024 lanes over chunks
__global__ __launch_bounds__(1024) void k_scan_i32(int n, const int32_t *__restrict__ in, int32_t *__restrict__ out) {
  scan_1024(n, in, out);
}

}  // namespace bar
}  // namespace fea

using namespace fea;
using namespace fea::bar;

extern "C" {

int fea_scan_i32(int n, const int32_t *in, int32_t *out, void *stream) {
  FEA_CHECK_ARG(n >= 0 && out && (n == 0 || in), "fea_scan_i32: bad arguments");
  k_scan_i32<<<1, 1024, 0, as_stream(stream)>>>(n, in, out);
  FEA_CHECK_LAUNCH("fea_scan_i32");
  return FEA_OK;
}

int fea_barrier_search(int n_query, const int32_t *qconn, const double *X, const double *U, const double *centroid,
                       const int32_t *bucket_start, const int32_t *bucket_items, const double *origin_host,
                       double cell, int table_bits, double radius, int self_contact, const int32_t *offsets,
                       int32_t *count, int32_t *pair_i, int32_t *pair_j, void *stream) {
  FEA_CHECK_ARG(n_query >= 0 && origin_host && cell > 0.0 && radius > 0.0 && radius <= cell && table_bits >= 1 &&
                    table_bits <= 28,
                "fea_barrier_search: bad arguments (radius must not exceed the cell size)");
  if (n_query == 0) return FEA_OK;
  FEA_CHECK_ARG(qconn && X && U && centroid && bucket_start && bucket_items, "fea_barrier_search: null input");
  FEA_CHECK_ARG(offsets != nullptr ? (pair_i && pair_j) : (count != nullptr), "fea_barrier_search: null output");
  const unsigned int nb = 1u << table_bits;
  k_bar_search<<<(n_query + kWaves - 1) / kWaves, kBlock, 0, as_stream(stream)>>>(
      n_query, qconn, X, U, centroid, bucket_start, bucket_items, origin_host[0], origin_host[1], origin_host[2],
      1.0 / cell, nb - 1, radius * radius, self_contact, offsets, count, pair_i, pair_j);
  FEA_CHECK_LAUNCH("fea_barrier_search");
  return FEA_OK;
}

int fea_barrier_active(const fea_barrier_params *P, int n_pairs, const int32_t *pair_i, const int32_t *pair_j,
                       const int32_t *conn1, const int32_t *conn2, const double *X, int32_t *flag, void *stream) {
  FEA_CHECK_ARG(P && n_pairs >= 0, "fea_barrier_active: bad arguments");
  if (n_pairs == 0) return FEA_OK;
  FEA_CHECK_ARG(pair_i && pair_j && conn1 && conn2 && X && flag, "fea_barrier_active: null argument");
  k_bar_active<<<blocks_for(n_pairs), kBlock, 0, as_stream(stream)>>>(*P, n_pairs, pair_i, pair_j, conn1, conn2, X,
                                                                     flag);
  FEA_CHECK_LAUNCH("fea_barrier_active");
  return FEA_OK;
}

int fea_barrier_missing(const fea_barrier_params *P, int n_pairs, const int32_t *pair_i, const int32_t *pair_j,
                        const int32_t *conn1, const int32_t *conn2, const double *X, const int64_t *keys,
                        int n_keys, int64_t stride, int32_t *missing, void *stream) {
  FEA_CHECK_ARG(P && n_pairs >= 0 && n_keys >= 0 && stride > 0 && missing, "fea_barrier_missing: bad arguments");
  if (n_pairs == 0) return FEA_OK;
  FEA_CHECK_ARG(pair_i && pair_j && conn1 && conn2 && X && (n_keys == 0 || keys), "fea_barrier_missing: null argument");
  k_bar_missing<<<blocks_for(n_pairs), kBlock, 0, as_stream(stream)>>>(*P, n_pairs, pair_i, pair_j, conn1, conn2, X,
                                                                      keys, n_keys, stride, missing);
  FEA_CHECK_LAUNCH("fea_barrier_missing");
  return FEA_OK;
}

int fea_barrier_energy_blocks(int n_pairs) { return n_pairs > 0 ? blocks_for(n_pairs) : 0; }

int fea_barrier_energy(const fea_barrier_params *P, int n_pairs, const int32_t *pair_i, const int32_t *pair_j,
                       const int32_t *conn1, const int32_t *conn2, const double *X, const double *U,
                       double *partial, void *stream) {
  FEA_CHECK_ARG(P && n_pairs >= 0, "fea_barrier_energy: bad arguments");
  if (n_pairs == 0) return FEA_OK;
  FEA_CHECK_ARG(pair_i && pair_j && conn1 && conn2 && X && U && partial, "fea_barrier_energy: null argument");
  k_bar_energy<<<blocks_for(n_pairs), kBlock, 0, as_stream(stream)>>>(*P, n_pairs, pair_i, pair_j, conn1, conn2, X, U,
                                                                     partial);
  FEA_CHECK_LAUNCH("fea_barrier_energy");
  return FEA_OK;
}

int fea_barrier_assemble(const fea_barrier_params *P, int stage_prec, int n_pairs, const int32_t *pair_i,
                         const int32_t *pair_j, const int32_t *conn1, const int32_t *conn2, const double *X,
                         const double *U, const int32_t *dst_k, void *kstage, const int32_t *dst_r, double *rstage,
                         void *stream) {
  FEA_CHECK_ARG(P && n_pairs >= 0 && (stage_prec == FEA_F32 || stage_prec == FEA_F64),
                "fea_barrier_assemble: bad arguments");
  FEA_CHECK_ARG((dst_k == nullptr) == (kstage == nullptr) && (dst_r == nullptr) == (rstage == nullptr),
                "fea_barrier_assemble: staging map / buffer mismatch");
  if (n_pairs == 0) return FEA_OK;
  FEA_CHECK_ARG(pair_i && pair_j && conn1 && conn2 && X && U, "fea_barrier_assemble: null argument");
  k_bar_assemble_wg<<<n_pairs, 64, 0, as_stream(stream)>>>(*P, n_pairs, pair_i, pair_j, conn1, conn2, X, U, dst_k,
                                                           kstage, stage_prec, dst_r, rstage);
  FEA_CHECK_LAUNCH("fea_barrier_assemble");
  return FEA_OK;
}

}  // extern "C"
