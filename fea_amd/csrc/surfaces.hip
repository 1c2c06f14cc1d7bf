// Surface-facet kernels: follower pressure (force, stiffness, enclosed volume)
// and dead-load traction on T3 / T6 / Q4 facets.
#include "common.h"

namespace fea {

// Reference surface shape functions / Gauss rules:
//   T3 triangle.py:22-40, T6 triangle.py:67-86, Q4 quadrilateral.py:22-40.
template <int ST>
struct Surf;

template <>
struct Surf<3> {
  static constexpr int nen = 3, ngp = 3;
  __device__ static void point(int g, double &x, double &y) {
    x = (g == 1) ? 2.0 / 3.0 : 1.0 / 6.0;
    y = (g == 2) ? 2.0 / 3.0 : 1.0 / 6.0;
  }
  __device__ static constexpr double weight(int) { return 1.0 / 6.0; }
  __device__ static void eval(int g, double N[3], double dN[2][3]) {
    double x, y;
    point(g, x, y);
    N[0] = 1.0 - x - y; N[1] = x; N[2] = y;
    dN[0][0] = -1.0; dN[0][1] = 1.0; dN[0][2] = 0.0;
    dN[1][0] = -1.0; dN[1][1] = 0.0; dN[1][2] = 1.0;
  }
};

template <>
struct Surf<6> {
  static constexpr int nen = 6, ngp = 3;
  __device__ static constexpr double weight(int) { return 1.0 / 6.0; }
  __device__ static void eval(int g, double N[6], double dN[2][6]) {
    double x, y;
    Surf<3>::point(g, x, y);
    const double L = 1.0 - x - y;
    N[0] = L * (1.0 - 2.0 * x - 2.0 * y);
    N[1] = x * (2.0 * x - 1.0);
    N[2] = y * (2.0 * y - 1.0);
    N[3] = 4.0 * x * L;
    N[4] = 4.0 * x * y;
    N[5] = 4.0 * y * L;
    dN[0][0] = -3.0 + 4.0 * x + 4.0 * y; dN[1][0] = -3.0 + 4.0 * x + 4.0 * y;
    dN[0][1] = 4.0 * x - 1.0;            dN[1][1] = 0.0;
    dN[0][2] = 0.0;                      dN[1][2] = 4.0 * y - 1.0;
    dN[0][3] = 4.0 - 8.0 * x - 4.0 * y;  dN[1][3] = -4.0 * x;
    dN[0][4] = 4.0 * y;                  dN[1][4] = 4.0 * x;
    dN[0][5] = -4.0 * y;                 dN[1][5] = 4.0 - 4.0 * x - 8.0 * y;
  }
};

template <>
struct Surf<4> {
  static constexpr int nen = 4, ngp = 4;
  static constexpr double q = 0.57735026918962576451;
  __device__ static constexpr double weight(int) { return 1.0; }
  __device__ static double sg(int a) { return (a == 1 || a == 2) ? 1.0 : -1.0; }
  __device__ static double sh(int a) { return (a >= 2) ? 1.0 : -1.0; }
  __device__ static void eval(int g, double N[4], double dN[2][4]) {
    const double x = q * sg(g), y = q * sh(g);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      N[a] = 0.25 * (1.0 + x * sg(a)) * (1.0 + y * sh(a));
      dN[0][a] = 0.25 * sg(a) * (1.0 + y * sh(a));
      dN[1][a] = 0.25 * sh(a) * (1.0 + x * sg(a));
    }
  }
};

// ra[i*3+m] = [x, x_xi, x_eta]_i at Gauss point g (pressure.py:95-105)
template <typename T, int ST>
__device__ __forceinline__ void ra_at(int g, const T pos[][3], T ra[9], double N[], double dN[][Surf<ST>::nen]) {
  constexpr int NEN = Surf<ST>::nen;
  Surf<ST>::eval(g, N, dN);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T a0 = T(0), a1 = T(0), a2 = T(0);
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
      a0 += T(N[c]) * pos[c][i];
      a1 += T(dN[0][c]) * pos[c][i];
      a2 += T(dN[1][c]) * pos[c][i];
    }
    ra[i * 3 + 0] = a0;
    ra[i * 3 + 1] = a1;
    ra[i * 3 + 2] = a2;
  }
}

// Derivatives of det A, A = [x, x_xi, x_eta] = sum_c pos_c (x) s_c with
// s_c = (N_c, dN_c/dxi, dN_c/deta) (pressure.py:95-127 forms them as
// det A * A^-T and det A (A^-1 A^-1 - A^-1 A^-1); those need A invertible and
// fail for a facet whose plane contains the origin).  Here, without division:
//   d det / d pos_a,i           = (cof A . s_a)_i,  cof A row i = r_{i+1} x r_{i+2};
//   d2 det / d pos_a,i d pos_b,j = eps_ijk (A (s_a x s_b))_k.
// They equal the reference's expressions wherever A is invertible.
template <typename T>
__device__ __forceinline__ void cofactor3(const T ra[9], T cof[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    cof[i * 3 + 0] = ra[j * 3 + 1] * ra[k * 3 + 2] - ra[j * 3 + 2] * ra[k * 3 + 1];
    cof[i * 3 + 1] = ra[j * 3 + 2] * ra[k * 3 + 0] - ra[j * 3 + 0] * ra[k * 3 + 2];
    cof[i * 3 + 2] = ra[j * 3 + 0] * ra[k * 3 + 1] - ra[j * 3 + 1] * ra[k * 3 + 0];
  }
}

// v = A (s_a x s_b)
template <typename T>
__device__ __forceinline__ void hess_axis(const T ra[9], const T sa[3], const T sb[3], T v[3]) {
  const T c0 = sa[1] * sb[2] - sa[2] * sb[1];
  const T c1 = sa[2] * sb[0] - sa[0] * sb[2];
  const T c2 = sa[0] * sb[1] - sa[1] * sb[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = ra[k * 3 + 0] * c0 + ra[k * 3 + 1] * c1 + ra[k * 3 + 2] * c2;
}

// One thread per (facet, node a): scale * dV/dx_a and scale * d2V/dx_a dx_b.
template <typename T, int ST>
__global__ __launch_bounds__(kBlock) void k_pressure(int n_facets, const int32_t *__restrict__ conn,
                                                     const double *__restrict__ X,
                                                     const double *__restrict__ U, double scale,
                                                     const int32_t *__restrict__ dst_k,
                                                     T *__restrict__ kstage,
                                                     const int32_t *__restrict__ dst_r,
                                                     double *__restrict__ rstage) {
  constexpr int NEN = Surf<ST>::nen, NGP = Surf<ST>::ngp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_facets * NEN) return;
  const int e = (int)(t / NEN), a = (int)(t - (int64_t)e * NEN);
  T pos[NEN][3];
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
    for (int i = 0; i < 3; ++i) pos[c][i] = (T)(X[3 * n + i] + U[3 * n + i]);
  }
  double f[3] = {0.0, 0.0, 0.0};
  T K[NEN][9];
#pragma unroll
  for (int b = 0; b < NEN; ++b)
#pragma unroll
    for (int k = 0; k < 9; ++k) K[b][k] = T(0);
  for (int g = 0; g < NGP; ++g) {
    double N[NEN], dN[2][NEN];
    T ra[9];
    ra_at<T, ST>(g, pos, ra, N, dN);
    T cof[9];
    cofactor3(ra, cof);
    const T ws = T(Surf<ST>::weight(g) / 3.0 * scale);
    const T sa[3] = {T(N[a]), T(dN[0][a]), T(dN[1][a])};
#pragma unroll
    for (int i = 0; i < 3; ++i)
      f[i] += (double)(ws * (cof[i * 3 + 0] * sa[0] + cof[i * 3 + 1] * sa[1] + cof[i * 3 + 2] * sa[2]));
    if (kstage != nullptr) {
#pragma unroll
      for (int b = 0; b < NEN; ++b) {
        const T sb[3] = {T(N[b]), T(dN[0][b]), T(dN[1][b])};
        T v[3];
        hess_axis(ra, sa, sb, v);
        // eps_ijk v_k
        K[b][1] += ws * v[2];
        K[b][2] -= ws * v[1];
        K[b][3] -= ws * v[2];
        K[b][5] += ws * v[0];
        K[b][6] += ws * v[1];
        K[b][7] -= ws * v[0];
      }
    }
  }
  if (kstage != nullptr) {
    const int64_t base = ((int64_t)e * NEN + a) * NEN;
#pragma unroll
    for (int b = 0; b < NEN; ++b) {
      const int d = dst_k[base + b];
      if (d >= 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) kstage[(int64_t)d * 9 + k] = K[b][k];
      }
    }
  }
  if (rstage != nullptr) {
    const int d = dst_r[(int64_t)e * NEN + a];
    if (d >= 0) {
      rstage[(int64_t)d * 3 + 0] = f[0];
      rstage[(int64_t)d * 3 + 1] = f[1];
      rstage[(int64_t)d * 3 + 2] = f[2];
    }
  }
}

// Matrix-free pressure tangent: out_a = sum_b K_ab lam_b with K_ab the Hessian
// block of scale * V (the k_pressure stiffness), one thread per (facet, a):
//   K_ab lam_b = scale w/3 sum_g lam_b x (A (s_a x s_b)).
template <int ST>
__global__ __launch_bounds__(kBlock) void k_pressure_hvp(int n_facets, const int32_t *__restrict__ conn,
                                                         const double *__restrict__ X,
                                                         const double *__restrict__ U, double scale,
                                                         const double *__restrict__ lam,
                                                         const int32_t *__restrict__ dst,
                                                         double *__restrict__ stage) {
  constexpr int NEN = Surf<ST>::nen, NGP = Surf<ST>::ngp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_facets * NEN) return;
  const int e = (int)(t / NEN), a = (int)(t - (int64_t)e * NEN);
  double pos[NEN][3], l[NEN][3];
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      pos[c][i] = X[3 * n + i] + U[3 * n + i];
      l[c][i] = lam[3 * n + i];
    }
  }
  double out[3] = {0.0, 0.0, 0.0};
  for (int g = 0; g < NGP; ++g) {
    double N[NEN], dN[2][NEN], ra[9];
    ra_at<double, ST>(g, pos, ra, N, dN);
    const double ws = Surf<ST>::weight(g) / 3.0 * scale;
    const double sa[3] = {N[a], dN[0][a], dN[1][a]};
#pragma unroll
    for (int b = 0; b < NEN; ++b) {
      const double sb[3] = {N[b], dN[0][b], dN[1][b]};
      double v[3];
      hess_axis(ra, sa, sb, v);
      // (eps_ijk v_k) lam_j = (lam x v)_i
      out[0] += ws * (l[b][1] * v[2] - l[b][2] * v[1]);
      out[1] += ws * (l[b][2] * v[0] - l[b][0] * v[2]);
      out[2] += ws * (l[b][0] * v[1] - l[b][1] * v[0]);
    }
  }
  const int64_t d = dst[(int64_t)e * NEN + a];
  if (d >= 0) {
    stage[d * 3 + 0] = out[0];
    stage[d * 3 + 1] = out[1];
    stage[d * 3 + 2] = out[2];
  }
}

template <int ST>
__global__ __launch_bounds__(kBlock) void k_volume(int n_facets, const int32_t *__restrict__ conn,
                                                   const double *__restrict__ X,
                                                   const double *__restrict__ U,
                                                   double *__restrict__ partial) {
  constexpr int NEN = Surf<ST>::nen, NGP = Surf<ST>::ngp;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (e < n_facets) {
    double pos[NEN][3];
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
      const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
      for (int i = 0; i < 3; ++i) pos[c][i] = X[3 * n + i] + U[3 * n + i];
    }
    for (int g = 0; g < NGP; ++g) {
      double N[NEN], dN[2][NEN], ra[9];
      ra_at<double, ST>(g, pos, ra, N, dN);
      v += Surf<ST>::weight(g) * det3(ra) / 3.0;
    }
  }
  const double s = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <int ST>
__global__ __launch_bounds__(kBlock) void k_traction(int n_facets, const int32_t *__restrict__ conn,
                                                     const double *__restrict__ X,
                                                     const double *__restrict__ U, double tx,
                                                     double ty, double tz, double scale,
                                                     const int32_t *__restrict__ dst_r,
                                                     double *__restrict__ rstage,
                                                     double *__restrict__ work_partial) {
  constexpr int NEN = Surf<ST>::nen, NGP = Surf<ST>::ngp;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double work = 0.0;
  if (e < n_facets) {
    double x[NEN][3];
    int64_t nd[NEN];
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
      nd[c] = conn[(int64_t)e * NEN + c];
#pragma unroll
      for (int i = 0; i < 3; ++i) x[c][i] = X[3 * nd[c] + i];
    }
    double fa[NEN];
#pragma unroll
    for (int c = 0; c < NEN; ++c) fa[c] = 0.0;
    for (int g = 0; g < NGP; ++g) {
      double N[NEN], dN[2][NEN];
      Surf<ST>::eval(g, N, dN);
      double j0[3], j1[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        j0[i] = 0.0;
        j1[i] = 0.0;
#pragma unroll
        for (int c = 0; c < NEN; ++c) {
          j0[i] += dN[0][c] * x[c][i];
          j1[i] += dN[1][c] * x[c][i];
        }
      }
      // sqrt(det(J^T J)) on the undeformed facet (basesurface.py:144-150)
      const double g00 = j0[0] * j0[0] + j0[1] * j0[1] + j0[2] * j0[2];
      const double g11 = j1[0] * j1[0] + j1[1] * j1[1] + j1[2] * j1[2];
      const double g01 = j0[0] * j1[0] + j0[1] * j1[1] + j0[2] * j1[2];
      const double dA = Surf<ST>::weight(g) * sqrt(g00 * g11 - g01 * g01);
#pragma unroll
      for (int c = 0; c < NEN; ++c) fa[c] += N[c] * dA;
    }
    const double t3[3] = {tx, ty, tz};
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
#pragma unroll
      for (int i = 0; i < 3; ++i) work += scale * fa[c] * t3[i] * U[3 * nd[c] + i];
      if (rstage != nullptr) {
        const int d = dst_r[(int64_t)e * NEN + c];
        if (d >= 0) {
#pragma unroll
          for (int i = 0; i < 3; ++i) rstage[(int64_t)d * 3 + i] = -scale * fa[c] * t3[i];
        }
      }
    }
  }
  if (work_partial != nullptr) {
    const double s = block_sum(work);
    if (threadIdx.x == 0) work_partial[blockIdx.x] = s;
  }
}

static bool valid_surf(int st) { return st == 3 || st == 4 || st == 6; }

template <typename T>
static void launch_pressure(int stt, int n, const int32_t *conn, const double *X, const double *U,
                            double scale, const int32_t *dk, T *ks, const int32_t *dr, double *rs,
                            hipStream_t st) {
  const int grid = blocks_for((int64_t)n * stt);
  switch (stt) {
    case 3: k_pressure<T, 3><<<grid, kBlock, 0, st>>>(n, conn, X, U, scale, dk, ks, dr, rs); break;
    case 4: k_pressure<T, 4><<<grid, kBlock, 0, st>>>(n, conn, X, U, scale, dk, ks, dr, rs); break;
    case 6: k_pressure<T, 6><<<grid, kBlock, 0, st>>>(n, conn, X, U, scale, dk, ks, dr, rs); break;
  }
}

}  // namespace fea

using namespace fea;

extern "C" int fea_surface_blocks(int n_facets) { return n_facets > 0 ? blocks_for(n_facets) : 0; }

extern "C" int fea_surface_pressure(int surf_type, int dtype, int n_facets, const int32_t *conn,
                                    const double *X, const double *U, double scale,
                                    const int32_t *dst_k, void *kstage, const int32_t *dst_r,
                                    double *rstage, void *stream) {
  FEA_CHECK_ARG(valid_surf(surf_type), "fea_surface_pressure: bad surface type %d", surf_type);
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_surface_pressure: bad dtype");
  FEA_CHECK_ARG((kstage == nullptr) == (dst_k == nullptr), "fea_surface_pressure: kstage/dst_k mismatch");
  FEA_CHECK_ARG((rstage == nullptr) == (dst_r == nullptr), "fea_surface_pressure: rstage/dst_r mismatch");
  if (n_facets <= 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U, "fea_surface_pressure: null input");
  hipStream_t st = as_stream(stream);
  if (dtype == FEA_F64)
    launch_pressure<double>(surf_type, n_facets, conn, X, U, scale, dst_k, (double *)kstage, dst_r, rstage, st);
  else
    launch_pressure<float>(surf_type, n_facets, conn, X, U, scale, dst_k, (float *)kstage, dst_r, rstage, st);
  FEA_CHECK_LAUNCH("fea_surface_pressure");
  return FEA_OK;
}

extern "C" int fea_surface_volume(int surf_type, int n_facets, const int32_t *conn, const double *X,
                                  const double *U, double *partial, void *stream) {
  FEA_CHECK_ARG(valid_surf(surf_type), "fea_surface_volume: bad surface type %d", surf_type);
  if (n_facets <= 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U && partial, "fea_surface_volume: null input");
  const int grid = blocks_for(n_facets);
  hipStream_t st = as_stream(stream);
  switch (surf_type) {
    case 3: k_volume<3><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, partial); break;
    case 4: k_volume<4><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, partial); break;
    case 6: k_volume<6><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, partial); break;
  }
  FEA_CHECK_LAUNCH("fea_surface_volume");
  return FEA_OK;
}

extern "C" int fea_surface_traction(int surf_type, int n_facets, const int32_t *conn,
                                    const double *X, const double *U, const double *traction_host,
                                    double scale, const int32_t *dst_r, double *rstage,
                                    double *work_partial, void *stream) {
  FEA_CHECK_ARG(valid_surf(surf_type), "fea_surface_traction: bad surface type %d", surf_type);
  FEA_CHECK_ARG(traction_host != nullptr, "fea_surface_traction: null traction");
  FEA_CHECK_ARG((rstage == nullptr) == (dst_r == nullptr), "fea_surface_traction: rstage/dst_r mismatch");
  if (n_facets <= 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U, "fea_surface_traction: null input");
  const int grid = blocks_for(n_facets);
  hipStream_t st = as_stream(stream);
  const double tx = traction_host[0], ty = traction_host[1], tz = traction_host[2];
  switch (surf_type) {
    case 3: k_traction<3><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, tx, ty, tz, scale, dst_r, rstage, work_partial); break;
    case 4: k_traction<4><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, tx, ty, tz, scale, dst_r, rstage, work_partial); break;
    case 6: k_traction<6><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, tx, ty, tz, scale, dst_r, rstage, work_partial); break;
  }
  FEA_CHECK_LAUNCH("fea_surface_traction");
  return FEA_OK;
}

extern "C" int fea_surface_pressure_hvp(int surf_type, int n_facets, const int32_t *conn, const double *X,
                                        const double *U, double scale, const double *lam,
                                        const int32_t *dst, double *stage, void *stream) {
  FEA_CHECK_ARG(valid_surf(surf_type), "fea_surface_pressure_hvp: bad surface type %d", surf_type);
  if (n_facets <= 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U && lam && dst && stage, "fea_surface_pressure_hvp: null input");
  const int grid = blocks_for((int64_t)n_facets * surf_type);
  hipStream_t st = as_stream(stream);
  switch (surf_type) {
    case 3: k_pressure_hvp<3><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, scale, lam, dst, stage); break;
    case 4: k_pressure_hvp<4><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, scale, lam, dst, stage); break;
    case 6: k_pressure_hvp<6><<<grid, kBlock, 0, st>>>(n_facets, conn, X, U, scale, lam, dst, stage); break;
  }
  FEA_CHECK_LAUNCH("fea_surface_pressure_hvp");
  return FEA_OK;
}
