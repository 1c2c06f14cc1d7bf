// Element kernels: residual + consistent tangent, residual only, strain energy,
// Gauss-point geometry export and lumped mass for C3D4 / C3D8 / C3D10.
#include "element_traits.h"

namespace fea {

// Residual + tangent.  A workgroup takes EPW elements, one thread per (element,
// node a).  The element's nodal coordinates / displacements and, per Gauss
// point, dN/dX, F, w S and the packed upper triangle of w D are staged in LDS:
// thread (e, g) computes Gauss point g's state once (float64), instead of every
// node thread recomputing all of them.  Thread (e, a) then forms its row block
// K_a. (3 x 3 nen) and R_a from the staged state, in passes over BH columns b
// at a time to bound the live accumulators.  The B^T D B contraction runs in T.
template <int ET>
struct RkCfg;
template <>
struct RkCfg<4> {
  static constexpr int EPW = 16, BH = 4;
};
template <>
struct RkCfg<8> {
  static constexpr int EPW = 8, BH = 4;
};
template <>
struct RkCfg<10> {
  static constexpr int EPW = 6, BH = 5;
};

template <typename T, int ET>
__global__ __launch_bounds__(RkCfg<ET>::EPW * Elem<ET>::nen) void k_elem_rk(
    int n_elem, const int32_t *__restrict__ conn, const double *__restrict__ X, const double *__restrict__ U,
    MatParams mp, const int32_t *__restrict__ dst_k, T *__restrict__ kstage, const int32_t *__restrict__ dst_r,
    double *__restrict__ rstage) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  constexpr int EPW = RkCfg<ET>::EPW, BH = RkCfg<ET>::BH;
  constexpr int O_F = 3 * NEN, O_S = O_F + 9, O_D = O_S + 6, GS = O_D + 21;  // per-GP layout
  __shared__ double sx[EPW][NEN][3], su[EPW][NEN][3];
  __shared__ T sg[EPW][NGP][GS];
  const int le = threadIdx.x / NEN, a = threadIdx.x - le * NEN;
  const int64_t e = (int64_t)blockIdx.x * EPW + le;
  const bool valid = e < n_elem;
  if (valid) {
    const int64_t n = conn[e * NEN + a];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      sx[le][a][i] = X[3 * n + i];
      su[le][a][i] = U[3 * n + i];
    }
  }
  __syncthreads();
  if (valid && a < NGP) {
    // Gauss point g = a of element e: state in float64 (near incompressibility
    // the deviatoric part of the tangent is a small difference of bulk-sized terms)
    double dNd[NEN][3];
    const double wd = gp_geometry<double, ET>(a, sx[le], dNd);
    double Hd[9];
    disp_grad<double, NEN>(su[le], dNd, Hd);
    GPState<double> sd;
    kinematics(Hd, sd);
    double Sd[6];
    stress(sd, mp, Sd);
    double Dd[6][6];
    tangent(sd, mp, Dd);
    T *o = sg[le][a];
#pragma unroll
    for (int c = 0; c < NEN; ++c)
#pragma unroll
      for (int i = 0; i < 3; ++i) o[3 * c + i] = (T)dNd[c][i];
#pragma unroll
    for (int k = 0; k < 9; ++k) o[O_F + k] = (T)sd.F[k];
#pragma unroll
    for (int I = 0; I < 6; ++I) o[O_S + I] = (T)(wd * Sd[I]);
    int q = 0;
#pragma unroll
    for (int I = 0; I < 6; ++I)
#pragma unroll
      for (int J = I; J < 6; ++J) o[O_D + q++] = (T)(wd * Dd[I][J]);
  }
  __syncthreads();
  if (!valid) return;
  double Ra[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int b0 = 0; b0 < NEN; b0 += BH) {
    T K[BH][9];
#pragma unroll
    for (int bb = 0; bb < BH; ++bb)
#pragma unroll
      for (int k = 0; k < 9; ++k) K[bb][k] = T(0);
    for (int g = 0; g < NGP; ++g) {
      const T *st = sg[le][g];
      T F[9], dNa[3];
#pragma unroll
      for (int k = 0; k < 9; ++k) F[k] = st[O_F + k];
#pragma unroll
      for (int i = 0; i < 3; ++i) dNa[i] = st[3 * a + i];
      T Ba[6][3];
      bmat(F, dNa, Ba);
      if (b0 == 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          T acc = T(0);
#pragma unroll
          for (int I = 0; I < 6; ++I) acc += st[O_S + I] * Ba[I][j];
          Ra[j] += (double)acc;
        }
      }
      // DB = (w D) Ba, D read from the packed upper triangle
      T DB[6][3];
#pragma unroll
      for (int I = 0; I < 6; ++I)
#pragma unroll
        for (int j = 0; j < 3; ++j) DB[I][j] = T(0);
#pragma unroll
      for (int I = 0; I < 6; ++I)
#pragma unroll
        for (int L = 0; L < 6; ++L) {
          const int r = I < L ? I : L, c = I < L ? L : I;
          const T d = st[O_D + r * 6 - (r * (r - 1)) / 2 + (c - r)];
#pragma unroll
          for (int j = 0; j < 3; ++j) DB[I][j] += d * Ba[L][j];
        }
      // (w S) dN_a (full symmetric S)
      T SdNa[3];
#pragma unroll
      for (int A = 0; A < 3; ++A)
        SdNa[A] = st[O_S + vidx(A, 0)] * dNa[0] + st[O_S + vidx(A, 1)] * dNa[1] + st[O_S + vidx(A, 2)] * dNa[2];
#pragma unroll
      for (int bb = 0; bb < BH; ++bb) {
        const int b = b0 + bb;
        if (b < NEN) {
          const T dNb[3] = {st[3 * b + 0], st[3 * b + 1], st[3 * b + 2]};
          T Bb[6][3];
          bmat(F, dNb, Bb);
          const T geo = SdNa[0] * dNb[0] + SdNa[1] * dNb[1] + SdNa[2] * dNb[2];
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              T acc = (j == k) ? geo : T(0);
#pragma unroll
              for (int I = 0; I < 6; ++I) acc += DB[I][j] * Bb[I][k];
              K[bb][j * 3 + k] += acc;
            }
        }
      }
    }
    if (kstage != nullptr) {
      const int64_t base = (e * NEN + a) * NEN;
#pragma unroll
      for (int bb = 0; bb < BH; ++bb) {
        const int b = b0 + bb;
        if (b < NEN) {
          const int d = dst_k[base + b];
          if (d >= 0) {
            T *o = kstage + (int64_t)d * 9;
#pragma unroll
            for (int k = 0; k < 9; ++k) o[k] = K[bb][k];
          }
        }
      }
    }
  }
  if (rstage != nullptr) {
    const int d = dst_r[e * NEN + a];
    if (d >= 0) {
      rstage[(int64_t)d * 3 + 0] = Ra[0];
      rstage[(int64_t)d * 3 + 1] = Ra[1];
      rstage[(int64_t)d * 3 + 2] = Ra[2];
    }
  }
}

// One thread per element: nodal forces R_a = sum_g w P dN_a with P = F S.
template <typename T, int ET>
__global__ __launch_bounds__(kBlock) void k_elem_r(int n_elem, const int32_t *__restrict__ conn,
                                                   const double *__restrict__ X,
                                                   const double *__restrict__ U, MatParams mp,
                                                   const int32_t *__restrict__ dst_r,
                                                   double *__restrict__ rstage) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  // In float, coordinates and displacements are taken relative to the element's first
  // node, formed in double: the Jacobian and grad u only see differences (the shape
  // function derivatives sum to zero), and a rigid translation of the element no longer
  // costs float digits of the strain.  In double the absolute values are used.
  constexpr bool kRel = sizeof(T) < sizeof(double);
  const int64_t n0 = conn[(int64_t)e * NEN];
  double x0[3] = {0.0, 0.0, 0.0}, u0[3] = {0.0, 0.0, 0.0};
  if (kRel) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x0[i] = X[3 * n0 + i];
      u0[i] = U[3 * n0 + i];
    }
  }
  T x[NEN][3], u[NEN][3];
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[c][i] = (T)(X[3 * n + i] - x0[i]);
      u[c][i] = (T)(U[3 * n + i] - u0[i]);
    }
  }
  T R[NEN][3];
#pragma unroll
  for (int c = 0; c < NEN; ++c) R[c][0] = R[c][1] = R[c][2] = T(0);
  for (int g = 0; g < NGP; ++g) {
    T dN[NEN][3];
    const T w = gp_geometry<T, ET>(g, x, dN);
    T H[9];
    disp_grad<T, NEN>(u, dN, H);
    GPState<T> s;
    kinematics(H, s);
    T S[6];
    stress(s, mp, S);
    T P[9];  // P = F S, scaled by w
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int A = 0; A < 3; ++A)
        P[j * 3 + A] = w * (s.F[j * 3 + 0] * S[vidx(0, A)] + s.F[j * 3 + 1] * S[vidx(1, A)] +
                            s.F[j * 3 + 2] * S[vidx(2, A)]);
#pragma unroll
    for (int c = 0; c < NEN; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        R[c][j] += P[j * 3 + 0] * dN[c][0] + P[j * 3 + 1] * dN[c][1] + P[j * 3 + 2] * dN[c][2];
  }
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    const int d = dst_r[(int64_t)e * NEN + c];
    if (d >= 0) {
      rstage[(int64_t)d * 3 + 0] = (double)R[c][0];
      rstage[(int64_t)d * 3 + 1] = (double)R[c][1];
      rstage[(int64_t)d * 3 + 2] = (double)R[c][2];
    }
  }
}

template <int ET>
__global__ __launch_bounds__(kBlock) void k_elem_energy(int n_elem, const int32_t *__restrict__ conn,
                                                        const double *__restrict__ X,
                                                        const double *__restrict__ U, MatParams mp,
                                                        double *__restrict__ partial) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (e < n_elem) {
    double x[NEN][3], u[NEN][3];
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
      const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        x[c][i] = X[3 * n + i];
        u[c][i] = U[3 * n + i];
      }
    }
    for (int g = 0; g < NGP; ++g) {
      double dN[NEN][3];
      const double w = gp_geometry<double, ET>(g, x, dN);
      double H[9];
      disp_grad<double, NEN>(u, dN, H);
      GPState<double> s;
      kinematics(H, s);
      acc += w * energy_density(s.I1, s.I2, s.J, s.Jm1, mp);
    }
  }
  const double tot = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <int ET>
__global__ __launch_bounds__(kBlock) void k_elem_prepare(int n_elem, const int32_t *__restrict__ conn,
                                                         const double *__restrict__ X,
                                                         double *__restrict__ dNdX,
                                                         double *__restrict__ wdet,
                                                         double *__restrict__ Nout) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double x[NEN][3];
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
    for (int i = 0; i < 3; ++i) x[c][i] = X[3 * n + i];
  }
  for (int g = 0; g < NGP; ++g) {
    double dN[NEN][3];
    const double w = gp_geometry<double, ET>(g, x, dN);
    const int64_t ge = (int64_t)g * n_elem + e;
    if (dNdX != nullptr) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < NEN; ++c) dNdX[(ge * 3 + i) * NEN + c] = dN[c][i];
    }
    if (wdet != nullptr) wdet[ge] = w;
    if (Nout != nullptr) {
      double N[NEN];
      Elem<ET>::shape(g, N);
#pragma unroll
      for (int c = 0; c < NEN; ++c) Nout[ge * NEN + c] = N[c];
    }
  }
}

template <int ET>
__global__ __launch_bounds__(kBlock) void k_elem_lumped_mass(int n_elem, const int32_t *__restrict__ conn,
                                                             const double *__restrict__ X,
                                                             const int32_t *__restrict__ dof_map,
                                                             double rho,
                                                             const int32_t *__restrict__ dst_r,
                                                             double *__restrict__ rstage) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_elem * NEN) return;
  const int e = (int)(t / NEN), a = (int)(t - (int64_t)e * NEN);
  const int d = dst_r[(int64_t)e * NEN + a];
  if (d < 0) return;
  double x[NEN][3];
  int nd[NEN];
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    nd[c] = conn[(int64_t)e * NEN + c];
#pragma unroll
    for (int i = 0; i < 3; ++i) x[c][i] = X[3 * (int64_t)nd[c] + i];
  }
  double m[3] = {0.0, 0.0, 0.0};
  for (int g = 0; g < NGP; ++g) {
    double dN[NEN][3];
    const double w = gp_geometry<double, ET>(g, x, dN);
    double N[NEN];
    Elem<ET>::shape(g, N);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sn = 0.0;
#pragma unroll
      for (int b = 0; b < NEN; ++b) sn += (dof_map[3 * (int64_t)nd[b] + j] >= 0) ? N[b] : 0.0;
      m[j] += rho * N[a] * w * sn;
    }
  }
  rstage[(int64_t)d * 3 + 0] = m[0];
  rstage[(int64_t)d * 3 + 1] = m[1];
  rstage[(int64_t)d * 3 + 2] = m[2];
}

// Consistent mass blocks (Element_3D.get_mass_matrix, C3base.py:348-384):
// M_ab = rho sum_g N_a N_b w detJ times the 3x3 identity; one thread per
// (element, node a) writes its row of blocks through the tangent's staging map.
template <int ET>
__global__ __launch_bounds__(kBlock) void k_elem_mass(int n_elem, const int32_t *__restrict__ conn,
                                                      const double *__restrict__ X, double rho,
                                                      const int32_t *__restrict__ dst_k,
                                                      double *__restrict__ kstage) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_elem * NEN) return;
  const int e = (int)(t / NEN), a = (int)(t - (int64_t)e * NEN);
  double x[NEN][3];
#pragma unroll
  for (int c = 0; c < NEN; ++c) {
    const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
    for (int i = 0; i < 3; ++i) x[c][i] = X[3 * n + i];
  }
  double m[NEN];
#pragma unroll
  for (int b = 0; b < NEN; ++b) m[b] = 0.0;
  for (int g = 0; g < NGP; ++g) {
    double dN[NEN][3], N[NEN];
    const double w = gp_geometry<double, ET>(g, x, dN);
    Elem<ET>::shape(g, N);
#pragma unroll
    for (int b = 0; b < NEN; ++b) m[b] += rho * N[a] * N[b] * w;
  }
  const int64_t base = ((int64_t)e * NEN + a) * NEN;
#pragma unroll
  for (int b = 0; b < NEN; ++b) {
    const int d = dst_k[base + b];
    if (d >= 0) {
      double *o = kstage + (int64_t)d * 9;
#pragma unroll
      for (int k = 0; k < 9; ++k) o[k] = (k == 0 || k == 4 || k == 8) ? m[b] : 0.0;
    }
  }
}

static bool valid_elem(int et) { return et == 4 || et == 8 || et == 10; }

static MatParams mat_from(const double *p) { return MatParams{p[0], p[1], p[2]}; }

template <typename T, int ET>
static void launch_rk_t(int n, const int32_t *conn, const double *X, const double *U, MatParams mp,
                        const int32_t *dk, T *ks, const int32_t *dr, double *rs, hipStream_t st) {
  constexpr int EPW = RkCfg<ET>::EPW;
  const int grid = (int)(((int64_t)n + EPW - 1) / EPW);
  k_elem_rk<T, ET><<<grid, EPW * Elem<ET>::nen, 0, st>>>(n, conn, X, U, mp, dk, ks, dr, rs);
}

template <typename T>
static int launch_rk(int et, int n, const int32_t *conn, const double *X, const double *U,
                     MatParams mp, const int32_t *dk, T *ks, const int32_t *dr, double *rs,
                     hipStream_t st) {
  switch (et) {
    case 4: launch_rk_t<T, 4>(n, conn, X, U, mp, dk, ks, dr, rs, st); break;
    case 8: launch_rk_t<T, 8>(n, conn, X, U, mp, dk, ks, dr, rs, st); break;
    case 10: launch_rk_t<T, 10>(n, conn, X, U, mp, dk, ks, dr, rs, st); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_residual_tangent");
  return FEA_OK;
}

template <typename T>
static int launch_r(int et, int n, const int32_t *conn, const double *X, const double *U,
                    MatParams mp, const int32_t *dr, double *rs, hipStream_t st) {
  const int grid = blocks_for(n);
  switch (et) {
    case 4: k_elem_r<T, 4><<<grid, kBlock, 0, st>>>(n, conn, X, U, mp, dr, rs); break;
    case 8: k_elem_r<T, 8><<<grid, kBlock, 0, st>>>(n, conn, X, U, mp, dr, rs); break;
    case 10: k_elem_r<T, 10><<<grid, kBlock, 0, st>>>(n, conn, X, U, mp, dr, rs); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_residual");
  return FEA_OK;
}

}  // namespace fea

using namespace fea;

extern "C" int fea_elem_residual_tangent(int elem_type, int dtype, int n_elem, const int32_t *conn,
                                         const double *X, const double *U,
                                         const double *params_host, const int32_t *dst_k,
                                         void *kstage, const int32_t *dst_r, double *rstage,
                                         void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_residual_tangent: bad element type %d", elem_type);
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_elem_residual_tangent: bad dtype %d", dtype);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr, "fea_elem_residual_tangent: bad arguments");
  FEA_CHECK_ARG((kstage == nullptr) == (dst_k == nullptr), "fea_elem_residual_tangent: kstage/dst_k mismatch");
  FEA_CHECK_ARG((rstage == nullptr) == (dst_r == nullptr), "fea_elem_residual_tangent: rstage/dst_r mismatch");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U, "fea_elem_residual_tangent: null input");
  const MatParams mp = mat_from(params_host);
  if (dtype == FEA_F64)
    return launch_rk<double>(elem_type, n_elem, conn, X, U, mp, dst_k, (double *)kstage, dst_r,
                             rstage, as_stream(stream));
  return launch_rk<float>(elem_type, n_elem, conn, X, U, mp, dst_k, (float *)kstage, dst_r, rstage,
                          as_stream(stream));
}

extern "C" int fea_elem_residual(int elem_type, int dtype, int n_elem, const int32_t *conn,
                                 const double *X, const double *U, const double *params_host,
                                 const int32_t *dst_r, double *rstage, void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_residual: bad element type %d", elem_type);
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_elem_residual: bad dtype %d", dtype);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr, "fea_elem_residual: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U && dst_r && rstage, "fea_elem_residual: null input");
  const MatParams mp = mat_from(params_host);
  if (dtype == FEA_F64)
    return launch_r<double>(elem_type, n_elem, conn, X, U, mp, dst_r, rstage, as_stream(stream));
  return launch_r<float>(elem_type, n_elem, conn, X, U, mp, dst_r, rstage, as_stream(stream));
}

extern "C" int fea_elem_energy_blocks(int n_elem) { return n_elem > 0 ? blocks_for(n_elem) : 0; }

extern "C" int fea_elem_energy(int elem_type, int n_elem, const int32_t *conn, const double *X,
                               const double *U, const double *params_host, double *partial,
                               void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_energy: bad element type %d", elem_type);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr, "fea_elem_energy: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U && partial, "fea_elem_energy: null input");
  const MatParams mp = mat_from(params_host);
  const int grid = blocks_for(n_elem);
  hipStream_t st = as_stream(stream);
  switch (elem_type) {
    case 4: k_elem_energy<4><<<grid, kBlock, 0, st>>>(n_elem, conn, X, U, mp, partial); break;
    case 8: k_elem_energy<8><<<grid, kBlock, 0, st>>>(n_elem, conn, X, U, mp, partial); break;
    case 10: k_elem_energy<10><<<grid, kBlock, 0, st>>>(n_elem, conn, X, U, mp, partial); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_energy");
  return FEA_OK;
}

extern "C" int fea_elem_prepare(int elem_type, int n_elem, const int32_t *conn, const double *X,
                                double *dNdX, double *wdet, double *N, void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_prepare: bad element type %d", elem_type);
  FEA_CHECK_ARG(n_elem >= 0, "fea_elem_prepare: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X, "fea_elem_prepare: null input");
  const int grid = blocks_for(n_elem);
  hipStream_t st = as_stream(stream);
  switch (elem_type) {
    case 4: k_elem_prepare<4><<<grid, kBlock, 0, st>>>(n_elem, conn, X, dNdX, wdet, N); break;
    case 8: k_elem_prepare<8><<<grid, kBlock, 0, st>>>(n_elem, conn, X, dNdX, wdet, N); break;
    case 10: k_elem_prepare<10><<<grid, kBlock, 0, st>>>(n_elem, conn, X, dNdX, wdet, N); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_prepare");
  return FEA_OK;
}

extern "C" int fea_elem_lumped_mass(int elem_type, int n_elem, const int32_t *conn,
                                    const double *X, const int32_t *dof_map, double rho,
                                    const int32_t *dst_r, double *rstage, void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_lumped_mass: bad element type %d", elem_type);
  FEA_CHECK_ARG(n_elem >= 0, "fea_elem_lumped_mass: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && dof_map && dst_r && rstage, "fea_elem_lumped_mass: null input");
  const int grid = blocks_for((int64_t)n_elem * elem_type);
  hipStream_t st = as_stream(stream);
  switch (elem_type) {
    case 4: k_elem_lumped_mass<4><<<grid, kBlock, 0, st>>>(n_elem, conn, X, dof_map, rho, dst_r, rstage); break;
    case 8: k_elem_lumped_mass<8><<<grid, kBlock, 0, st>>>(n_elem, conn, X, dof_map, rho, dst_r, rstage); break;
    case 10: k_elem_lumped_mass<10><<<grid, kBlock, 0, st>>>(n_elem, conn, X, dof_map, rho, dst_r, rstage); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_lumped_mass");
  return FEA_OK;
}

extern "C" int fea_elem_mass(int elem_type, int n_elem, const int32_t *conn, const double *X, double rho,
                             const int32_t *dst_k, double *kstage, void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_mass: bad element type %d", elem_type);
  FEA_CHECK_ARG(n_elem >= 0, "fea_elem_mass: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && dst_k && kstage, "fea_elem_mass: null input");
  const int grid = blocks_for((int64_t)n_elem * elem_type);
  hipStream_t st = as_stream(stream);
  switch (elem_type) {
    case 4: k_elem_mass<4><<<grid, kBlock, 0, st>>>(n_elem, conn, X, rho, dst_k, kstage); break;
    case 8: k_elem_mass<8><<<grid, kBlock, 0, st>>>(n_elem, conn, X, rho, dst_k, kstage); break;
    case 10: k_elem_mass<10><<<grid, kBlock, 0, st>>>(n_elem, conn, X, rho, dst_k, kstage); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_mass");
  return FEA_OK;
}
