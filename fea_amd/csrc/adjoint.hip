// Hand-written adjoint (vector-Jacobian product) of the element internal force.
//
// For f_a = sum_g w detJ0 P(F) G_a  (G_a = dN_a/dX, F = I + u^T G, P = F S)
// and an adjoint field lambda, the scalar phi = lambda . f_int has, per Gauss
// point (M = lambda^T G, the adjoint field's gradient; H = F - I):
//   d phi / d u_b     = wJ Q G_b,          Q = dP[M] = M S + F (CC : sym(F^T M))
//   d phi / d X_b     = wJ Z G_b,          Z = (P : M) I - H^T Q - M^T P
//   d phi / d theta_k = wJ S_k : sym(F^T M),  S_k = dS/dtheta_k
// (theta = {w1, w2, kv}; S is linear in them).  X enters through G and detJ0:
// dG = -G (dX^T G) and d detJ0 = detJ0 tr(dX^T G).  Everything is recomputed
// from X, U and lambda (nothing is stored from the forward pass), in float64.
// The reference differentiates its assembly with torch autograd
// (C3base.py:413-520 under requires_grad); this kernel replaces that.
#include "common.h"
#include "element_traits.h"

namespace fea {

template <int ET>
__global__ __launch_bounds__(kBlock) void k_elem_adjoint(int n_elem, const int32_t *__restrict__ conn,
                                                         const double *__restrict__ X,
                                                         const double *__restrict__ U, MatParams mp,
                                                         const double *__restrict__ lam,
                                                         const int32_t *__restrict__ dst,
                                                         double *__restrict__ gx_stage,
                                                         double *__restrict__ gu_stage,
                                                         double *__restrict__ gp_partial) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double gth[3] = {0.0, 0.0, 0.0};
  if (e < n_elem) {
    double x[NEN][3], u[NEN][3], l[NEN][3];
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
      const int64_t n = conn[(int64_t)e * NEN + c];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        x[c][i] = X[3 * n + i];
        u[c][i] = U[3 * n + i];
        l[c][i] = lam[3 * n + i];
      }
    }
    double gx[NEN][3], gu[NEN][3];
#pragma unroll
    for (int c = 0; c < NEN; ++c)
#pragma unroll
      for (int i = 0; i < 3; ++i) gx[c][i] = gu[c][i] = 0.0;
    const MatParams unit[3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int g = 0; g < NGP; ++g) {
      double G[NEN][3];
      const double w = gp_geometry<double, ET>(g, x, G);
      double H[9], M[9];
      disp_grad<double, NEN>(u, G, H);
      disp_grad<double, NEN>(l, G, M);
      GPState<double> s;
      kinematics(H, s);
      double S[6], D[6][6];
      stress(s, mp, S);
      tangent(s, mp, D);
      const double *F = s.F;
      double P[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int A = 0; A < 3; ++A)
          P[i * 3 + A] = F[i * 3 + 0] * S[vidx(0, A)] + F[i * 3 + 1] * S[vidx(1, A)] + F[i * 3 + 2] * S[vidx(2, A)];
      // dE = sym(F^T M) in Voigt form with engineering shear
      double FtM[9];
#pragma unroll
      for (int A = 0; A < 3; ++A)
#pragma unroll
        for (int B = 0; B < 3; ++B)
          FtM[A * 3 + B] = F[0 * 3 + A] * M[0 * 3 + B] + F[1 * 3 + A] * M[1 * 3 + B] + F[2 * 3 + A] * M[2 * 3 + B];
      double dE[6];
#pragma unroll
      for (int I = 0; I < 6; ++I) {
        const int A = vA(I), B = vB(I);
        dE[I] = (I < 3) ? FtM[A * 3 + A] : FtM[A * 3 + B] + FtM[B * 3 + A];
      }
      double dS[6];
#pragma unroll
      for (int I = 0; I < 6; ++I) {
        double acc = 0.0;
#pragma unroll
        for (int J = 0; J < 6; ++J) acc += D[I][J] * dE[J];
        dS[I] = acc;
      }
      // material-parameter sensitivities: S_k : dE
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double Sk[6];
        stress(s, unit[k], Sk);
        double acc = 0.0;
#pragma unroll
        for (int I = 0; I < 6; ++I) acc += Sk[I] * dE[I];
        gth[k] += w * acc;
      }
      double Q[9], PM = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int J = 0; J < 3; ++J) {
          double q = 0.0;
#pragma unroll
          for (int K = 0; K < 3; ++K) q += M[i * 3 + K] * S[vidx(K, J)] + F[i * 3 + K] * dS[vidx(K, J)];
          Q[i * 3 + J] = q;
          PM += P[i * 3 + J] * M[i * 3 + J];
        }
      double Z[9];
#pragma unroll
      for (int K = 0; K < 3; ++K)
#pragma unroll
        for (int J = 0; J < 3; ++J) {
          double z = (K == J) ? PM : 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) z -= H[i * 3 + K] * Q[i * 3 + J] + M[i * 3 + K] * P[i * 3 + J];
          Z[K * 3 + J] = z;
        }
#pragma unroll
      for (int c = 0; c < NEN; ++c)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          gx[c][K] += w * (Z[K * 3 + 0] * G[c][0] + Z[K * 3 + 1] * G[c][1] + Z[K * 3 + 2] * G[c][2]);
          gu[c][K] += w * (Q[K * 3 + 0] * G[c][0] + Q[K * 3 + 1] * G[c][1] + Q[K * 3 + 2] * G[c][2]);
        }
    }
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
      const int64_t d = dst[(int64_t)e * NEN + c];
      if (d < 0) continue;
      if (gx_stage != nullptr) {
        gx_stage[d * 3 + 0] = gx[c][0];
        gx_stage[d * 3 + 1] = gx[c][1];
        gx_stage[d * 3 + 2] = gx[c][2];
      }
      if (gu_stage != nullptr) {
        gu_stage[d * 3 + 0] = gu[c][0];
        gu_stage[d * 3 + 1] = gu[c][1];
        gu_stage[d * 3 + 2] = gu[c][2];
      }
    }
  }
  if (gp_partial != nullptr) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double t = block_sum(gth[k]);
      if (threadIdx.x == 0) gp_partial[(int64_t)k * gridDim.x + blockIdx.x] = t;
    }
  }
}

}  // namespace fea

using namespace fea;

extern "C" int fea_elem_adjoint_blocks(int n_elem) { return n_elem > 0 ? blocks_for(n_elem) : 0; }

extern "C" int fea_elem_adjoint(int elem_type, int n_elem, const int32_t *conn, const double *X,
                                const double *U, const double *params, const double *lam,
                                const int32_t *dst, double *gx_stage, double *gu_stage,
                                double *gp_partial, void *stream) {
  FEA_CHECK_ARG(elem_type == 4 || elem_type == 8 || elem_type == 10, "fea_elem_adjoint: bad element type %d",
                elem_type);
  FEA_CHECK_ARG(n_elem >= 0 && params != nullptr, "fea_elem_adjoint: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U && lam && dst, "fea_elem_adjoint: null input");
  FEA_CHECK_ARG(gx_stage || gu_stage || gp_partial, "fea_elem_adjoint: no output requested");
  const MatParams mp{params[0], params[1], params[2]};
  const int grid = blocks_for(n_elem);
  hipStream_t st = as_stream(stream);
  switch (elem_type) {
    case 4: k_elem_adjoint<4><<<grid, kBlock, 0, st>>>(n_elem, conn, X, U, mp, lam, dst, gx_stage, gu_stage, gp_partial); break;
    case 8: k_elem_adjoint<8><<<grid, kBlock, 0, st>>>(n_elem, conn, X, U, mp, lam, dst, gx_stage, gu_stage, gp_partial); break;
    case 10: k_elem_adjoint<10><<<grid, kBlock, 0, st>>>(n_elem, conn, X, U, mp, lam, dst, gx_stage, gu_stage, gp_partial); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_adjoint");
  return FEA_OK;
}
