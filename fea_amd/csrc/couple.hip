// Kinematic coupling of node sets to reference points: Rodrigues kinematics and the
// chain-rule condensation of the extended system onto the generalized coordinates
// (replaces Couple.modify_RGC / modify_R_K / _calculate_Ydotz, constraints/couple.py:36-270,
// and the reduction of _assemble_reduced_Matrix, assembly.py:357-387).
#include "common.h"

namespace fea {

// Coefficients of R(z) y - y = A(s) (z x y) + B(s) (z x (z x y)), s = |z|^2, and their
// first and second s-derivatives.  Series below s = 2.5 (terms (-s)^n / (2n+1)! and
// (-s)^n / (2n+2)!, 20 terms: the remainder is below 1e-25), closed forms above
// (A' = (cos t - A) / 2s, B' = (A - 2B) / 2s, A'' = (-A/2 - 3A') / 2s, B'' = (A' - 4B') / 2s).
struct RotCoef {
  double A, B, A1, B1, A2, B2;
};

__device__ __forceinline__ RotCoef rot_coef(double s) {
  RotCoef c;
  if (s < 2.5) {
    double A = 0.0, B = 0.0, A1 = 0.0, B1 = 0.0, A2 = 0.0, B2 = 0.0;
    double f1 = 1.0, f2 = 0.5;   // 1/(2n+1)!, 1/(2n+2)!
    double pm2 = 0.0, pm1 = 0.0, p = 1.0;   // (-s)^(n-2), (-s)^(n-1), (-s)^n
    for (int n = 0; n < 20; ++n) {
      A += p * f1;
      B += p * f2;
      if (n >= 1) {
        A1 -= n * pm1 * f1;      // d/ds (-s)^n = -n (-s)^(n-1)
        B1 -= n * pm1 * f2;
      }
      if (n >= 2) {
        A2 += n * (n - 1) * pm2 * f1;
        B2 += n * (n - 1) * pm2 * f2;
      }
      pm2 = pm1;
      pm1 = p;
      p *= -s;
      f1 /= (double)((2 * n + 2) * (2 * n + 3));
      f2 /= (double)((2 * n + 3) * (2 * n + 4));
    }
    c.A = A; c.B = B; c.A1 = A1; c.B1 = B1; c.A2 = A2; c.B2 = B2;
  } else {
    const double t = sqrt(s);
    const double st = sin(t), ct = cos(t);
    c.A = st / t;
    c.B = (1.0 - ct) / s;
    c.A1 = (ct - c.A) / (2.0 * s);
    c.B1 = (c.A - 2.0 * c.B) / (2.0 * s);
    c.A2 = (-0.5 * c.A - 3.0 * c.A1) / (2.0 * s);
    c.B2 = (c.A1 - 4.0 * c.B1) / (2.0 * s);
  }
  return c;
}

// One point y: Y = R(z) y - y and, when Yd != nullptr, dY_i/dz_p at [i*3+p] and
// d2Y_i/dz_p dz_q at [i*9 + p*3 + q].
__device__ void rodrigues_point(const double z[3], const double y[3], double Y[3], double *Yd,
                                double *Ydd) {
  const double s = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
  const RotCoef c = rot_coef(s);
  const double zy = z[0] * y[0] + z[1] * y[1] + z[2] * y[2];
  const double zxy[3] = {z[1] * y[2] - z[2] * y[1], z[2] * y[0] - z[0] * y[2], z[0] * y[1] - z[1] * y[0]};
  double W[3];                                   // z x (z x y) = z (z.y) - s y
  for (int i = 0; i < 3; ++i) W[i] = z[i] * zy - s * y[i];
  for (int i = 0; i < 3; ++i) Y[i] = c.A * zxy[i] + c.B * W[i];
  if (Yd == nullptr) return;
  // E[i][p] = d(z x y)_i / dz_p = eps_ipk y_k; dW[i][p] = d W_i / d z_p
  const double E[9] = {0.0, y[2], -y[1], -y[2], 0.0, y[0], y[1], -y[0], 0.0};
  double dW[9];
  for (int i = 0; i < 3; ++i)
    for (int p = 0; p < 3; ++p)
      dW[i * 3 + p] = (i == p ? zy : 0.0) + z[i] * y[p] - 2.0 * z[p] * y[i];
  for (int i = 0; i < 3; ++i)
    for (int p = 0; p < 3; ++p)
      Yd[i * 3 + p] = 2.0 * c.A1 * z[p] * zxy[i] + c.A * E[i * 3 + p] + 2.0 * c.B1 * z[p] * W[i] +
                      c.B * dW[i * 3 + p];
  if (Ydd == nullptr) return;
  for (int i = 0; i < 3; ++i)
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) {
        const double dpq = (p == q) ? 1.0 : 0.0;
        const double d2W = (i == p ? y[q] : 0.0) + (i == q ? y[p] : 0.0) - 2.0 * dpq * y[i];
        Ydd[i * 9 + p * 3 + q] =
            4.0 * c.A2 * z[p] * z[q] * zxy[i] + 2.0 * c.A1 * dpq * zxy[i] +
            2.0 * c.A1 * (z[p] * E[i * 3 + q] + z[q] * E[i * 3 + p]) + 4.0 * c.B2 * z[p] * z[q] * W[i] +
            2.0 * c.B1 * dpq * W[i] + 2.0 * c.B1 * (z[p] * dW[i * 3 + q] + z[q] * dW[i * 3 + p]) +
            c.B * d2W;
      }
}

// couple index of coupled node b (cnode_ptr is short: one entry per couple)
__device__ __forceinline__ int couple_of(const int32_t *cnode_ptr, int n_couples, int b) {
  int k = 0;
  while (k + 1 < n_couples && cnode_ptr[k + 1] <= b) ++k;
  return k;
}

__global__ __launch_bounds__(kBlock) void k_couple_kin(fea_couple_map m, double *__restrict__ state,
                                                       int with_derivs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= m.n_cnodes) return;
  const int k = couple_of(m.cnode_ptr, m.n_couples, b);
  const double *rp = state + m.rp_state[k];
  const double z[3] = {rp[3], rp[4], rp[5]};
  const double y[3] = {m.y[3 * b], m.y[3 * b + 1], m.y[3 * b + 2]};
  double Y[3];
  rodrigues_point(z, y, Y, with_derivs ? m.ydot + 9 * (int64_t)b : nullptr,
                  with_derivs ? m.ydot2 + 27 * (int64_t)b : nullptr);
  const int64_t node = m.cnode[b];
  for (int c = 0; c < 3; ++c) state[3 * node + c] = rp[c] + Y[c];
}

// T(q, d): derivative of coupled DOF q = 3b + c by the point's coordinate d (v then z)
__device__ __forceinline__ double tcol(const double *ydot, int q, int d) {
  return d < 3 ? ((q % 3) == d ? 1.0 : 0.0) : ydot[3 * (int64_t)q + (d - 3)];
}

// node-node entries of the reduced matrix and node rows of the reduced residual
template <typename TO>
__global__ __launch_bounds__(kBlock) void k_couple_ff(fea_couple_map m, const double *__restrict__ K_ext,
                                                      const double *__restrict__ R_ext, TO *__restrict__ val,
                                                      double *__restrict__ R_red) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (K_ext != nullptr && i < m.nnz_red) {
    const int src = m.ff_src[i];
    if (src >= 0) val[i] = (TO)K_ext[src];
  }
  if (i < m.n_red_nodes) R_red[i] = R_ext[m.rf_src[i]];
}

// (node row, point DOF) entries and their transposes: one thread per (row, couple) pair
template <typename TO>
__global__ __launch_bounds__(kBlock) void k_couple_pairs(fea_couple_map m, const double *__restrict__ K_ext,
                                                         TO *__restrict__ val) {
  const int pr = blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= m.n_pairs) return;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = m.pair_ptr[pr]; e < m.pair_ptr[pr + 1]; ++e) {
    const double v = K_ext[m.pe_slot[e]];
    const int q = m.pe_dof[e];
    acc[q % 3] += v;
    const double *yd = m.ydot + 3 * (int64_t)q;
    acc[3] += v * yd[0];
    acc[4] += v * yd[1];
    acc[5] += v * yd[2];
  }
  for (int d = 0; d < 6; ++d) {
    const int s1 = m.pair_row_slot[6 * pr + d];
    const int s2 = m.pair_col_slot[6 * pr + d];
    if (s1 >= 0) val[s1] = (TO)acc[d];
    if (s2 >= 0) val[s2] = (TO)acc[d];
  }
}

// per coupled DOF q: the 6x6 block contribution T(q,:)^T [sum_q' K(q,q') T(q',:)] plus
// R_q d2u_q/dz2 (36 values), and R_q T(q,:) (6 values)
__global__ __launch_bounds__(kBlock) void k_couple_rows(fea_couple_map m, const double *__restrict__ K_ext,
                                                        const double *__restrict__ R_ext) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * m.n_cnodes) return;
  double *out = m.partial + 42 * (int64_t)q;
  const double r = R_ext[m.crow_ext[q]];
  double t[6];
  for (int d = 0; d < 6; ++d) t[d] = tcol(m.ydot, q, d);
  for (int d = 0; d < 6; ++d) out[36 + d] = r * t[d];
  if (K_ext == nullptr) return;
  double u[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = m.cr_ptr[q]; e < m.cr_ptr[q + 1]; ++e) {
    const double v = K_ext[m.ce_slot[e]];
    const int q2 = m.ce_dof[e];
    u[q2 % 3] += v;
    const double *yd = m.ydot + 3 * (int64_t)q2;
    u[3] += v * yd[0];
    u[4] += v * yd[1];
    u[5] += v * yd[2];
  }
  const double *yd2 = m.ydot2 + 9 * (int64_t)q;    // [p*3 + q'] of this component
  for (int d = 0; d < 6; ++d)
    for (int e = 0; e < 6; ++e) {
      double x = t[d] * u[e];
      if (d >= 3 && e >= 3) x += r * yd2[(d - 3) * 3 + (e - 3)];
      out[d * 6 + e] = x;
    }
}

// one workgroup per couple: fixed-order sums of its rows' partials into the point block
// and the point residual
template <typename TO>
__global__ __launch_bounds__(kBlock) void k_couple_reduce(fea_couple_map m, int with_k, double load_scale,
                                                          TO *__restrict__ val, double *__restrict__ R_red) {
  constexpr int NJ = 42, NS = kBlock / NJ;      // 6 slices of 42 outputs
  __shared__ double red[NS][NJ];
  const int k = blockIdx.x;
  const int q0 = 3 * m.cnode_ptr[k], q1 = 3 * m.cnode_ptr[k + 1];
  const int t = threadIdx.x;
  if (t < NS * NJ) {
    const int j = t % NJ, sl = t / NJ;
    double acc = 0.0;
    if (with_k || j >= 36)
      for (int q = q0 + sl; q < q1; q += NS) acc += m.partial[NJ * (int64_t)q + j];
    red[sl][j] = acc;
  }
  __syncthreads();
  if (t < NJ) {
    double acc = 0.0;
    for (int sl = 0; sl < NS; ++sl) acc += red[sl][t];
    if (t < 36) {
      const int s = m.rp_slot[36 * k + t];
      if (with_k && s >= 0) val[s] = (TO)acc;
    } else {
      const int d = t - 36;
      const int row = m.rp_row[6 * k + d];
      if (row >= 0) R_red[row] = acc - load_scale * m.rp_load[6 * k + d];
    }
  }
}

template <typename TO>
static int condense_t(const fea_couple_map *m, const double *K_ext, const double *R_ext, double load_scale,
                      TO *val, double *R_red, hipStream_t st) {
  const fea_couple_map M = *m;
  const int64_t n_ff = (K_ext != nullptr && M.nnz_red > M.n_red_nodes) ? M.nnz_red : M.n_red_nodes;
  if (n_ff > 0) k_couple_ff<TO><<<blocks_for(n_ff), kBlock, 0, st>>>(M, K_ext, R_ext, val, R_red);
  if (K_ext != nullptr && M.n_pairs > 0)
    k_couple_pairs<TO><<<blocks_for(M.n_pairs), kBlock, 0, st>>>(M, K_ext, val);
  if (M.n_cnodes > 0) k_couple_rows<<<blocks_for(3 * (int64_t)M.n_cnodes), kBlock, 0, st>>>(M, K_ext, R_ext);
  if (M.n_couples > 0)
    k_couple_reduce<TO><<<M.n_couples, kBlock, 0, st>>>(M, K_ext != nullptr, load_scale, val, R_red);
  FEA_CHECK_LAUNCH("fea_couple_condense");
  return FEA_OK;
}

}  // namespace fea

extern "C" {

int fea_couple_map_size(void) { return (int)sizeof(fea_couple_map); }

int fea_couple_kinematics(const fea_couple_map *m, double *state, int with_derivs, void *stream) {
  FEA_CHECK_ARG(m != nullptr && state != nullptr, "fea_couple_kinematics: null map or state");
  FEA_CHECK_ARG(m->n_cnodes >= 0 && m->n_couples >= 0, "fea_couple_kinematics: negative sizes");
  FEA_CHECK_ARG(!with_derivs || (m->ydot != nullptr && m->ydot2 != nullptr),
                "fea_couple_kinematics: derivative workspaces missing");
  if (m->n_cnodes == 0) return FEA_OK;
  fea::k_couple_kin<<<fea::blocks_for(m->n_cnodes), fea::kBlock, 0, fea::as_stream(stream)>>>(*m, state,
                                                                                            with_derivs);
  FEA_CHECK_LAUNCH("fea_couple_kinematics");
  return FEA_OK;
}

int fea_couple_condense(const fea_couple_map *m, int val_prec, const double *K_ext, const double *R_ext,
                        double load_scale, void *val_red, double *R_red, void *stream) {
  FEA_CHECK_ARG(m != nullptr && R_ext != nullptr && R_red != nullptr, "fea_couple_condense: null argument");
  FEA_CHECK_ARG(K_ext == nullptr || val_red != nullptr, "fea_couple_condense: tangent output missing");
  FEA_CHECK_ARG(m->n_couples == 0 || m->partial != nullptr, "fea_couple_condense: partial workspace missing");
  hipStream_t st = fea::as_stream(stream);
  if (val_prec == FEA_F64)
    return fea::condense_t<double>(m, K_ext, R_ext, load_scale, (double *)val_red, R_red, st);
  if (val_prec == FEA_F32)
    return fea::condense_t<float>(m, K_ext, R_ext, load_scale, (float *)val_red, R_red, st);
  fea::set_error("fea_couple_condense: unknown precision %d", val_prec);
  return FEA_ERR_ARG;
}

}  // extern "C"
