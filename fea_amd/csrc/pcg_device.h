// Device building blocks shared by the single-system PCG (pcg.hip) and the
// batched multi-system PCG (batched.hip): workspace layout, the node-grouped
// CSR row-group product and the deterministic partial-sum reductions.
#pragma once

#include "common.h"

namespace fea {


// ws layout (float64); partial rings as described in k_pcg_update / k_pcg_dir
enum : int {
  G1_MAX = 2048,  // max workgroups of the SpMV stage (p.q partials; 2048 measured best)
  PCG_G = 512,    // workgroups of the vector stages (r.z / b.z partials)
  W_ITER = 0, W_DONE = 1, W_RZ = 2, W_BZ = 3, W_G1 = 4,
  W_RZH = 5,      // r.z of iterations of parity 0 / 1, written by the direction stage's
                  // workgroup 0 (slot it & 1) and read from the next iteration on
                  // (slot (it - 1) & 1) instead of re-summing the partials
  PCG_RING = 2048, // most workgroups of the init / update stages (r.z, b.z partials per parity)
  W_PQ = 16, W_RZP = W_PQ + G1_MAX, W_BZP = W_RZP + 2 * PCG_RING, W_END = W_BZP + PCG_RING
};

// MODE 0: one row per group; MODE 1: node groups (rows share the column list);
// MODE 2: node groups whose column list is made of whole neighbour-node triples
// (3 consecutive columns per neighbour, 3 rows per group): lane j takes
// neighbour j's 3 columns with one index load, one 3-wide x load and three
// 3-wide value loads.  In MODE 2 `col` is the compact triple-column array: one entry
// (the first column) per neighbour triple of each group, group g's at
// rowptr[grp[g]] / 9 + j (every earlier group holds 3 rows of a multiple of 3 entries),
// so the index stream is a third of the CSR column array.
template <typename TM, typename TV, int SW, int MODE>
__device__ __forceinline__ int group_spmv(int g, int lane, const int32_t *__restrict__ grp,
                                          const int32_t *__restrict__ rowptr,
                                          const int32_t *__restrict__ col, const TM *__restrict__ val,
                                          const TV *__restrict__ x, TV out[3]) {
  const int r0 = MODE ? grp[g] : g;
  const int nr = MODE ? grp[g + 1] - r0 : 1;
  const int s0 = rowptr[r0];
  const int L = rowptr[r0 + 1] - s0;
  TV a0 = TV(0), a1 = TV(0), a2 = TV(0);
  if (MODE == 2) {
    const int nb = L / 3;
    for (int j = lane; j < nb; j += SW) {
      const int o = s0 + 3 * j;
      const int c = col[s0 / 9 + j];
      const TV x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
      a0 += (TV)val[o] * x0 + (TV)val[o + 1] * x1 + (TV)val[o + 2] * x2;
      a1 += (TV)val[o + L] * x0 + (TV)val[o + L + 1] * x1 + (TV)val[o + L + 2] * x2;
      a2 += (TV)val[o + 2 * L] * x0 + (TV)val[o + 2 * L + 1] * x1 + (TV)val[o + 2 * L + 2] * x2;
    }
  } else {
    for (int t = lane; t < L; t += SW) {
      const TV xv = x[col[s0 + t]];
      a0 += (TV)val[s0 + t] * xv;
      if (MODE == 1) {
        if (nr > 1) a1 += (TV)val[s0 + L + t] * xv;
        if (nr > 2) a2 += (TV)val[s0 + 2 * L + t] * xv;
      }
    }
  }
#pragma unroll
  for (int off = SW / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, SW);
    if (MODE) {
      a1 += __shfl_xor(a1, off, SW);
      a2 += __shfl_xor(a2, off, SW);
    }
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
  return r0;
}

// ---------------------------------------------------------------- node-block Jacobi
// M = blockdiag(K_gg) over the row groups g (the free rows of a node, nr <= 3): z = M^-1 r
// per group with the symmetric inverse stored as 6 floats (00, 01, 02, 11, 12, 22) -- the
// bytes of the FP64 diagonal it replaces.  Entries of rows / columns beyond nr are zero.
__device__ __forceinline__ void bj_apply(const float *__restrict__ B, const double r[3], double z[3]) {
  const double b00 = B[0], b01 = B[1], b02 = B[2], b11 = B[3], b12 = B[4], b22 = B[5];
  z[0] = b00 * r[0] + b01 * r[1] + b02 * r[2];
  z[1] = b01 * r[0] + b11 * r[1] + b12 * r[2];
  z[2] = b02 * r[0] + b12 * r[1] + b22 * r[2];
}

// B = the packed float inverse of the symmetric part of the nr x nr block A (entries beyond
// nr ignored); a block that is not positive definite (leading minors) falls back to the
// point-Jacobi inverse 1 / |A_ii| of its rows
__device__ inline void bj_invert_store(double A[3][3], int nr, float *__restrict__ B) {
  for (int a = 0; a < 3; ++a)
    for (int b = a + 1; b < 3; ++b) A[a][b] = A[b][a] = 0.5 * (A[a][b] + A[b][a]);
  for (int a = nr; a < 3; ++a) {
    for (int b = 0; b < 3; ++b) A[a][b] = A[b][a] = 0.0;
    A[a][a] = 1.0;  // pad: identity on absent rows
  }
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[1][2];
  const double c01 = A[0][2] * A[1][2] - A[0][1] * A[2][2];
  const double c02 = A[0][1] * A[1][2] - A[0][2] * A[1][1];
  const double c11 = A[0][0] * A[2][2] - A[0][2] * A[0][2];
  const double c12 = A[0][1] * A[0][2] - A[0][0] * A[1][2];
  const double c22 = A[0][0] * A[1][1] - A[0][1] * A[0][1];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const bool spd = A[0][0] > 0.0 && c22 > 0.0 && det > 0.0 && det == det;
  double o[6];
  if (spd) {
    const double id = 1.0 / det;
    o[0] = c00 * id; o[1] = c01 * id; o[2] = c02 * id; o[3] = c11 * id; o[4] = c12 * id; o[5] = c22 * id;
  } else {
    for (int k = 0; k < 6; ++k) o[k] = 0.0;
    const double d0 = fabs(A[0][0]), d1 = fabs(A[1][1]), d2 = fabs(A[2][2]);
    o[0] = d0 == 0.0 ? 1.0 : 1.0 / d0;
    o[3] = d1 == 0.0 ? 1.0 : 1.0 / d1;
    o[5] = d2 == 0.0 ? 1.0 : 1.0 / d2;
  }
  if (nr < 3) { o[2] = o[4] = o[5] = 0.0; }
  if (nr < 2) { o[1] = o[3] = 0.0; }
  for (int k = 0; k < 6; ++k) B[k] = (float)o[k];
}

template <typename TV>
__device__ __forceinline__ void bj_load(const TV *__restrict__ v, int r0, int nr, double o[3]) {
  o[0] = (double)v[r0];
  o[1] = nr > 1 ? (double)v[r0 + 1] : 0.0;
  o[2] = nr > 2 ? (double)v[r0 + 2] : 0.0;
}

// sum of g partials, identical in every workgroup (all threads get the result)
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int g) {
  __shared__ double res;
  double acc = 0.0;
  for (int b = threadIdx.x; b < g; b += blockDim.x) acc += part[b];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) res = s;
  __syncthreads();
  const double out = res;
  __syncthreads();
  return out;
}

__device__ __forceinline__ void write_partial(double *part, double v) {
  const double s = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

}  // namespace fea
