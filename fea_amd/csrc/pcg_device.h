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
