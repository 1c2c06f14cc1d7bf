// CSR SpMV and device-resident Jacobi-PCG (replaces the CSR `A @ p` and the CG
// loop of _linear_solver.py:4-42 on the scaled system of static_implicit.py:314-321).
//
// Precision is split between the matrix (TM: the HBM-heavy values) and the
// vectors / recurrences (TV): (double,double) plain FP64, (float,double) mixed
// (FP32 matrix storage, FP64 vectors and accumulation), (float,float) FP32.
//
// Node-grouped SpMV: in the free-DOF CSR every free row of a node has the same
// column list (columns come from the node's neighbours), and a node's free
// rows are consecutive.  With a group table grp[g] = first row of group g
// (rows grp[g] .. grp[g+1]-1, at most 3, all of length L), a SW-lane sub-wave
// reads each column index and gathers each x value once for all rows of the
// node, cutting index + gather traffic by up to 3x.  grp == NULL selects one
// row per group (plain CSR).
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <mutex>
#include <vector>

#include "pcg_device.h"

namespace fea {

// y = A x   (or y = b - A x when b != nullptr); grid-stride over groups
template <typename TM, typename TV, int SW, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(int ngroups, const int32_t *__restrict__ grp,
                                                 const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ col,
                                                 const TM *__restrict__ val, const TV *__restrict__ x,
                                                 const TV *__restrict__ b, TV *__restrict__ y) {
  constexpr int GPB = kBlock / SW;
  const int lane = threadIdx.x % SW;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / SW);
    if (g >= ngroups) continue;
    TV o[3];
    const int r0 = group_spmv<TM, TV, SW, MODE>(g, lane, grp, rowptr, col, val, x, o);
    if (lane == 0) {
      const int nr = MODE ? grp[g + 1] - r0 : 1;
      for (int j = 0; j < nr; ++j) y[r0 + j] = (b != nullptr) ? b[r0 + j] - o[j] : o[j];
    }
  }
}

// z = dinv r, p = z; partials of r.z (ring slot 0) and b.(dinv b)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const TV *__restrict__ dinv,
                                                     const TV *__restrict__ b,
                                                     const TV *__restrict__ r, TV *__restrict__ z,
                                                     TV *__restrict__ p, double *__restrict__ ws) {
  double rz = 0.0, bz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const TV zi = dinv[i] * r[i];
    z[i] = zi;
    p[i] = zi;
    rz += (double)r[i] * (double)zi;
    bz += (double)b[i] * (double)dinv[i] * (double)b[i];
  }
  write_partial(ws + W_RZP, rz);
  write_partial(ws + W_BZP, bz);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws[W_ITER] = 0.0;
    ws[W_DONE] = 0.0;
  }
}

// iteration i = ITER+1: q = A p and p.q partials; workgroup 0 advances ITER and
// records the SpMV grid (the number of p.q partials)
template <typename TM, typename TV, int SW, int MODE>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(int ngroups, const int32_t *__restrict__ grp,
                                                     const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col,
                                                     const TM *__restrict__ val,
                                                     const TV *__restrict__ p, TV *__restrict__ q,
                                                     double *__restrict__ ws) {
  if (ws[W_DONE] != 0.0) return;
  constexpr int GPB = kBlock / SW;
  const int lane = threadIdx.x % SW;
  double pq = 0.0;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / SW);
    if (g >= ngroups) continue;
    TV o[3];
    const int r0 = group_spmv<TM, TV, SW, MODE>(g, lane, grp, rowptr, col, val, p, o);
    if (lane == 0) {
      const int nr = MODE ? grp[g + 1] - r0 : 1;
      for (int j = 0; j < nr; ++j) {
        q[r0 + j] = o[j];
        pq += (double)o[j] * (double)p[r0 + j];
      }
    }
  }
  write_partial(ws + W_PQ, pq);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws[W_ITER] += 1.0;
    ws[W_G1] = (double)gridDim.x;
  }
}

// The vector stages are latency bound (a flag, the partial sums, then ~3 elements per
// thread): each thread issues the vector loads of its first kPre elements, which do not
// depend on the scalars, before the flag test and the partial sums.  Same arithmetic in
// the same order as a plain grid-stride loop.
constexpr int kPre = 4;

// alpha = rz_{i-1} / pq; x += alpha p; r -= alpha q; z = dinv r; r.z partials (ring i&1)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, const TV *__restrict__ dinv,
                                                       TV *__restrict__ x, TV *__restrict__ r,
                                                       TV *__restrict__ z, const TV *__restrict__ p,
                                                       const TV *__restrict__ q, double *__restrict__ ws) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  TV vx[kPre], vp[kPre], vr[kPre], vq[kPre], vd[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * stride;
    if (i < n) {
      vx[k] = x[i];
      vp[k] = p[i];
      vr[k] = r[i];
      vq[k] = q[i];
      vd[k] = dinv[i];
    }
  }
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  const int g1 = (int)ws[W_G1];
  // r.z of the previous iteration: stored by its direction stage (a re-sum of the same
  // partials in the same order gives the same bits); summed here only in iteration 1
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, gridDim.x);
  const double pq = sum_partials(ws + W_PQ, g1);
  const TV alpha = (TV)(pq != 0.0 ? rz_old / pq : 0.0);  // b = A x exactly: no step
  double rz = 0.0;
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * stride;
    if (i < n) {
      x[i] = vx[k] + alpha * vp[k];
      const TV ri = vr[k] - alpha * vq[k];
      r[i] = ri;
      const TV zi = vd[k] * ri;
      z[i] = zi;
      rz += (double)ri * (double)zi;
    }
  }
  for (int64_t i = i0 + kPre * stride; i < n; i += stride) {
    x[i] += alpha * p[i];
    const TV ri = r[i] - alpha * q[i];
    r[i] = ri;
    const TV zi = dinv[i] * ri;
    z[i] = zi;
    rz += (double)ri * (double)zi;
  }
  write_partial(ws + W_RZP + (it & 1) * PCG_RING, rz);
}

// convergence (rz_i / bz < tol, _linear_solver.py:37; r.z = 0 is an exact solution, which
// also ends a solve with b = 0 where the ratio is 0/0) and p = z + beta p
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, const TV *__restrict__ z, TV *__restrict__ p,
                                                    double *__restrict__ ws, double tol, int max_iter, int gu,
                                                    int gu_start) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  TV vz[kPre], vp[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * stride;
    if (i < n) {
      vz[k] = z[i];
      vp[k] = p[i];
    }
  }
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  // gu: the number of this iteration's r.z partials (workgroups of the stage that wrote them);
  // gu_start: that of the start stage's r.z and b.z partials (read in iteration 1)
  const double rz_new = sum_partials(ws + W_RZP + (it & 1) * PCG_RING, gu);
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, gu_start);
  // b.z is rewritten with the same value by every direction stage: from iteration 2 on
  // it is read back (workgroup 0 of this launch may be storing the identical bits)
  const double bz = it > 1 ? ws[W_BZ] : sum_partials(ws + W_BZP, gu_start);
  const bool done = (rz_new == 0.0) || (rz_new / bz < tol) || (it >= max_iter) || !(rz_new == rz_new);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws[W_RZ] = rz_new;
    ws[W_RZH + (it & 1)] = rz_new;
    ws[W_BZ] = bz;
    if (done) ws[W_DONE] = 1.0;
  }
  if (done) return;
  const TV beta = (TV)(rz_new / rz_old);
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * stride;
    if (i < n) p[i] = vz[k] + beta * vp[k];
  }
  for (int64_t i = i0 + kPre * stride; i < n; i += stride) p[i] = z[i] + beta * p[i];
}

// binv[g] = (K_gg)^-1 from the group's diagonal block: entry (a, b) of row r0 + a sits at
// diag_slot[r0 + a] + b - a when the row's sorted columns hold r0 .. r0 + nr - 1 (checked;
// absent entries are zero).  A block that is not positive definite falls back to the
// point-Jacobi inverse 1 / |K_ii| of its rows.
template <typename TM>
__global__ __launch_bounds__(kBlock) void k_block_jacobi_inv(int ng, const int32_t *__restrict__ grp,
                                                             const int32_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ col,
                                                             const int32_t *__restrict__ diag_slot,
                                                             const TM *__restrict__ val, float *__restrict__ binv) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int r0 = grp[g];
  const int nr = grp[g + 1] - r0;
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < nr; ++a) {
    const int d = diag_slot[r0 + a];
    if (d < 0) continue;
    for (int b = 0; b < nr; ++b) {
      const int sl = d + b - a;
      if (sl >= rowptr[r0 + a] && sl < rowptr[r0 + a + 1] && col[sl] == r0 + b) A[a][b] = (double)val[sl];
    }
  }
  bj_invert_store(A, nr, binv + 6 * (int64_t)g);
}

// block-Jacobi start stage: z = M^-1 r, p = z; partials of r.z (ring slot 0) and b.(M^-1 b)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pcg_init_bj(int ng, const int32_t *__restrict__ grp,
                                                        const float *__restrict__ binv, const TV *__restrict__ b,
                                                        const TV *__restrict__ r, TV *__restrict__ z,
                                                        TV *__restrict__ p, double *__restrict__ ws) {
  double rz = 0.0, bz = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    const int r0 = grp[g], nr = grp[g + 1] - r0;
    double rr[3], bb[3], zz[3], zb[3];
    bj_load(r, r0, nr, rr);
    bj_load(b, r0, nr, bb);
    bj_apply(binv + 6 * g, rr, zz);
    bj_apply(binv + 6 * g, bb, zb);
    for (int a = 0; a < nr; ++a) {
      const TV zi = (TV)zz[a];
      z[r0 + a] = zi;
      p[r0 + a] = zi;
      rz += rr[a] * (double)zi;
      bz += bb[a] * zb[a];
    }
  }
  write_partial(ws + W_RZP, rz);
  write_partial(ws + W_BZP, bz);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws[W_ITER] = 0.0;
    ws[W_DONE] = 0.0;
  }
}

// block-Jacobi update stage: as k_pcg_update with z = M^-1 r per row group.  With the
// vector grid (~one group per thread on a large system) each thread loads its first
// group's operands before the flag test and the partial sums, as k_pcg_update does.
template <typename TV>
__device__ __forceinline__ double bj_update_group(int r0, int nr, const float *__restrict__ B, TV alpha,
                                                  const TV vx[3], const TV vp[3], const TV vr[3], const TV vq[3],
                                                  TV *__restrict__ x, TV *__restrict__ r, TV *__restrict__ z) {
  double rr[3] = {0.0, 0.0, 0.0}, zz[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < nr) {
      x[r0 + a] = vx[a] + alpha * vp[a];
      const TV ri = vr[a] - alpha * vq[a];
      r[r0 + a] = ri;
      rr[a] = (double)ri;
    }
  }
  bj_apply(B, rr, zz);
  double rz = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < nr) {
      const TV zi = (TV)zz[a];
      z[r0 + a] = zi;
      rz += rr[a] * (double)zi;
    }
  }
  return rz;
}

template <typename TV>
__device__ __forceinline__ void bj_fetch(int64_t g, const int32_t *__restrict__ grp, const float *__restrict__ binv,
                                         const TV *__restrict__ x, const TV *__restrict__ r,
                                         const TV *__restrict__ p, const TV *__restrict__ q, int &r0, int &nr,
                                         float B[6], TV vx[3], TV vp[3], TV vr[3], TV vq[3]) {
  r0 = grp[g];
  nr = grp[g + 1] - r0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < nr) {
      vx[a] = x[r0 + a];
      vp[a] = p[r0 + a];
      vr[a] = r[r0 + a];
      vq[a] = q[r0 + a];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) B[k] = binv[6 * g + k];
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pcg_update_bj(int ng, const int32_t *__restrict__ grp,
                                                          const float *__restrict__ binv, TV *__restrict__ x,
                                                          TV *__restrict__ r, TV *__restrict__ z,
                                                          const TV *__restrict__ p, const TV *__restrict__ q,
                                                          double *__restrict__ ws) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t g0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int r0 = 0, nr = 0;
  float B[6];
  TV vx[3], vp[3], vr[3], vq[3];
  if (g0 < ng) bj_fetch(g0, grp, binv, x, r, p, q, r0, nr, B, vx, vp, vr, vq);
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  const int g1 = (int)ws[W_G1];
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, gridDim.x);
  const double pq = sum_partials(ws + W_PQ, g1);
  const TV alpha = (TV)(pq != 0.0 ? rz_old / pq : 0.0);
  double rz = 0.0;
  if (g0 < ng) rz += bj_update_group(r0, nr, B, alpha, vx, vp, vr, vq, x, r, z);
  for (int64_t g = g0 + stride; g < ng; g += stride) {
    bj_fetch(g, grp, binv, x, r, p, q, r0, nr, B, vx, vp, vr, vq);
    rz += bj_update_group(r0, nr, B, alpha, vx, vp, vr, vq, x, r, z);
  }
  write_partial(ws + W_RZP + (it & 1) * PCG_RING, rz);
}

// block-Jacobi update stage for triple-structured systems (every row group 3 rows: row i is
// row i % 3 of group i / 3): one thread per row, coalesced like k_pcg_update, a block
// stepping over RB = 255 rows (whole groups); the three new residuals of a group meet in
// LDS for z = M^-1 r.  The first kPre steps are loaded before the flag and the partial sums.
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pcg_update_bj3(int n, const float *__restrict__ binv,
                                                           TV *__restrict__ x, TV *__restrict__ r,
                                                           TV *__restrict__ z, const TV *__restrict__ p,
                                                           const TV *__restrict__ q, double *__restrict__ ws) {
  constexpr int RB = kBlock - kBlock % 3;
  __shared__ double sr[kPre][kBlock];
  const int t = threadIdx.x;
  const int a = t % 3;                      // row of the group (RB and the block origin are multiples of 3)
  // packed symmetric inverse (00, 01, 02, 11, 12, 22): row a is (ia0, ia1, ia2)
  const int ia0 = a;
  const int ia1 = a == 0 ? 1 : (a == 1 ? 3 : 4);
  const int ia2 = a == 0 ? 2 : (a == 1 ? 4 : 5);
  const bool lane = t < RB;
  const int64_t step = (int64_t)gridDim.x * RB;
  const int64_t i0 = (int64_t)blockIdx.x * RB + t;
  TV vx[kPre], vp[kPre], vr[kPre], vq[kPre];
  float b0[kPre], b1[kPre], b2[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * step;
    if (lane && i < n) {
      vx[k] = x[i];
      vp[k] = p[i];
      vr[k] = r[i];
      vq[k] = q[i];
      const float *B = binv + 6 * (i / 3);
      b0[k] = B[ia0];
      b1[k] = B[ia1];
      b2[k] = B[ia2];
    }
  }
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  const int g1 = (int)ws[W_G1];
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, gridDim.x);
  const double pq = sum_partials(ws + W_PQ, g1);
  const TV alpha = (TV)(pq != 0.0 ? rz_old / pq : 0.0);
  double rz = 0.0;
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * step;
    if (lane && i < n) {
      x[i] = vx[k] + alpha * vp[k];
      vr[k] = vr[k] - alpha * vq[k];
      r[i] = vr[k];
      sr[k][t] = (double)vr[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t i = i0 + k * step;
    if (lane && i < n) {
      const int tb = t - a;
      const TV zi = (TV)((double)b0[k] * sr[k][tb] + (double)b1[k] * sr[k][tb + 1] + (double)b2[k] * sr[k][tb + 2]);
      z[i] = zi;
      rz += (double)vr[k] * (double)zi;
    }
  }
  // rows beyond the prefetched steps (systems above kPre * gridDim.x * RB rows)
  for (int64_t base = (int64_t)kPre * step; base < n; base += step) {
    __syncthreads();                        // sr[0] of the previous step has been read
    const int64_t i = base + (int64_t)blockIdx.x * RB + t;
    const bool on = lane && i < n;
    TV ri = TV(0);
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    if (on) {
      x[i] += alpha * p[i];
      ri = r[i] - alpha * q[i];
      r[i] = ri;
      sr[0][t] = (double)ri;
      const float *B = binv + 6 * (i / 3);
      c0 = B[ia0];
      c1 = B[ia1];
      c2 = B[ia2];
    }
    __syncthreads();
    if (on) {
      const int tb = t - a;
      const TV zi = (TV)((double)c0 * sr[0][tb] + (double)c1 * sr[0][tb + 1] + (double)c2 * sr[0][tb + 2]);
      z[i] = zi;
      rz += (double)ri * (double)zi;
    }
  }
  write_partial(ws + W_RZP + (it & 1) * PCG_RING, rz);
}

template <typename TM, typename TV>
__global__ __launch_bounds__(kBlock) void k_jacobi_inv(int n, const int32_t *__restrict__ diag_slot,
                                                       const TM *__restrict__ val, TV *__restrict__ dinv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = diag_slot[r];
  const TV d = (s >= 0) ? fabs((TV)val[s]) : TV(0);
  dinv[r] = (d == TV(0)) ? TV(1) : TV(1) / d;
}

// ------------------------------------------------------------------ host side
struct SpmvShape {
  int sw;
  int mode;
  int ngroups;
};

// n_groups < 0 marks a triple-structured group table (every group: 3 rows whose
// column list is whole neighbour triples; col is then the compact triple-column
// array); |n_groups| is the group count.
static SpmvShape shape_of(int n, int64_t nnz, int n_groups, const int32_t *grp) {
  SpmvShape s;
  s.mode = (grp == nullptr) ? 0 : (n_groups < 0 ? 2 : 1);
  s.ngroups = (grp == nullptr) ? n : (n_groups < 0 ? -n_groups : n_groups);
  const double len = n > 0 ? (double)nnz / n : 0.0;  // row length
  const double work = (s.mode == 2) ? len / 3.0 : len;  // lane work items per group
  s.sw = work > 48 ? 32 : (work > 24 ? 16 : (work > 12 ? 8 : 4));
  // triple mode: a 16-lane sub-wave up to 32 neighbour triples per node (27 for a
  // brick mesh: two passes) measured 9 % faster per CG iteration than 32 lanes
  if (s.mode == 2) s.sw = work > 32 ? 32 : (work > 8 ? 16 : 8);
  return s;
}

static int spmv_grid(const SpmvShape &s) {
  const int64_t g = ((int64_t)s.ngroups * s.sw + kBlock - 1) / kBlock;
  return (int)(g < 1 ? 1 : (g > G1_MAX ? G1_MAX : g));
}

#define FEA_SW_SWITCH(MACRO)                                                         \
  do {                                                                               \
    if (s.mode == 2) {                                                               \
      switch (s.sw) { case 8: MACRO(8, 2); break; case 16: MACRO(16, 2); break;      \
                      default: MACRO(32, 2); }                                       \
    } else if (s.mode == 1) {                                                        \
      switch (s.sw) { case 4: MACRO(4, 1); break; case 8: MACRO(8, 1); break;        \
                      case 16: MACRO(16, 1); break; default: MACRO(32, 1); }         \
    } else {                                                                         \
      switch (s.sw) { case 4: MACRO(4, 0); break; case 8: MACRO(8, 0); break;        \
                      case 16: MACRO(16, 0); break; default: MACRO(32, 0); }         \
    }                                                                                \
  } while (0)

template <typename TM, typename TV>
static void launch_spmv(const SpmvShape &s, const int32_t *grp, const int32_t *rp, const int32_t *c,
                        const TM *v, const TV *x, const TV *b, TV *y, hipStream_t st) {
  const int grid = spmv_grid(s);
#define FEA_SPMV(SW, MD) k_spmv<TM, TV, SW, MD><<<grid, kBlock, 0, st>>>(s.ngroups, grp, rp, c, v, x, b, y)
  FEA_SW_SWITCH(FEA_SPMV);
#undef FEA_SPMV
}

template <typename TM, typename TV>
static void launch_pcg_spmv(const SpmvShape &s, const int32_t *grp, const int32_t *rp, const int32_t *c,
                            const TM *v, const TV *p, TV *q, double *ws, hipStream_t st) {
  const int grid = spmv_grid(s);
#define FEA_PSPMV(SW, MD) k_pcg_spmv<TM, TV, SW, MD><<<grid, kBlock, 0, st>>>(s.ngroups, grp, rp, c, v, p, q, ws)
  FEA_SW_SWITCH(FEA_PSPMV);
#undef FEA_PSPMV
}

// workgroups of the init / update stages (their partial count): a sweep of 256 .. 2048 on
// the 418 176-row beam measured 512 best (within 4 % across the range)
static int vec_grid() { return PCG_G; }

struct PcgArgs {
  int n;
  int64_t nnz;
  int n_groups;
  const int32_t *grp, *rowptr, *col;
  const void *val, *dinv, *b;
  void *x, *r, *z, *p, *q;
  int precond;  // 0: point Jacobi (dinv: vec_dtype[n]); 1: node-block Jacobi (dinv: float[6 n_groups])
  double *ws;
  double tol;
  int max_iter, iters;
  double *status_host;
  hipStream_t st;
};

template <typename TM, typename TV>
static int pcg_start_t(const PcgArgs &a) {
  const SpmvShape s = shape_of(a.n, a.nnz, a.n_groups, a.grp);
  launch_spmv<TM, TV>(s, a.grp, a.rowptr, a.col, (const TM *)a.val, (const TV *)a.x, (const TV *)a.b,
                      (TV *)a.r, a.st);
  if (a.precond == 1)
    k_pcg_init_bj<TV><<<vec_grid(), kBlock, 0, a.st>>>(s.ngroups, a.grp, (const float *)a.dinv, (const TV *)a.b,
                                                       (const TV *)a.r, (TV *)a.z, (TV *)a.p, a.ws);
  else
    k_pcg_init<TV><<<vec_grid(), kBlock, 0, a.st>>>(a.n, (const TV *)a.dinv, (const TV *)a.b, (const TV *)a.r,
                                                    (TV *)a.z, (TV *)a.p, a.ws);
  FEA_CHECK_LAUNCH("fea_pcg_start");
  return FEA_OK;
}

template <typename TM, typename TV>
static int pcg_iterate_t(const PcgArgs &a) {
  const SpmvShape s = shape_of(a.n, a.nnz, a.n_groups, a.grp);
  const int gu = vec_grid();
  for (int it = 0; it < a.iters; ++it) {
    launch_pcg_spmv<TM, TV>(s, a.grp, a.rowptr, a.col, (const TM *)a.val, (const TV *)a.p, (TV *)a.q,
                            a.ws, a.st);
    if (a.precond == 1 && s.mode == 2)
      k_pcg_update_bj3<TV><<<gu, kBlock, 0, a.st>>>(a.n, (const float *)a.dinv, (TV *)a.x, (TV *)a.r, (TV *)a.z,
                                                   (const TV *)a.p, (const TV *)a.q, a.ws);
    else if (a.precond == 1)
      k_pcg_update_bj<TV><<<gu, kBlock, 0, a.st>>>(s.ngroups, a.grp, (const float *)a.dinv, (TV *)a.x, (TV *)a.r,
                                                  (TV *)a.z, (const TV *)a.p, (const TV *)a.q, a.ws);
    else
      k_pcg_update<TV><<<gu, kBlock, 0, a.st>>>(a.n, (const TV *)a.dinv, (TV *)a.x, (TV *)a.r, (TV *)a.z,
                                               (const TV *)a.p, (const TV *)a.q, a.ws);
    k_pcg_dir<TV><<<PCG_G, kBlock, 0, a.st>>>(a.n, (const TV *)a.z, (TV *)a.p, a.ws, a.tol, a.max_iter, gu, gu);
  }
  FEA_CHECK_LAUNCH("fea_pcg_iterate");
  if (a.status_host != nullptr) {
    double h[4];
    FEA_HIP(hipMemcpyAsync(h, a.ws, sizeof(h), hipMemcpyDeviceToHost, a.st), "fea_pcg_iterate");
    FEA_HIP(hipStreamSynchronize(a.st), "fea_pcg_iterate");
    for (int i = 0; i < 4; ++i) a.status_host[i] = h[i];
  }
  return FEA_OK;
}

static bool valid_combo(int m, int v) {
  return (m == FEA_F64 && v == FEA_F64) || (m == FEA_F32 && v == FEA_F64) || (m == FEA_F32 && v == FEA_F32);
}

}  // namespace fea

using namespace fea;

#define FEA_DISPATCH(m, v, fn, arg)                              \
  ((m) == FEA_F64 ? fn<double, double>(arg)                      \
                  : ((v) == FEA_F64 ? fn<float, double>(arg) : fn<float, float>(arg)))

extern "C" int64_t fea_pcg_ws_size(int) { return (int64_t)W_END; }

extern "C" int fea_spmv(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                        const int32_t *rowptr, const int32_t *col, const void *val, const void *x, void *y,
                        void *stream) {
  FEA_CHECK_ARG(valid_combo(mat_dtype, vec_dtype), "fea_spmv: unsupported precision pair");
  FEA_CHECK_ARG(n >= 0, "fea_spmv: bad sizes");
  if (n == 0) return FEA_OK;
  FEA_CHECK_ARG(rowptr && col && val && x && y, "fea_spmv: null argument");
  const SpmvShape s = shape_of(n, nnz, n_groups, grp);
  hipStream_t st = as_stream(stream);
  if (mat_dtype == FEA_F64)
    launch_spmv<double, double>(s, grp, rowptr, col, (const double *)val, (const double *)x, nullptr, (double *)y, st);
  else if (vec_dtype == FEA_F64)
    launch_spmv<float, double>(s, grp, rowptr, col, (const float *)val, (const double *)x, nullptr, (double *)y, st);
  else
    launch_spmv<float, float>(s, grp, rowptr, col, (const float *)val, (const float *)x, nullptr, (float *)y, st);
  FEA_CHECK_LAUNCH("fea_spmv");
  return FEA_OK;
}

extern "C" int fea_block_jacobi_inv(int mat_dtype, int n, int n_groups, const int32_t *grp, const int32_t *rowptr,
                                    const int32_t *col, const int32_t *diag_slot, const void *val, float *binv,
                                    void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_block_jacobi_inv: bad matrix dtype");
  const int ng = n_groups < 0 ? -n_groups : n_groups;
  FEA_CHECK_ARG(n >= 0 && ng >= 0, "fea_block_jacobi_inv: bad sizes");
  if (ng == 0) return FEA_OK;
  FEA_CHECK_ARG(grp && rowptr && col && diag_slot && val && binv, "fea_block_jacobi_inv: null argument");
  hipStream_t st = as_stream(stream);
  const int g = blocks_for(ng);
  if (mat_dtype == FEA_F64)
    k_block_jacobi_inv<double><<<g, kBlock, 0, st>>>(ng, grp, rowptr, col, diag_slot, (const double *)val, binv);
  else
    k_block_jacobi_inv<float><<<g, kBlock, 0, st>>>(ng, grp, rowptr, col, diag_slot, (const float *)val, binv);
  FEA_CHECK_LAUNCH("fea_block_jacobi_inv");
  return FEA_OK;
}

extern "C" int fea_jacobi_inv(int mat_dtype, int vec_dtype, int n, const int32_t *diag_slot, const void *val,
                              void *dinv, void *stream) {
  FEA_CHECK_ARG(valid_combo(mat_dtype, vec_dtype), "fea_jacobi_inv: unsupported precision pair");
  FEA_CHECK_ARG(n >= 0, "fea_jacobi_inv: bad n");
  if (n == 0) return FEA_OK;
  FEA_CHECK_ARG(diag_slot && val && dinv, "fea_jacobi_inv: null argument");
  hipStream_t st = as_stream(stream);
  const int g = blocks_for(n);
  if (mat_dtype == FEA_F64)
    k_jacobi_inv<double, double><<<g, kBlock, 0, st>>>(n, diag_slot, (const double *)val, (double *)dinv);
  else if (vec_dtype == FEA_F64)
    k_jacobi_inv<float, double><<<g, kBlock, 0, st>>>(n, diag_slot, (const float *)val, (double *)dinv);
  else
    k_jacobi_inv<float, float><<<g, kBlock, 0, st>>>(n, diag_slot, (const float *)val, (float *)dinv);
  FEA_CHECK_LAUNCH("fea_jacobi_inv");
  return FEA_OK;
}

extern "C" int fea_pcg_start(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups,
                             const int32_t *grp, const int32_t *rowptr, const int32_t *col, const void *val,
                             int precond, const void *dinv, const void *b, void *x, void *r, void *z, void *p,
                             double *ws, void *stream) {
  FEA_CHECK_ARG(valid_combo(mat_dtype, vec_dtype), "fea_pcg_start: unsupported precision pair");
  FEA_CHECK_ARG(n > 0 && rowptr && col && val && dinv && b && x && r && z && p && ws,
                "fea_pcg_start: bad arguments");
  FEA_CHECK_ARG(precond == 0 || (precond == 1 && grp != nullptr && n_groups != 0),
                "fea_pcg_start: block Jacobi needs the row groups");
  PcgArgs a{n, nnz, n_groups, grp, rowptr, col, val, dinv, b, x, r, z, p, nullptr, precond, ws, 0.0, 0, 0,
            nullptr, as_stream(stream)};
  return FEA_DISPATCH(mat_dtype, vec_dtype, pcg_start_t, a);
}

extern "C" int fea_pcg_iterate(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups,
                               const int32_t *grp, const int32_t *rowptr, const int32_t *col,
                               const void *val, int precond, const void *dinv, void *x, void *r, void *z,
                               void *p, void *q, double *ws, double tol, int max_iter, int iters,
                               double *status_host, void *stream) {
  FEA_CHECK_ARG(valid_combo(mat_dtype, vec_dtype), "fea_pcg_iterate: unsupported precision pair");
  FEA_CHECK_ARG(n > 0 && iters >= 0 && max_iter >= 1 && rowptr && col && val && dinv && x && r && z && p &&
                    q && ws,
                "fea_pcg_iterate: bad arguments");
  FEA_CHECK_ARG(precond == 0 || (precond == 1 && grp != nullptr && n_groups != 0),
                "fea_pcg_iterate: block Jacobi needs the row groups");
  PcgArgs a{n, nnz, n_groups, grp, rowptr, col, val, dinv, nullptr, x, r, z, p, q, precond, ws, tol, max_iter,
            iters, status_host, as_stream(stream)};
  return FEA_DISPATCH(mat_dtype, vec_dtype, pcg_iterate_t, a);
}

extern "C" int fea_pcg_spmv(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                            const int32_t *rowptr, const int32_t *col, const void *val, const void *p, void *q,
                            double *ws, void *stream) {
  FEA_CHECK_ARG(valid_combo(mat_dtype, vec_dtype), "fea_pcg_spmv: unsupported precision pair");
  FEA_CHECK_ARG(n > 0 && rowptr && col && val && p && q && ws, "fea_pcg_spmv: bad arguments");
  const SpmvShape s = shape_of(n, nnz, n_groups, grp);
  hipStream_t st = as_stream(stream);
  if (mat_dtype == FEA_F64)
    launch_pcg_spmv<double, double>(s, grp, rowptr, col, (const double *)val, (const double *)p, (double *)q, ws, st);
  else if (vec_dtype == FEA_F64)
    launch_pcg_spmv<float, double>(s, grp, rowptr, col, (const float *)val, (const double *)p, (double *)q, ws, st);
  else
    launch_pcg_spmv<float, float>(s, grp, rowptr, col, (const float *)val, (const float *)p, (float *)q, ws, st);
  FEA_CHECK_LAUNCH("fea_pcg_spmv");
  return FEA_OK;
}

// ---------------------------------------------------------------- allocation-free PCG
// The PCG workspace (scalars and the partial-sum rings, W_END doubles = 65 664 B) of
// fea_pcg_run_noalloc lives in this file-scope device array: one copy per GPU, held
// until the process ends.  Host code never passes `scratch` itself to a kernel: the
// array's device address is looked up once per device with hipGetSymbolAddress and
// kept in g_scratch_dev.
namespace fea {
constexpr int kNoallocWsDoubles = W_END;
constexpr int kMaxDevices = 64;
}  // namespace fea

static __device__ __align__(256) double scratch[fea::kNoallocWsDoubles];

namespace fea {
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

// the current device's copy of `scratch`, looked up on the first call on that device
static int scratch_for_current_device(double **out) {
  int dev = -1;
  FEA_HIP(hipGetDevice(&dev), "fea_pcg_run_noalloc: hipGetDevice");
  FEA_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "fea_pcg_run_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void *ptr = nullptr;
    FEA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "fea_pcg_run_noalloc: hipGetSymbolAddress");
    FEA_CHECK_ARG(ptr != nullptr, "fea_pcg_run_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_dev[dev];
  return FEA_OK;
}

// zero the scalar header of the workspace (iteration counter, done flag, stored r.z):
// a fill kernel rather than a memset so that a captured graph replays it correctly
__global__ void k_ws_clear_header(double *__restrict__ ws) {
  if (threadIdx.x < W_PQ) ws[threadIdx.x] = 0.0;
}
}  // namespace fea

extern "C" int fea_pcg_run_noalloc(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups,
                                   const int32_t *grp, const int32_t *rowptr, const int32_t *col,
                                   const void *val, int precond, const void *dinv, const void *b, void *x,
                                   void *r, void *z, void *p, void *q, double tol, int max_iter, int iters,
                                   double *status_host, void *stream) {
  FEA_CHECK_ARG(valid_combo(mat_dtype, vec_dtype), "fea_pcg_run_noalloc: unsupported precision pair");
  FEA_CHECK_ARG(n > 0 && iters >= 0 && max_iter >= 1 && rowptr && col && val && dinv && x && r && z && p && q,
                "fea_pcg_run_noalloc: bad arguments");
  FEA_CHECK_ARG(precond == 0 || (precond == 1 && grp != nullptr && n_groups != 0),
                "fea_pcg_run_noalloc: block Jacobi needs the row groups");
  FEA_CHECK_ARG(fea_pcg_ws_size(n) <= (int64_t)kNoallocWsDoubles,
                "fea_pcg_run_noalloc: n = %d needs %lld workspace doubles, the static scratch holds %d", n,
                (long long)fea_pcg_ws_size(n), kNoallocWsDoubles);
  double *ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != FEA_OK) return rc;
  hipStream_t st = as_stream(stream);
  if (b != nullptr) {
    k_ws_clear_header<<<1, 64, 0, st>>>(ws);
    FEA_CHECK_LAUNCH("fea_pcg_run_noalloc");
    PcgArgs a0{n, nnz, n_groups, grp, rowptr, col, val, dinv, b, x, r, z, p, nullptr, precond, ws, 0.0, 0, 0,
               nullptr, st};
    const int rs = FEA_DISPATCH(mat_dtype, vec_dtype, pcg_start_t, a0);
    if (rs != FEA_OK) return rs;
  }
  PcgArgs a{n, nnz, n_groups, grp, rowptr, col, val, dinv, nullptr, x, r, z, p, q, precond, ws, tol, max_iter,
            iters, status_host, st};
  return FEA_DISPATCH(mat_dtype, vec_dtype, pcg_iterate_t, a);
}

// ---------------------------------------------------------------- smoothed-aggregation AMG
// V-cycle of the hierarchy built on the host (amg_setup.cpp).  Level 0 is the PCG's own
// matrix (node-grouped CSR, current values) smoothed with its node-block Jacobi inverses;
// levels >= 1 hold the Galerkin operators as CSR whose rows come in groups of 6 (one
// aggregate's coarse DOFs) sharing one column list, with 6x6 block inverses; the coarsest
// level is solved with its dense inverse.  The smoother is Chebyshev of D^-1 A on
// [lambda_max / 30, 1.1 lambda_max], the same polynomial before and after the coarse
// correction, so the V-cycle is a symmetric positive definite preconditioner for CG.
// Operators, prolongators, smoother blocks and the coarse inverse are stored in float;
// every vector and every accumulation is float64: rounding the preconditioner's input
// residual to float costs ~20 % more CG iterations on stiff systems (the error of a
// residual dominated by stiff components lands on the smooth part of M^-1 r), while float
// storage of the operators is a fixed symmetric perturbation that costs none.  Every
// kernel of the iteration path returns at once when the solve is marked converged.
namespace fea {

constexpr int AMG_SW = 64;  // lanes per 6-row group / per aggregate in the coarse-level kernels
typedef _Float16 half_t;    // prolongator values: IEEE binary16

static int amg_grid(int64_t items, int per_block) {
  const int64_t g = (items + per_block - 1) / per_block;
  return (int)(g < 1 ? 1 : (g > G1_MAX ? G1_MAX : g));
}

// o[m] = (A x)[6 g + m] for a level >= 1 operator: the group's 6 rows are consecutive CSR
// rows of equal length L sharing the column list, row m at s0 + m L
template <int SW>
__device__ __forceinline__ void group6_spmv(int g, int lane, const int32_t *__restrict__ rowptr,
                                            const int32_t *__restrict__ col, const float *__restrict__ val,
                                            const double *__restrict__ x, double o[6]) {
  const int s0 = rowptr[6 * g];
  const int L = rowptr[6 * g + 1] - s0;
  double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = lane; t < L; t += SW) {
    const double xv = x[col[s0 + t]];
#pragma unroll
    for (int m = 0; m < 6; ++m) a[m] += (double)val[s0 + m * L + t] * xv;
  }
#pragma unroll
  for (int off = SW / 2; off > 0; off >>= 1)
#pragma unroll
    for (int m = 0; m < 6; ++m) a[m] += __shfl_xor(a[m], off, SW);
#pragma unroll
  for (int m = 0; m < 6; ++m) o[m] = a[m];
}

__device__ __forceinline__ double half_lo(uint32_t w) { return (double)(float)__builtin_bit_cast(half_t, (uint16_t)(w & 0xffffu)); }
__device__ __forceinline__ double half_hi(uint32_t w) { return (double)(float)__builtin_bit_cast(half_t, (uint16_t)(w >> 16)); }

// group_spmv (triple mode) on the quads of k_to_half: row r's triple j of group g is quad
// rowptr[r0] / 3 + r * nb + j
template <int SW>
__device__ __forceinline__ void group_spmv_h4(int g, int lane, const int32_t *__restrict__ grp,
                                              const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                              const uint2 *__restrict__ q4, const double *__restrict__ x,
                                              double out[3]) {
  const int r0 = grp[g];
  const int s0 = rowptr[r0];
  const int nb = (rowptr[r0 + 1] - s0) / 3;
  const uint2 *base = q4 + s0 / 3;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int j = lane; j < nb; j += SW) {
    const int c = col[s0 / 9 + j];
    const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
    const uint2 u = base[j], v = base[nb + j], w = base[2 * nb + j];
    a0 += half_lo(u.x) * x0 + half_hi(u.x) * x1 + half_lo(u.y) * x2;
    a1 += half_lo(v.x) * x0 + half_hi(v.x) * x1 + half_lo(v.y) * x2;
    a2 += half_lo(w.x) * x0 + half_hi(w.x) * x1 + half_lo(w.y) * x2;
  }
#pragma unroll
  for (int off = SW / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, SW);
    a1 += __shfl_xor(a1, off, SW);
    a2 += __shfl_xor(a2, off, SW);
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
}

// the level-0 product of a group: the matrix's own values, or (TM = uint2) its binary16 quads
template <typename TM, int SW, int MODE>
__device__ __forceinline__ int level0_spmv(int g, int lane, const int32_t *__restrict__ grp,
                                           const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                           const TM *__restrict__ val, const double *__restrict__ x, double o[3]) {
  if constexpr (std::is_same<TM, uint2>::value) {
    static_assert(MODE == 2, "binary16 quads exist for triple-structured matrices only");
    group_spmv_h4<SW>(g, lane, grp, rowptr, col, val, x, o);
    return grp[g];
  } else {
    return group_spmv<TM, double, SW, MODE>(g, lane, grp, rowptr, col, val, x, o);
  }
}

__device__ __forceinline__ bool amg_done(const double *done) { return done != nullptr && *done != 0.0; }

// one Chebyshev step on level 0: r = b - A x (FIRST: x = 0, no product), d = c1 d + c2 M^-1 r,
// xo = x + d with M the node-block diagonal (6 packed floats per group).  RZ (the last
// post-smoothing step of a PCG iteration's V-cycle, xo = z, b = the PCG residual): also the
// workgroup's partial of b.xo = r.z into ring slot (iteration & 1) of the PCG workspace rz_ws.
// vscale: for binary16 values, the factor that undoes their scaling (vscale[1]); else null.
template <typename TM, int SW, int MODE, int FIRST, int RZ = 0>
__global__ __launch_bounds__(kBlock) void k_amg_cheb0(int ngroups, const int32_t *__restrict__ grp,
                                                      const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const TM *__restrict__ val,
                                                      const float *__restrict__ binv, const double *__restrict__ b,
                                                      const double *__restrict__ x, double *__restrict__ xo,
                                                      double *__restrict__ d, double c1, double c2,
                                                      const double *__restrict__ done, double *__restrict__ rz_ws,
                                                      const double *__restrict__ vscale) {
  if (amg_done(done)) return;
  constexpr int GPB = kBlock / SW;
  const int lane = threadIdx.x % SW;
  const double vs = vscale != nullptr ? vscale[1] : 1.0;
  double rz = 0.0;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / SW);
    if (g >= ngroups) continue;
    double o[3] = {0.0, 0.0, 0.0};
    if (!FIRST) {
      level0_spmv<TM, SW, MODE>(g, lane, grp, rowptr, col, val, x, o);
      o[0] *= vs;
      o[1] *= vs;
      o[2] *= vs;
    }
    // every lane holds the group's products (butterfly sums): lane a updates row a
    const int r0 = grp[g], nr = grp[g + 1] - r0;
    if (lane < nr) {
      double rr[3] = {0.0, 0.0, 0.0}, w[3];
      for (int a = 0; a < nr; ++a) rr[a] = b[r0 + a] - o[a];
      bj_apply(binv + 6 * (int64_t)g, rr, w);
      const int a = lane;
      const double dn = (c1 != 0.0 ? c1 * d[r0 + a] : 0.0) + c2 * w[a];
      d[r0 + a] = dn;
      const double xn = (FIRST ? 0.0 : x[r0 + a]) + dn;
      xo[r0 + a] = xn;
      if (RZ) rz += b[r0 + a] * xn;
    }
  }
  if (RZ) {
    const int it = (int)rz_ws[W_ITER];
    write_partial(rz_ws + W_RZP + (it & 1) * PCG_RING, rz);
  }
}

// r = b - A x on level 0
template <typename TM, int SW, int MODE>
__global__ __launch_bounds__(kBlock) void k_amg_resid0(int ngroups, const int32_t *__restrict__ grp,
                                                       const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col, const TM *__restrict__ val,
                                                       const double *__restrict__ b, const double *__restrict__ x,
                                                       double *__restrict__ r, const double *__restrict__ done,
                                                       const double *__restrict__ vscale) {
  if (amg_done(done)) return;
  constexpr int GPB = kBlock / SW;
  const int lane = threadIdx.x % SW;
  const double vs = vscale != nullptr ? vscale[1] : 1.0;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / SW);
    if (g >= ngroups) continue;
    double o[3];
    const int r0 = level0_spmv<TM, SW, MODE>(g, lane, grp, rowptr, col, val, x, o);
    if (lane < grp[g + 1] - r0) r[r0 + lane] = b[r0 + lane] - vs * (lane == 0 ? o[0] : (lane == 1 ? o[1] : o[2]));
  }
}

// the same two kernels on a level >= 1 (6-row groups, 36-float block inverses)
template <int FIRST>
__global__ __launch_bounds__(kBlock) void k_amg_cheb_c(int ngroups, const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col, const float *__restrict__ val,
                                                       const float *__restrict__ dinv, const double *__restrict__ b,
                                                       const double *__restrict__ x, double *__restrict__ xo,
                                                       double *__restrict__ d, double c1, double c2,
                                                       const double *__restrict__ done) {
  if (amg_done(done)) return;
  constexpr int GPB = kBlock / AMG_SW;
  const int lane = threadIdx.x % AMG_SW;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / AMG_SW);
    if (g >= ngroups) continue;
    double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!FIRST) group6_spmv<AMG_SW>(g, lane, rowptr, col, val, x, o);
    if (lane < 6) {  // lane m updates row m of the group
      double rr[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) rr[k] = b[6 * g + k] - o[k];
      const int m = lane;
      const float *D = dinv + 36 * (int64_t)g + 6 * m;
      double w = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) w += (double)D[k] * rr[k];
      const double dn = (c1 != 0.0 ? c1 * d[6 * g + m] : 0.0) + c2 * w;
      d[6 * g + m] = dn;
      xo[6 * g + m] = (FIRST ? 0.0 : x[6 * g + m]) + dn;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_amg_resid_c(int ngroups, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col,
                                                        const float *__restrict__ val, const double *__restrict__ b,
                                                        const double *__restrict__ x, double *__restrict__ r,
                                                        const double *__restrict__ done) {
  if (amg_done(done)) return;
  constexpr int GPB = kBlock / AMG_SW;
  const int lane = threadIdx.x % AMG_SW;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / AMG_SW);
    if (g >= ngroups) continue;
    double o[6];
    group6_spmv<AMG_SW>(g, lane, rowptr, col, val, x, o);
    if (lane < 6) {
      double om = o[0];
#pragma unroll
      for (int m = 1; m < 6; ++m)
        if (lane == m) om = o[m];
      r[6 * g + lane] = b[6 * g + lane] - om;
    }
  }
}

// bc = P^T r: one wave per coarse aggregate over its transposed entries, whose 6 values
// (ptval, binary16) are stored in entry order so a wave reads them contiguously
__global__ __launch_bounds__(kBlock) void k_amg_restrict(int nagg, const int32_t *__restrict__ pt_ptr,
                                                         const int32_t *__restrict__ pt_row,
                                                         const half_t *__restrict__ ptval,
                                                         const double *__restrict__ r, double *__restrict__ bc,
                                                         const double *__restrict__ done) {
  if (amg_done(done)) return;
  constexpr int GPB = kBlock / AMG_SW;
  const int lane = threadIdx.x % AMG_SW;
  for (int64_t a0 = (int64_t)blockIdx.x * GPB; a0 < nagg; a0 += (int64_t)gridDim.x * GPB) {
    const int I = (int)(a0 + threadIdx.x / AMG_SW);
    if (I >= nagg) continue;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int t = pt_ptr[I] + lane; t < pt_ptr[I + 1]; t += AMG_SW) {
      const double rv = r[pt_row[t]];
      const half_t *p = ptval + 6 * (int64_t)t;
#pragma unroll
      for (int m = 0; m < 6; ++m) acc[m] += (double)p[m] * rv;
    }
#pragma unroll
    for (int off = AMG_SW / 2; off > 0; off >>= 1)
#pragma unroll
      for (int m = 0; m < 6; ++m) acc[m] += __shfl_xor(acc[m], off, AMG_SW);
    if (lane == 0)
#pragma unroll
      for (int m = 0; m < 6; ++m) bc[6 * I + m] = acc[m];
  }
}

// x += P xc: one thread per fine row over its 6-column blocks
__global__ __launch_bounds__(kBlock) void k_amg_prolong(int n, const int32_t *__restrict__ prow,
                                                        const int32_t *__restrict__ pagg,
                                                        const half_t *__restrict__ pval, const double *__restrict__ xc,
                                                        double *__restrict__ x, const double *__restrict__ done) {
  if (amg_done(done)) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = prow[i]; k < prow[i + 1]; ++k) {
      const half_t *p = pval + 6 * (int64_t)k;
      const double *c = xc + 6 * (int64_t)pagg[k];
#pragma unroll
      for (int m = 0; m < 6; ++m) s += (double)p[m] * c[m];
    }
    x[i] += s;
  }
}

// max |val| as the bits of a non-negative float (they order like the values) into *amax
template <typename TM>
__global__ __launch_bounds__(kBlock) void k_absmax(int64_t n, const TM *__restrict__ val, unsigned int *__restrict__ amax) {
  float m = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf((float)val[i]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(amax, __float_as_uint(m));
}

__global__ void k_absmax_reset(unsigned int *amax) { *amax = 0u; }

// val16 = binary16(2^k val), 2^k putting max |val| in [2^14, 2^15); scale[1] = 2^-k.  The
// values come in triples (a row's entries of one neighbour node, triple-structured
// matrices); each triple t is stored as the quad val16[4t .. 4t+3] (the fourth zero), so a
// lane reads a row's triple with one 8-byte load.
template <typename TM>
__global__ __launch_bounds__(kBlock) void k_to_half(int64_t ntrip, const TM *__restrict__ val,
                                                    const unsigned int *__restrict__ amax, uint2 *__restrict__ out,
                                                    double *__restrict__ scale) {
  const float m = __uint_as_float(*amax);
  int e = 0;
  if (m > 0.f) frexpf(m, &e);  // m = f 2^e, f in [1/2, 1)
  const int k = m > 0.f ? 15 - e : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) scale[1] = ldexp(1.0, -k);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntrip; t += (int64_t)gridDim.x * blockDim.x) {
    const half_t h0 = (half_t)ldexp((double)val[3 * t], k), h1 = (half_t)ldexp((double)val[3 * t + 1], k),
                 h2 = (half_t)ldexp((double)val[3 * t + 2], k);
    uint2 q;
    q.x = (uint32_t)__builtin_bit_cast(uint16_t, h0) | ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
    q.y = (uint32_t)__builtin_bit_cast(uint16_t, h2);
    out[t] = q;
  }
}

// x = inv b on the coarsest level: one wave per row.  Rows of a multiple of 4 entries are read
// as float4 (a lane takes 4 consecutive entries per step) with two accumulators.
__global__ __launch_bounds__(kBlock) void k_amg_dense(int n, const float *__restrict__ inv,
                                                      const double *__restrict__ b, double *__restrict__ x,
                                                      const double *__restrict__ done) {
  if (amg_done(done)) return;
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  double s = 0.0, t = 0.0;
  if ((n & 3) == 0) {
    const float4 *r4 = reinterpret_cast<const float4 *>(inv + row * n);
    const int n4 = n >> 2;
    for (int j = lane; j < n4; j += 64) {
      const float4 v = r4[j];
      const double *bj = b + 4 * j;
      s += (double)v.x * bj[0] + (double)v.y * bj[1];
      t += (double)v.z * bj[2] + (double)v.w * bj[3];
    }
  } else {
    for (int j = lane; j < n; j += 64) s += (double)inv[row * n + j] * b[j];
  }
  s += t;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) x[row] = s;
}

// Chebyshev step coefficients for lambda in [rho / 30, 1.1 rho]: step 0 (d = r / theta),
// then d = c1 d + c2 r with the three-term recurrence
static void cheb_coeffs(double rho, int deg, double *c1, double *c2) {
  const double lmax = 1.1 * rho, lmin = rho / 30.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  double rk = 1.0 / sigma;
  c1[0] = 0.0;
  c2[0] = 1.0 / theta;
  for (int k = 1; k < deg; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rk);
    c1[k] = rn * rk;
    c2[k] = 2.0 * rn / delta;
    rk = rn;
  }
}

constexpr int AMG_MAX_LEVELS = 12, AMG_MAX_DEG = 8;

template <typename TM>
struct AmgFine {
  SpmvShape s;
  const int32_t *grp, *rowptr, *col;
  const TM *val;
  const float *binv;
  double rho;
  const uint2 *val16;        // binary16 quads for the smoothing and residual (or null)
  const double *val16_scale;
};

// the first Chebyshev step from x = 0 on level 0: d = xo = c2 M^-1 b, one thread per group
__global__ __launch_bounds__(kBlock) void k_amg_cheb0_first(int ngroups, const int32_t *__restrict__ grp,
                                                            const float *__restrict__ binv,
                                                            const double *__restrict__ b, double *__restrict__ xo,
                                                            double *__restrict__ d, double c2,
                                                            const double *__restrict__ done) {
  if (amg_done(done)) return;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += (int64_t)gridDim.x * blockDim.x) {
    const int r0 = grp[g], nr = grp[g + 1] - r0;
    double rr[3], w[3];
    bj_load(b, r0, nr, rr);
    bj_apply(binv + 6 * g, rr, w);
    for (int a = 0; a < nr; ++a) {
      const double dn = c2 * w[a];
      d[r0 + a] = dn;
      xo[r0 + a] = dn;
    }
  }
}

// The PCG update fused with the V-cycle's first level-0 step: alpha = r.z / p.q (re-summed
// partials, as the stages of the plain PCG do), x += alpha p, r -= alpha q, then the first Chebyshev step
// from 0 on the new r: d = xo = c2 M^-1 r.  One thread per row group.  gu_rz: the number of
// r.z partials in ring slot 0 (the start stage's), read in iteration 1 only.
__global__ __launch_bounds__(kBlock) void k_amg_pcg_update_first(int ngroups, const int32_t *__restrict__ grp,
                                                                 const float *__restrict__ binv,
                                                                 double *__restrict__ x, double *__restrict__ r,
                                                                 const double *__restrict__ p,
                                                                 const double *__restrict__ q,
                                                                 double *__restrict__ xo, double *__restrict__ d,
                                                                 double c2, double *__restrict__ ws, int gu_rz) {
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  const int g1 = (int)ws[W_G1];
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, gu_rz);
  const double pq = sum_partials(ws + W_PQ, g1);
  const double alpha = pq != 0.0 ? rz_old / pq : 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += (int64_t)gridDim.x * blockDim.x) {
    const int r0 = grp[g], nr = grp[g + 1] - r0;
    double rr[3] = {0.0, 0.0, 0.0}, w[3];
    for (int a = 0; a < nr; ++a) {
      const int i = r0 + a;
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      r[i] = ri;
      rr[a] = ri;
    }
    bj_apply(binv + 6 * g, rr, w);
    for (int a = 0; a < nr; ++a) {
      const double dn = c2 * w[a];
      d[r0 + a] = dn;
      xo[r0 + a] = dn;
    }
  }
}

// the level-0 products of the smoothing steps and the residual read VT values (the matrix's
// own, or its binary16 copy with the scale vscale)
// triple-structured matrices only (the binary16 quads)
#define FEA_SW_SWITCH_TRIPLE(MACRO)                                        \
  do {                                                                     \
    switch (s.sw) { case 8: MACRO(8, 2); break; case 16: MACRO(16, 2); break; \
                    default: MACRO(32, 2); }                               \
  } while (0)

template <typename VT>
static void cheb0_v(const SpmvShape &s, const int32_t *grp, const int32_t *rowptr, const int32_t *col,
                    const VT *val, const double *vscale, const float *binv, const double *b, const double *x,
                    double *xo, double *d, double c1, double c2, const double *done, hipStream_t st, double *rz_ws) {
  const int grid = spmv_grid(s);
  if (rz_ws != nullptr) {
#define FEA_AC2(SW, MD) \
  k_amg_cheb0<VT, SW, MD, 0, 1><<<grid, kBlock, 0, st>>>(s.ngroups, grp, rowptr, col, val, binv, b, x, xo, d, c1, c2, done, rz_ws, vscale)
    if constexpr (std::is_same<VT, uint2>::value) FEA_SW_SWITCH_TRIPLE(FEA_AC2); else FEA_SW_SWITCH(FEA_AC2);
#undef FEA_AC2
  } else {
#define FEA_AC1(SW, MD) \
  k_amg_cheb0<VT, SW, MD, 0><<<grid, kBlock, 0, st>>>(s.ngroups, grp, rowptr, col, val, binv, b, x, xo, d, c1, c2, done, nullptr, vscale)
    if constexpr (std::is_same<VT, uint2>::value) FEA_SW_SWITCH_TRIPLE(FEA_AC1); else FEA_SW_SWITCH(FEA_AC1);
#undef FEA_AC1
  }
}

template <typename TM>
static void cheb0(const AmgFine<TM> &f, bool first, const double *b, const double *x, double *xo, double *d,
                  double c1, double c2, const double *done, hipStream_t st, double *rz_ws = nullptr) {
  const auto &s = f.s;
  if (first)
    k_amg_cheb0_first<<<amg_grid(s.ngroups, kBlock), kBlock, 0, st>>>(s.ngroups, f.grp, f.binv, b, xo, d, c2, done);
  else if (f.val16 != nullptr)
    cheb0_v<uint2>(s, f.grp, f.rowptr, f.col, f.val16, f.val16_scale, f.binv, b, x, xo, d, c1, c2, done, st, rz_ws);
  else
    cheb0_v<TM>(s, f.grp, f.rowptr, f.col, f.val, nullptr, f.binv, b, x, xo, d, c1, c2, done, st, rz_ws);
}

template <typename TM>
static void resid0(const AmgFine<TM> &f, const double *b, const double *x, double *r, const double *done,
                   hipStream_t st) {
  const int grid = spmv_grid(f.s);
  const auto &s = f.s;
  if (f.val16 != nullptr) {
#define FEA_AR0(SW, MD) \
  k_amg_resid0<uint2, SW, MD><<<grid, kBlock, 0, st>>>(s.ngroups, f.grp, f.rowptr, f.col, f.val16, b, x, r, done, f.val16_scale)
    FEA_SW_SWITCH_TRIPLE(FEA_AR0);
#undef FEA_AR0
  } else {
#define FEA_AR0(SW, MD) \
  k_amg_resid0<TM, SW, MD><<<grid, kBlock, 0, st>>>(s.ngroups, f.grp, f.rowptr, f.col, f.val, b, x, r, done, nullptr)
    FEA_SW_SWITCH(FEA_AR0);
#undef FEA_AR0
  }
}

static void cheb_c(const fea_amg_level &L, bool first, const double *b, const double *x, double *xo, double c1,
                   double c2, const double *done, hipStream_t st) {
  const int ng = L.n / 6;
  const int grid = amg_grid(ng, kBlock / AMG_SW);
  if (first)
    k_amg_cheb_c<1><<<grid, kBlock, 0, st>>>(ng, L.rowptr, L.col, L.val, L.dinv, b, x, xo, L.d, c1, c2, done);
  else
    k_amg_cheb_c<0><<<grid, kBlock, 0, st>>>(ng, L.rowptr, L.col, L.val, L.dinv, b, x, xo, L.d, c1, c2, done);
}

// The PCG work a V-cycle takes over inside a PCG iteration (amg_pcg_iterate_t): the update
// x += alpha p, r -= alpha q in its first level-0 step (b0 is r), and the r.z partials in its
// last (gu_rz: the number of partials of ring slot 0, written by the start stage).
struct PcgFuse {
  double *x, *r;
  const double *p, *q;
  double *ws;
  int gu_rz;
};

// z = V(b) on the hierarchy (b, z: level-0 vectors; z must not alias b)
template <typename TM>
static void amg_vcycle(const AmgFine<TM> &f, const fea_amg_level *lv, int nl, int deg, const double *b0, double *z0,
                       const double *done, hipStream_t st, const PcgFuse *pf = nullptr) {
  double c1[AMG_MAX_DEG], c2[AMG_MAX_DEG];
  double *cur[AMG_MAX_LEVELS], *oth[AMG_MAX_LEVELS];
  for (int l = 0; l + 1 < nl; ++l) {  // down: pre-smooth, residual, restriction
    const fea_amg_level &L = lv[l];
    const double *bl = l == 0 ? b0 : L.b;
    cheb_coeffs(l == 0 ? f.rho : L.rho, deg, c1, c2);
    double *xa = L.x, *xb = L.x2;
    if (l == 0 && pf != nullptr)
      k_amg_pcg_update_first<<<amg_grid(f.s.ngroups, kBlock), kBlock, 0, st>>>(
          f.s.ngroups, f.grp, f.binv, pf->x, pf->r, pf->p, pf->q, xa, L.d, c2[0], pf->ws, pf->gu_rz);
    else if (l == 0)
      cheb0(f, true, bl, nullptr, xa, L.d, c1[0], c2[0], done, st);
    else
      cheb_c(L, true, bl, nullptr, xa, c1[0], c2[0], done, st);
    for (int k = 1; k < deg; ++k) {
      if (l == 0)
        cheb0(f, false, bl, xa, xb, L.d, c1[k], c2[k], done, st);
      else
        cheb_c(L, false, bl, xa, xb, c1[k], c2[k], done, st);
      std::swap(xa, xb);
    }
    cur[l] = xa;
    oth[l] = xb;
    if (l == 0)
      resid0(f, bl, xa, L.r, done, st);
    else
      k_amg_resid_c<<<amg_grid(L.n / 6, kBlock / AMG_SW), kBlock, 0, st>>>(L.n / 6, L.rowptr, L.col, L.val, bl, xa,
                                                                           L.r, done);
    k_amg_restrict<<<amg_grid(L.nagg, kBlock / AMG_SW), kBlock, 0, st>>>(L.nagg, L.pt_ptr, L.pt_row,
                                                                        (const half_t *)L.ptval, L.r, lv[l + 1].b,
                                                                        done);
  }
  const fea_amg_level &C = lv[nl - 1];
  k_amg_dense<<<(C.n + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, st>>>(C.n, C.dense_inv, C.b, C.x, done);
  for (int l = nl - 2; l >= 0; --l) {  // up: prolongation, post-smooth
    const fea_amg_level &L = lv[l];
    const double *bl = l == 0 ? b0 : L.b;
    const double *xc = (l + 1 == nl - 1) ? C.x : cur[l + 1];
    k_amg_prolong<<<amg_grid(L.n, kBlock), kBlock, 0, st>>>(L.n, L.prow, L.pagg, (const half_t *)L.pval, xc,
                                                             cur[l], done);
    cheb_coeffs(l == 0 ? f.rho : L.rho, deg, c1, c2);
    double *x = cur[l], *o = oth[l];
    for (int k = 0; k < deg; ++k) {
      const bool last = l == 0 && k == deg - 1;
      double *out = last ? z0 : o;
      if (l == 0)
        cheb0(f, false, bl, x, out, L.d, c1[k], c2[k], done, st, (last && pf != nullptr) ? pf->ws : nullptr);
      else
        cheb_c(L, false, bl, x, out, c1[k], c2[k], done, st);
      o = x;
      x = out;
    }
    cur[l] = x;
  }
}

// ---- PCG stages around the V-cycle
// partials of b.(M^-1 b) for the stopping test
__global__ __launch_bounds__(kBlock) void k_amg_bz(int n, const double *__restrict__ b, const double *__restrict__ zb,
                                                   double *__restrict__ ws) {
  double bz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bz += b[i] * zb[i];
  write_partial(ws + W_BZP, bz);
}

// p = z (= M^-1 r from the V-cycle), partials of r.z (ring slot 0); reset the counters
__global__ __launch_bounds__(kBlock) void k_amg_pcg_init(int n, const double *__restrict__ r,
                                                         const double *__restrict__ z, double *__restrict__ p,
                                                         double *__restrict__ ws) {
  double rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double zi = z[i];
    p[i] = zi;
    rz += r[i] * zi;
  }
  write_partial(ws + W_RZP, rz);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws[W_ITER] = 0.0;
    ws[W_DONE] = 0.0;
  }
}

// y = s D^-1 A x (float) and partials of y.y: one step of the power iteration for lambda_max
template <typename TM, int SW, int MODE>
__global__ __launch_bounds__(kBlock) void k_amg_power0(int ngroups, const int32_t *__restrict__ grp,
                                                       const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col, const TM *__restrict__ val,
                                                       const float *__restrict__ binv, const float *__restrict__ x,
                                                       float *__restrict__ y, float s, double *__restrict__ part) {
  constexpr int GPB = kBlock / SW;
  const int lane = threadIdx.x % SW;
  double yy = 0.0;
  for (int64_t g0 = (int64_t)blockIdx.x * GPB; g0 < ngroups; g0 += (int64_t)gridDim.x * GPB) {
    const int g = (int)(g0 + threadIdx.x / SW);
    if (g >= ngroups) continue;
    float o[3];
    const int r0 = group_spmv<TM, float, SW, MODE>(g, lane, grp, rowptr, col, val, x, o);
    if (lane == 0) {
      const int nr = grp[g + 1] - r0;
      double rr[3] = {0.0, 0.0, 0.0}, w[3];
      for (int a = 0; a < nr; ++a) rr[a] = (double)o[a];
      bj_apply(binv + 6 * (int64_t)g, rr, w);
      for (int a = 0; a < nr; ++a) {
        const float v = s * (float)w[a];
        y[r0 + a] = v;
        yy += (double)v * (double)v;
      }
    }
  }
  write_partial(part, yy);
}

template <typename TM>
static AmgFine<TM> amg_fine(int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                            const int32_t *col, const void *val, const float *binv, double rho,
                            const fea_amg_level *lv0 = nullptr) {
  AmgFine<TM> f;
  f.s = shape_of(n, nnz, n_groups, grp);
  f.grp = grp;
  f.rowptr = rowptr;
  f.col = col;
  f.val = (const TM *)val;
  f.binv = binv;
  f.rho = rho;
  f.val16 = lv0 != nullptr ? (const uint2 *)lv0->val16 : nullptr;
  f.val16_scale = lv0 != nullptr ? lv0->val16_scale : nullptr;
  return f;
}

struct AmgPcgArgs {
  int n;
  int64_t nnz;
  int n_groups;
  const int32_t *grp, *rowptr, *col;
  const void *val;
  const float *binv;
  double rho0;
  const fea_amg_level *lv;
  int nl, deg;
  const double *b;
  double *x, *r, *z, *p, *q;
  double *ws;
  double tol;
  int max_iter, iters;
  double *status_host;
  hipStream_t st;
};

template <typename TM>
static int amg_pcg_start_t(const AmgPcgArgs &a) {
  const AmgFine<TM> f = amg_fine<TM>(a.n, a.nnz, a.n_groups, a.grp, a.rowptr, a.col, a.val, a.binv, a.rho0, a.lv);
  const int gu = vec_grid();
  launch_spmv<TM, double>(f.s, a.grp, a.rowptr, a.col, f.val, a.x, a.b, a.r, a.st);  // r = b - A x
  amg_vcycle<TM>(f, a.lv, a.nl, a.deg, a.b, a.z, nullptr, a.st);                    // z = M^-1 b
  k_amg_bz<<<gu, kBlock, 0, a.st>>>(a.n, a.b, a.z, a.ws);
  amg_vcycle<TM>(f, a.lv, a.nl, a.deg, a.r, a.z, nullptr, a.st);                    // z = M^-1 r
  k_amg_pcg_init<<<gu, kBlock, 0, a.st>>>(a.n, a.r, a.z, a.p, a.ws);
  FEA_CHECK_LAUNCH("fea_pcg_start_amg");
  return FEA_OK;
}

template <typename TM>
static int amg_pcg_iterate_t(const AmgPcgArgs &a) {
  const AmgFine<TM> f = amg_fine<TM>(a.n, a.nnz, a.n_groups, a.grp, a.rowptr, a.col, a.val, a.binv, a.rho0, a.lv);
  const int gu = vec_grid();
  const double *done = a.ws + W_DONE;
  // the PCG update rides in the V-cycle's first level-0 step and the r.z partials in its last
  // (one partial per workgroup of that step's grid)
  const PcgFuse pf{a.x, a.r, a.p, a.q, a.ws, gu};
  const int g_rz = spmv_grid(f.s);
  for (int it = 0; it < a.iters; ++it) {
    launch_pcg_spmv<TM, double>(f.s, a.grp, a.rowptr, a.col, f.val, a.p, a.q, a.ws, a.st);
    amg_vcycle<TM>(f, a.lv, a.nl, a.deg, a.r, a.z, done, a.st, &pf);
    k_pcg_dir<double><<<PCG_G, kBlock, 0, a.st>>>(a.n, a.z, a.p, a.ws, a.tol, a.max_iter, g_rz, gu);
  }
  FEA_CHECK_LAUNCH("fea_pcg_iterate_amg");
  if (a.status_host != nullptr) {
    double h[4];
    FEA_HIP(hipMemcpyAsync(h, a.ws, sizeof(h), hipMemcpyDeviceToHost, a.st), "fea_pcg_iterate_amg");
    FEA_HIP(hipStreamSynchronize(a.st), "fea_pcg_iterate_amg");
    for (int i = 0; i < 4; ++i) a.status_host[i] = h[i];
  }
  return FEA_OK;
}

static int check_amg_levels(const fea_amg_level *lv, int nl, int deg, int n, const char *who) {
  FEA_CHECK_ARG(lv != nullptr && nl >= 2 && nl <= AMG_MAX_LEVELS, "%s: 2 .. %d levels expected", who,
                AMG_MAX_LEVELS);
  FEA_CHECK_ARG(deg >= 1 && deg <= AMG_MAX_DEG, "%s: smoother degree 1 .. %d", who, AMG_MAX_DEG);
  FEA_CHECK_ARG(lv[0].n == n, "%s: level 0 has %d rows, the matrix %d", who, lv[0].n, n);
  FEA_CHECK_ARG(lv[0].val16 == nullptr || lv[0].val16_scale != nullptr, "%s: level-0 binary16 values need their scale",
                who);
  for (int l = 0; l < nl; ++l) {
    const fea_amg_level &L = lv[l];
    FEA_CHECK_ARG(L.x && L.x2 && L.d && L.r && (l == 0 || L.b), "%s: level %d work vectors missing", who, l);
    if (l + 1 < nl) {
      FEA_CHECK_ARG(L.prow && L.pagg && L.pval && L.pt_ptr && L.ptval && L.pt_row && L.nagg * 6 == lv[l + 1].n,
                    "%s: level %d prolongator inconsistent with level %d", who, l, l + 1);
      FEA_CHECK_ARG(l == 0 || (L.rowptr && L.col && L.val && L.dinv && L.n % 6 == 0 && L.rho > 0.0),
                    "%s: level %d operator / smoother missing", who, l);
    } else {
      FEA_CHECK_ARG(L.dense_inv != nullptr && L.n > 0, "%s: the coarsest level needs its dense inverse", who);
    }
  }
  return FEA_OK;
}

}  // namespace fea

extern "C" int fea_amg_level_size(void) { return (int)sizeof(fea_amg_level); }

extern "C" int fea_amg_half_values(int mat_dtype, int64_t nnz, const void *val, uint16_t *val16, double *scale,
                                   void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_amg_half_values: bad matrix dtype");
  FEA_CHECK_ARG(nnz > 0 && nnz % 3 == 0 && val && val16 && scale,
                "fea_amg_half_values: bad arguments (values of a triple-structured matrix expected)");
  hipStream_t st = as_stream(stream);
  unsigned int *amax = reinterpret_cast<unsigned int *>(scale);
  const int grid = (int)std::min<int64_t>(2048, (nnz + kBlock - 1) / kBlock);
  k_absmax_reset<<<1, 1, 0, st>>>(amax);
  if (mat_dtype == FEA_F64) {
    k_absmax<double><<<grid, kBlock, 0, st>>>(nnz, (const double *)val, amax);
    k_to_half<double><<<grid, kBlock, 0, st>>>(nnz / 3, (const double *)val, amax, (uint2 *)val16, scale);
  } else {
    k_absmax<float><<<grid, kBlock, 0, st>>>(nnz, (const float *)val, amax);
    k_to_half<float><<<grid, kBlock, 0, st>>>(nnz / 3, (const float *)val, amax, (uint2 *)val16, scale);
  }
  FEA_CHECK_LAUNCH("fea_amg_half_values");
  return FEA_OK;
}

extern "C" int fea_amg_vcycle(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                              const int32_t *rowptr, const int32_t *col, const void *val, const float *binv,
                              double rho0, const fea_amg_level *levels, int n_levels, int degree, const double *b,
                              double *z, void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_amg_vcycle: bad matrix dtype");
  FEA_CHECK_ARG(n > 0 && grp && n_groups != 0 && rowptr && col && val && binv && b && z && b != z && rho0 > 0.0,
                "fea_amg_vcycle: bad arguments (node groups and block-Jacobi inverses are required)");
  const int rc = check_amg_levels(levels, n_levels, degree, n, "fea_amg_vcycle");
  if (rc != FEA_OK) return rc;
  FEA_CHECK_ARG(levels[0].val16 == nullptr || n_groups < 0,
                "fea_amg_vcycle: level-0 binary16 values need a triple-structured matrix (n_groups < 0)");
  hipStream_t st = as_stream(stream);
  if (mat_dtype == FEA_F64)
    amg_vcycle<double>(amg_fine<double>(n, nnz, n_groups, grp, rowptr, col, val, binv, rho0, levels), levels,
                       n_levels, degree, b, z, nullptr, st);
  else
    amg_vcycle<float>(amg_fine<float>(n, nnz, n_groups, grp, rowptr, col, val, binv, rho0, levels), levels,
                      n_levels, degree, b, z, nullptr, st);
  FEA_CHECK_LAUNCH("fea_amg_vcycle");
  return FEA_OK;
}

extern "C" int fea_amg_smooth0(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                               const int32_t *rowptr, const int32_t *col, const void *val, const float *binv,
                               const double *b, const double *x, double *xo, double *d, double c1, double c2,
                               void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_amg_smooth0: bad matrix dtype");
  FEA_CHECK_ARG(n > 0 && grp && n_groups != 0 && rowptr && col && val && binv && b && xo && d && x != xo,
                "fea_amg_smooth0: bad arguments");
  const AmgFine<float> ff = amg_fine<float>(n, nnz, n_groups, grp, rowptr, col, val, binv, 1.0);
  const AmgFine<double> fd = amg_fine<double>(n, nnz, n_groups, grp, rowptr, col, val, binv, 1.0);
  hipStream_t st = as_stream(stream);
  if (mat_dtype == FEA_F64)
    cheb0(fd, x == nullptr, b, x, xo, d, c1, c2, nullptr, st);
  else
    cheb0(ff, x == nullptr, b, x, xo, d, c1, c2, nullptr, st);
  FEA_CHECK_LAUNCH("fea_amg_smooth0");
  return FEA_OK;
}

extern "C" int fea_amg_smooth0_half(int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                                    const int32_t *col, const uint16_t *val16, const double *scale,
                                    const float *binv, const double *b, const double *x, double *xo, double *d,
                                    double c1, double c2, void *stream) {
  FEA_CHECK_ARG(n > 0 && grp && n_groups < 0 && rowptr && col && val16 && scale && binv && b && xo && d && x != xo,
                "fea_amg_smooth0_half: bad arguments (a triple-structured matrix, n_groups < 0, is required)");
  AmgFine<float> f = amg_fine<float>(n, nnz, n_groups, grp, rowptr, col, nullptr, binv, 1.0);
  f.val16 = (const uint2 *)val16;
  f.val16_scale = scale;
  cheb0(f, x == nullptr, b, x, xo, d, c1, c2, nullptr, as_stream(stream));
  FEA_CHECK_LAUNCH("fea_amg_smooth0_half");
  return FEA_OK;
}

extern "C" int fea_amg_power(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                             const int32_t *rowptr, const int32_t *col, const void *val, const float *binv, float *x,
                             float *y, double *partial, int iters, double *rho_out, void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_amg_power: bad matrix dtype");
  FEA_CHECK_ARG(n > 0 && grp && n_groups != 0 && rowptr && col && val && binv && x && y && partial && rho_out &&
                    iters >= 1,
                "fea_amg_power: bad arguments");
  hipStream_t st = as_stream(stream);
  const SpmvShape s = shape_of(n, nnz, n_groups, grp);
  const int grid = spmv_grid(s);
  std::vector<double> h(grid);
  // x has unit 2-norm; every step maps a unit input to D^-1 A of it, whose norm estimates
  // lambda_max, and the next step scales its image back by that norm
  float scale = 1.f;
  double lam = 0.0;
  for (int it = 0; it < iters; ++it) {
    float *in = (it & 1) ? y : x, *out = (it & 1) ? x : y;
    if (mat_dtype == FEA_F64) {
#define FEA_PW(SW, MD) \
  k_amg_power0<double, SW, MD><<<grid, kBlock, 0, st>>>(s.ngroups, grp, rowptr, col, (const double *)val, binv, in, out, scale, partial)
      FEA_SW_SWITCH(FEA_PW);
#undef FEA_PW
    } else {
#define FEA_PW(SW, MD) \
  k_amg_power0<float, SW, MD><<<grid, kBlock, 0, st>>>(s.ngroups, grp, rowptr, col, (const float *)val, binv, in, out, scale, partial)
      FEA_SW_SWITCH(FEA_PW);
#undef FEA_PW
    }
    FEA_CHECK_LAUNCH("fea_amg_power");
    FEA_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * grid, hipMemcpyDeviceToHost, st), "fea_amg_power");
    FEA_HIP(hipStreamSynchronize(st), "fea_amg_power");
    double yy = 0.0;
    for (int k = 0; k < grid; ++k) yy += h[k];
    const double ny = std::sqrt(yy);
    lam = ny;
    if (!(ny > 0.0)) break;
    scale = (float)(1.0 / ny);
  }
  *rho_out = lam;
  return FEA_OK;
}

extern "C" int fea_pcg_start_amg(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                                 const int32_t *rowptr, const int32_t *col, const void *val, const float *binv,
                                 double rho0, const fea_amg_level *levels, int n_levels, int degree, const double *b,
                                 double *x, double *r, double *z, double *p, double *ws, void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_pcg_start_amg: bad matrix dtype");
  FEA_CHECK_ARG(n > 0 && grp && n_groups != 0 && rowptr && col && val && binv && b && x && r && z && p && ws &&
                    rho0 > 0.0,
                "fea_pcg_start_amg: bad arguments");
  const int rc = check_amg_levels(levels, n_levels, degree, n, "fea_pcg_start_amg");
  if (rc != FEA_OK) return rc;
  FEA_CHECK_ARG(levels[0].val16 == nullptr || n_groups < 0,
                "fea_pcg_start_amg: level-0 binary16 values need a triple-structured matrix (n_groups < 0)");
  AmgPcgArgs a{n, nnz, n_groups, grp, rowptr, col, val, binv, rho0, levels, n_levels, degree, b, x, r, z, p,
               nullptr, ws, 0.0, 0, 0, nullptr, as_stream(stream)};
  return mat_dtype == FEA_F64 ? amg_pcg_start_t<double>(a) : amg_pcg_start_t<float>(a);
}

extern "C" int fea_pcg_iterate_amg(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                                   const int32_t *rowptr, const int32_t *col, const void *val, const float *binv,
                                   double rho0, const fea_amg_level *levels, int n_levels, int degree, double *x,
                                   double *r, double *z, double *p, double *q, double *ws, double tol, int max_iter,
                                   int iters, double *status_host, void *stream) {
  FEA_CHECK_ARG(mat_dtype == FEA_F64 || mat_dtype == FEA_F32, "fea_pcg_iterate_amg: bad matrix dtype");
  FEA_CHECK_ARG(n > 0 && grp && n_groups != 0 && rowptr && col && val && binv && x && r && z && p && q && ws &&
                    rho0 > 0.0 && iters >= 0 && max_iter >= 1,
                "fea_pcg_iterate_amg: bad arguments");
  const int rc = check_amg_levels(levels, n_levels, degree, n, "fea_pcg_iterate_amg");
  if (rc != FEA_OK) return rc;
  FEA_CHECK_ARG(levels[0].val16 == nullptr || n_groups < 0,
                "fea_pcg_iterate_amg: level-0 binary16 values need a triple-structured matrix (n_groups < 0)");
  AmgPcgArgs a{n, nnz, n_groups, grp, rowptr, col, val, binv, rho0, levels, n_levels, degree, nullptr, x, r, z, p,
               q, ws, tol, max_iter, iters, status_host, as_stream(stream)};
  return mat_dtype == FEA_F64 ? amg_pcg_iterate_t<double>(a) : amg_pcg_iterate_t<float>(a);
}
