// Batched Jacobi-PCG: one launch per CG stage advances every system of a batch
// of independent SPD systems (one per static-solve request in the solve server).
//
// Each system keeps its own CSR, vectors and float64 workspace (the layout of
// the single-system PCG in pcg.hip).  A launch's workgroups are partitioned
// among the systems by prefix tables (spmv_start / vec_start: first workgroup
// of each system); a workgroup finds its system with one wave-wide ballot over
// the table.  The number of workgroups a system gets, and so the order of every
// one of its partial sums, depends only on that system's own size: a system
// produces bit-identical iterates alone and in any batch.
//
// Converged systems (ws[W_DONE] set) return at once, so a batch keeps running
// until its slowest member finishes without disturbing the others.  A grid may be
// larger than the last range (the caller pads launch sizes so that one captured graph
// serves several batch layouts): its extra workgroups return at once.
#include <climits>

#include "pcg_device.h"

namespace fea {

constexpr int kBatchSW = 16;  // sub-wave per node group in batched SpMV (fixed: batch-independent)
constexpr int kSteps = 4;     // grid-stride steps per workgroup a system's grid is sized for

// Launch dimensions of a batched stage: the system count and the stage's total of virtual
// workgroups, by value, or (hdr != NULL) from a device header {n_sys, spmv_total,
// vec_total} so that a captured launch serves any batch layout.  A physical workgroup
// takes virtual workgroups blockIdx.x, blockIdx.x + gridDim.x, ...; each does the work a
// workgroup of the exact grid does, so results do not depend on the grid size.
struct BatchDims {
  int n_sys, total;
  const int32_t *hdr;
};

// index k of the system whose workgroup range [start[k], start[k+1]) holds b: the number of
// systems whose range ends at or before b.  Every lane loads its (up to 16) table entries
// first, so a workgroup waits for one load latency, not one per 64 systems (n_sys is at
// most FEA_PCG_BATCH_MAX, checked by every entry point).
__device__ __forceinline__ int find_system(const int32_t *__restrict__ start, int n_sys, int b) {
  constexpr int kSteps = (FEA_PCG_BATCH_MAX + 63) / 64;
  const int lane = (int)(threadIdx.x & 63);
  int end[kSteps];
#pragma unroll
  for (int t = 0; t < kSteps; ++t) {
    const int k = t * 64 + lane;
    end[t] = k < n_sys ? start[k + 1] : INT_MAX;
  }
  int cnt = 0;
#pragma unroll
  for (int t = 0; t < kSteps; ++t) cnt += __popcll(__ballot(end[t] <= b));
  return __builtin_amdgcn_readfirstlane(cnt);
}

template <typename TM, typename TV>
__device__ __forceinline__ int sys_group_spmv(const fea_pcg_system &S, int g, int lane, const TV *x,
                                              TV o[3]) {
  if (S.n_groups < 0)
    return group_spmv<TM, TV, kBatchSW, 2>(g, lane, S.grp, S.rowptr, S.col, (const TM *)S.val, x, o);
  return group_spmv<TM, TV, kBatchSW, 1>(g, lane, S.grp, S.rowptr, S.col, (const TM *)S.val, x, o);
}

// r = b - A x and dinv = 1/|diag A| (1 where the diagonal is 0) for every system
template <typename TM, typename TV>
__device__ __forceinline__ void batch_residual_vb(int n_sys, const fea_pcg_system *__restrict__ sys,
                                                           const int32_t *__restrict__ spmv_start, int vb) {
  const int k = find_system(spmv_start, n_sys, vb);
  if (k >= n_sys) return;  // a workgroup of a padded grid, past the last range
  const fea_pcg_system S = sys[k];
  const int lb = vb - spmv_start[k];
  const int ngroups = S.n_groups < 0 ? -S.n_groups : S.n_groups;
  constexpr int GPB = kBlock / kBatchSW;
  const int lane = threadIdx.x % kBatchSW;
  const TV *x = (const TV *)S.x;
  const TV *b = (const TV *)S.b;
  TV *r = (TV *)S.r;
  TV *dinv = (TV *)S.dinv;
  const TM *val = (const TM *)S.val;
  for (int64_t g0 = (int64_t)lb * GPB; g0 < ngroups; g0 += (int64_t)S.spmv_blocks * GPB) {
    const int g = (int)(g0 + threadIdx.x / kBatchSW);
    if (g >= ngroups) continue;
    TV o[3];
    const int r0 = sys_group_spmv<TM, TV>(S, g, lane, x, o);
    if (lane == 0) {
      const int nr = S.grp[g + 1] - r0;
      for (int j = 0; j < nr; ++j) r[r0 + j] = b[r0 + j] - o[j];
      if (S.precond == 1) {
        // the node's diagonal block: entry (a, c) at diag_slot[r0 + a] + c - a.  On a
        // triple-structured system the node's own columns are one aligned triple of every
        // row of the group; otherwise the CSR columns confirm each entry
        double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < nr; ++a) {
          const int d = S.diag_slot[r0 + a];
          if (d < 0) continue;
          for (int c = 0; c < nr; ++c) {
            const int sl = d + c - a;
            const bool ok = S.n_groups < 0 ||
                            (sl >= S.rowptr[r0 + a] && sl < S.rowptr[r0 + a + 1] && S.col[sl] == r0 + c);
            if (ok) A[a][c] = (double)val[sl];
          }
        }
        bj_invert_store(A, nr, (float *)S.dinv + 6 * (int64_t)g);
      } else {
        for (int j = 0; j < nr; ++j) {
          const int s = S.diag_slot[r0 + j];
          const TV d = (s >= 0) ? fabs((TV)val[s]) : TV(0);
          dinv[r0 + j] = (d == TV(0)) ? TV(1) : TV(1) / d;
        }
      }
    }
  }
}

template <typename TM, typename TV>
__global__ __launch_bounds__(kBlock) void k_batch_residual(BatchDims d, const fea_pcg_system *__restrict__ sys,
                                                           const int32_t *__restrict__ spmv_start) {
  const int n_sys = d.hdr != nullptr ? d.hdr[0] : d.n_sys;
  const int total = d.hdr != nullptr ? d.hdr[1] : d.total;
  for (int b = blockIdx.x; b < total; b += gridDim.x) batch_residual_vb<TM, TV>(n_sys, sys, spmv_start, b);
}

// z = dinv r, p = z; r.z and b.(dinv b) partials; iteration counter reset
template <typename TV>
__device__ __forceinline__ void batch_init_vb(int n_sys, const fea_pcg_system *__restrict__ sys,
                                                       const int32_t *__restrict__ vec_start, int vb) {
  const int k = find_system(vec_start, n_sys, vb);
  if (k >= n_sys) return;  // a workgroup of a padded grid, past the last range
  const fea_pcg_system S = sys[k];
  const int lb = vb - vec_start[k];
  const TV *dinv = (const TV *)S.dinv;
  const TV *b = (const TV *)S.b;
  const TV *r = (const TV *)S.r;
  TV *z = (TV *)S.z;
  TV *p = (TV *)S.p;
  double rz = 0.0, bz = 0.0;
  if (S.precond == 1) {
    // node-block Jacobi: one row group per thread
    const int ng = S.n_groups < 0 ? -S.n_groups : S.n_groups;
    const float *binv = (const float *)S.dinv;
    for (int64_t g = (int64_t)lb * blockDim.x + threadIdx.x; g < ng; g += (int64_t)S.vec_blocks * blockDim.x) {
      const int r0 = S.grp[g], nr = S.grp[g + 1] - r0;
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
  } else {
    for (int64_t i = (int64_t)lb * blockDim.x + threadIdx.x; i < S.n; i += (int64_t)S.vec_blocks * blockDim.x) {
      const TV zi = dinv[i] * r[i];
      z[i] = zi;
      p[i] = zi;
      rz += (double)r[i] * (double)zi;
      bz += (double)b[i] * (double)dinv[i] * (double)b[i];
    }
  }
  double *ws = S.ws;
  const double srz = block_sum(rz);
  const double sbz = block_sum(bz);
  if (threadIdx.x == 0) {
    ws[W_RZP + lb] = srz;
    ws[W_BZP + lb] = sbz;
  }
  if (lb == 0 && threadIdx.x == 0) {
    ws[W_ITER] = 0.0;
    ws[W_DONE] = 0.0;
    ws[W_G1] = (double)S.spmv_blocks;
  }
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_batch_init(BatchDims d, const fea_pcg_system *__restrict__ sys,
                                                       const int32_t *__restrict__ vec_start) {
  const int n_sys = d.hdr != nullptr ? d.hdr[0] : d.n_sys;
  const int total = d.hdr != nullptr ? d.hdr[2] : d.total;
  for (int b = blockIdx.x; b < total; b += gridDim.x) batch_init_vb<TV>(n_sys, sys, vec_start, b);
}

// q = A p and p.q partials; the system's first workgroup advances its iteration count
template <typename TM, typename TV>
__device__ __forceinline__ void batch_spmv_vb(int n_sys, const fea_pcg_system *__restrict__ sys,
                                                       const int32_t *__restrict__ spmv_start, int vb) {
  const int k = find_system(spmv_start, n_sys, vb);
  if (k >= n_sys) return;  // a workgroup of a padded grid, past the last range
  const fea_pcg_system S = sys[k];
  double *ws = S.ws;
  if (ws[W_DONE] != 0.0) return;
  const int lb = vb - spmv_start[k];
  const int ngroups = S.n_groups < 0 ? -S.n_groups : S.n_groups;
  constexpr int GPB = kBlock / kBatchSW;
  const int lane = threadIdx.x % kBatchSW;
  const TV *p = (const TV *)S.p;
  TV *q = (TV *)S.q;
  double pq = 0.0;
  for (int64_t g0 = (int64_t)lb * GPB; g0 < ngroups; g0 += (int64_t)S.spmv_blocks * GPB) {
    const int g = (int)(g0 + threadIdx.x / kBatchSW);
    if (g >= ngroups) continue;
    TV o[3];
    const int r0 = sys_group_spmv<TM, TV>(S, g, lane, p, o);
    if (lane == 0) {
      const int nr = S.grp[g + 1] - r0;
      for (int j = 0; j < nr; ++j) {
        q[r0 + j] = o[j];
        pq += (double)o[j] * (double)p[r0 + j];
      }
    }
  }
  const double s = block_sum(pq);
  if (threadIdx.x == 0) ws[W_PQ + lb] = s;
  if (lb == 0 && threadIdx.x == 0) ws[W_ITER] += 1.0;
}

template <typename TM, typename TV>
__global__ __launch_bounds__(kBlock) void k_batch_spmv(BatchDims d, const fea_pcg_system *__restrict__ sys,
                                                       const int32_t *__restrict__ spmv_start) {
  const int n_sys = d.hdr != nullptr ? d.hdr[0] : d.n_sys;
  const int total = d.hdr != nullptr ? d.hdr[1] : d.total;
  for (int b = blockIdx.x; b < total; b += gridDim.x) batch_spmv_vb<TM, TV>(n_sys, sys, spmv_start, b);
}

// alpha = rz_{i-1} / pq; x += alpha p; r -= alpha q; z = dinv r; r.z partials
template <typename TV>
__device__ __forceinline__ void batch_update_vb(int n_sys, const fea_pcg_system *__restrict__ sys,
                                                         const int32_t *__restrict__ vec_start, int vb) {
  const int k = find_system(vec_start, n_sys, vb);
  if (k >= n_sys) return;  // a workgroup of a padded grid, past the last range
  const fea_pcg_system S = sys[k];
  double *ws = S.ws;
  if (ws[W_DONE] != 0.0) return;
  const int lb = vb - vec_start[k];
  const int it = (int)ws[W_ITER];
  // r.z of the previous iteration: stored by its direction stage (a re-sum of the same
  // partials in the same order gives the same bits); summed here only in iteration 1
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, S.vec_blocks);
  const double pq = sum_partials(ws + W_PQ, S.spmv_blocks);
  const TV alpha = (TV)(pq != 0.0 ? rz_old / pq : 0.0);  // b = A x exactly: no step
  const TV *dinv = (const TV *)S.dinv;
  const TV *p = (const TV *)S.p;
  const TV *q = (const TV *)S.q;
  TV *x = (TV *)S.x;
  TV *r = (TV *)S.r;
  TV *z = (TV *)S.z;
  double rz = 0.0;
  if (S.precond == 1) {
    // node-block Jacobi: one row group per thread
    const int ng = S.n_groups < 0 ? -S.n_groups : S.n_groups;
    const float *binv = (const float *)S.dinv;
    for (int64_t g = (int64_t)lb * blockDim.x + threadIdx.x; g < ng; g += (int64_t)S.vec_blocks * blockDim.x) {
      const int r0 = S.grp[g], nr = S.grp[g + 1] - r0;
      double rr[3] = {0.0, 0.0, 0.0}, zz[3];
      for (int a = 0; a < nr; ++a) {
        const int i = r0 + a;
        x[i] += alpha * p[i];
        const TV ri = r[i] - alpha * q[i];
        r[i] = ri;
        rr[a] = (double)ri;
      }
      bj_apply(binv + 6 * g, rr, zz);
      for (int a = 0; a < nr; ++a) {
        const TV zi = (TV)zz[a];
        z[r0 + a] = zi;
        rz += rr[a] * (double)zi;
      }
    }
  } else {
    for (int64_t i = (int64_t)lb * blockDim.x + threadIdx.x; i < S.n; i += (int64_t)S.vec_blocks * blockDim.x) {
      x[i] += alpha * p[i];
      const TV ri = r[i] - alpha * q[i];
      r[i] = ri;
      const TV zi = dinv[i] * ri;
      z[i] = zi;
      rz += (double)ri * (double)zi;
    }
  }
  const double s = block_sum(rz);
  if (threadIdx.x == 0) ws[W_RZP + (it & 1) * PCG_RING + lb] = s;
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_batch_update(BatchDims d, const fea_pcg_system *__restrict__ sys,
                                                         const int32_t *__restrict__ vec_start) {
  const int n_sys = d.hdr != nullptr ? d.hdr[0] : d.n_sys;
  const int total = d.hdr != nullptr ? d.hdr[2] : d.total;
  for (int b = blockIdx.x; b < total; b += gridDim.x) batch_update_vb<TV>(n_sys, sys, vec_start, b);
}

// convergence test (r.z / b.(dinv b) < tol or max_iter) and p = z + beta p
template <typename TV>
__device__ __forceinline__ void batch_dir_vb(int n_sys, const fea_pcg_system *__restrict__ sys,
                                                      const int32_t *__restrict__ vec_start, int vb) {
  const int k = find_system(vec_start, n_sys, vb);
  if (k >= n_sys) return;  // a workgroup of a padded grid, past the last range
  const fea_pcg_system S = sys[k];
  double *ws = S.ws;
  if (ws[W_DONE] != 0.0) return;
  const int lb = vb - vec_start[k];
  const int it = (int)ws[W_ITER];
  const double rz_new = sum_partials(ws + W_RZP + (it & 1) * PCG_RING, S.vec_blocks);
  const double rz_old = it > 1 ? ws[W_RZH + ((it - 1) & 1)]
                               : sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_RING, S.vec_blocks);
  // b.z is rewritten with the same value by every direction stage: from iteration 2 on
  // it is read back (workgroup 0 of this launch may be storing the identical bits)
  const double bz = it > 1 ? ws[W_BZ] : sum_partials(ws + W_BZP, S.vec_blocks);
  const bool done = (rz_new == 0.0) || (rz_new / bz < S.tol) || (it >= S.max_iter) || !(rz_new == rz_new);
  // every workgroup of the system reads W_DONE at its start: only a later launch may see it set
  if (lb == 0 && threadIdx.x == 0) {
    ws[W_RZ] = rz_new;
    ws[W_RZH + (it & 1)] = rz_new;
    ws[W_BZ] = bz;
    if (done) ws[W_DONE] = 1.0;
  }
  if (done) return;
  const TV beta = (TV)(rz_new / rz_old);
  const TV *z = (const TV *)S.z;
  TV *p = (TV *)S.p;
  for (int64_t i = (int64_t)lb * blockDim.x + threadIdx.x; i < S.n; i += (int64_t)S.vec_blocks * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_batch_dir(BatchDims d, const fea_pcg_system *__restrict__ sys,
                                                      const int32_t *__restrict__ vec_start) {
  const int n_sys = d.hdr != nullptr ? d.hdr[0] : d.n_sys;
  const int total = d.hdr != nullptr ? d.hdr[2] : d.total;
  for (int b = blockIdx.x; b < total; b += gridDim.x) batch_dir_vb<TV>(n_sys, sys, vec_start, b);
}

// status[2k] = iterations, status[2k+1] = done flag of system k
__global__ __launch_bounds__(kBlock) void k_batch_status(int n_arg, const int32_t *__restrict__ hdr,
                                                         const fea_pcg_system *__restrict__ sys,
                                                         double *__restrict__ status) {
  const int n_sys = hdr != nullptr ? hdr[0] : n_arg;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_sys) return;
  const double *ws = sys[k].ws;
  status[2 * k] = ws[W_ITER];
  status[2 * k + 1] = ws[W_DONE];
}

static bool batch_combo(int m, int v) {
  return (m == FEA_F64 && v == FEA_F64) || (m == FEA_F32 && v == FEA_F64) || (m == FEA_F32 && v == FEA_F32);
}

template <typename TM, typename TV>
static void batch_start_t(int n_sys, const fea_pcg_system *sys, const int32_t *ss, const int32_t *vs, int st_tot,
                          int vt_tot, hipStream_t st) {
  k_batch_residual<TM, TV><<<st_tot, kBlock, 0, st>>>(BatchDims{n_sys, st_tot, nullptr}, sys, ss);
  k_batch_init<TV><<<vt_tot, kBlock, 0, st>>>(BatchDims{n_sys, vt_tot, nullptr}, sys, vs);
}

// grids of st_grid / vt_grid workgroups; with hdr the counts come from the device header
template <typename TM, typename TV>
static void batch_iterate_t(int n_sys, const int32_t *hdr, const fea_pcg_system *sys, const int32_t *ss,
                            const int32_t *vs, int st_tot, int vt_tot, int st_grid, int vt_grid, int iters,
                            hipStream_t st) {
  const BatchDims ds{n_sys, st_tot, hdr}, dv{n_sys, vt_tot, hdr};
  for (int i = 0; i < iters; ++i) {
    k_batch_spmv<TM, TV><<<st_grid, kBlock, 0, st>>>(ds, sys, ss);
    k_batch_update<TV><<<vt_grid, kBlock, 0, st>>>(dv, sys, vs);
    k_batch_dir<TV><<<vt_grid, kBlock, 0, st>>>(dv, sys, vs);
  }
}

}  // namespace fea

using namespace fea;

extern "C" int fea_pcg_system_size(void) { return (int)sizeof(fea_pcg_system); }

extern "C" int fea_pcg_batch_layout(int n_sys, fea_pcg_system *sys_host, int32_t *spmv_start_host,
                                    int32_t *vec_start_host) {
  FEA_CHECK_ARG(n_sys > 0 && n_sys <= FEA_PCG_BATCH_MAX, "fea_pcg_batch_layout: n_sys %d outside 1..%d", n_sys,
                FEA_PCG_BATCH_MAX);
  FEA_CHECK_ARG(sys_host && spmv_start_host && vec_start_host, "fea_pcg_batch_layout: null argument");
  int64_t s_acc = 0, v_acc = 0;
  spmv_start_host[0] = 0;
  vec_start_host[0] = 0;
  for (int k = 0; k < n_sys; ++k) {
    fea_pcg_system &S = sys_host[k];
    FEA_CHECK_ARG(S.n > 0 && S.n_groups != 0 && S.grp && S.rowptr && S.col && S.diag_slot && S.val && S.dinv &&
                      S.b && S.x && S.r && S.z && S.p && S.q && S.ws && S.max_iter >= 1,
                  "fea_pcg_batch_layout: system %d is incomplete", k);
    FEA_CHECK_ARG(S.precond == 0 || S.precond == 1, "fea_pcg_batch_layout: system %d: precond must be 0 or 1", k);
    const int64_t ng = S.n_groups < 0 ? -(int64_t)S.n_groups : S.n_groups;
    // kSteps grid-stride steps per workgroup (as the single-system grids give), so a
    // workgroup's fixed start (finding its system, loading the descriptor) is amortised
    int64_t sb = (ng * kBatchSW + kSteps * kBlock - 1) / (kSteps * kBlock);
    sb = sb < 1 ? 1 : (sb > G1_MAX ? G1_MAX : sb);
    int64_t vb = ((int64_t)S.n + kSteps * kBlock - 1) / (kSteps * kBlock);
    vb = vb > PCG_G ? PCG_G : vb;
    S.spmv_blocks = (int32_t)sb;
    S.vec_blocks = (int32_t)vb;
    s_acc += sb;
    v_acc += vb;
    FEA_CHECK_ARG(s_acc < (1LL << 30) && v_acc < (1LL << 30), "fea_pcg_batch_layout: batch too large");
    spmv_start_host[k + 1] = (int32_t)s_acc;
    vec_start_host[k + 1] = (int32_t)v_acc;
  }
  return FEA_OK;
}

extern "C" int fea_pcg_batch_start(int mat_dtype, int vec_dtype, int n_sys, const fea_pcg_system *sys,
                                   const int32_t *spmv_start, const int32_t *vec_start, int spmv_total,
                                   int vec_total, void *stream) {
  FEA_CHECK_ARG(batch_combo(mat_dtype, vec_dtype), "fea_pcg_batch_start: unsupported precision pair");
  FEA_CHECK_ARG(n_sys > 0 && n_sys <= FEA_PCG_BATCH_MAX && spmv_total > 0 && vec_total > 0,
                "fea_pcg_batch_start: bad sizes");
  FEA_CHECK_ARG(sys && spmv_start && vec_start, "fea_pcg_batch_start: null argument");
  hipStream_t st = as_stream(stream);
  if (mat_dtype == FEA_F64)
    batch_start_t<double, double>(n_sys, sys, spmv_start, vec_start, spmv_total, vec_total, st);
  else if (vec_dtype == FEA_F64)
    batch_start_t<float, double>(n_sys, sys, spmv_start, vec_start, spmv_total, vec_total, st);
  else
    batch_start_t<float, float>(n_sys, sys, spmv_start, vec_start, spmv_total, vec_total, st);
  FEA_CHECK_LAUNCH("fea_pcg_batch_start");
  return FEA_OK;
}

extern "C" int fea_pcg_batch_spmv(int mat_dtype, int vec_dtype, int n_sys, const fea_pcg_system *sys,
                                  const int32_t *spmv_start, int spmv_total, void *stream) {
  FEA_CHECK_ARG(batch_combo(mat_dtype, vec_dtype), "fea_pcg_batch_spmv: unsupported precision pair");
  FEA_CHECK_ARG(n_sys > 0 && n_sys <= FEA_PCG_BATCH_MAX && spmv_total > 0, "fea_pcg_batch_spmv: bad sizes");
  FEA_CHECK_ARG(sys && spmv_start, "fea_pcg_batch_spmv: null argument");
  hipStream_t st = as_stream(stream);
  if (mat_dtype == FEA_F64)
    k_batch_spmv<double, double><<<spmv_total, kBlock, 0, st>>>(BatchDims{n_sys, spmv_total, nullptr}, sys,
                                                                spmv_start);
  else if (vec_dtype == FEA_F64)
    k_batch_spmv<float, double><<<spmv_total, kBlock, 0, st>>>(BatchDims{n_sys, spmv_total, nullptr}, sys,
                                                               spmv_start);
  else
    k_batch_spmv<float, float><<<spmv_total, kBlock, 0, st>>>(BatchDims{n_sys, spmv_total, nullptr}, sys,
                                                              spmv_start);
  FEA_CHECK_LAUNCH("fea_pcg_batch_spmv");
  return FEA_OK;
}

extern "C" int fea_pcg_batch_iterate(int mat_dtype, int vec_dtype, int n_sys, const fea_pcg_system *sys,
                                     const int32_t *spmv_start, const int32_t *vec_start, int spmv_total,
                                     int vec_total, int iters, double *status, void *stream) {
  FEA_CHECK_ARG(batch_combo(mat_dtype, vec_dtype), "fea_pcg_batch_iterate: unsupported precision pair");
  FEA_CHECK_ARG(n_sys > 0 && n_sys <= FEA_PCG_BATCH_MAX && spmv_total > 0 && vec_total > 0 && iters >= 0,
                "fea_pcg_batch_iterate: bad sizes");
  FEA_CHECK_ARG(sys && spmv_start && vec_start, "fea_pcg_batch_iterate: null argument");
  hipStream_t st = as_stream(stream);
  const int S = spmv_total, V = vec_total;
  if (mat_dtype == FEA_F64)
    batch_iterate_t<double, double>(n_sys, nullptr, sys, spmv_start, vec_start, S, V, S, V, iters, st);
  else if (vec_dtype == FEA_F64)
    batch_iterate_t<float, double>(n_sys, nullptr, sys, spmv_start, vec_start, S, V, S, V, iters, st);
  else
    batch_iterate_t<float, float>(n_sys, nullptr, sys, spmv_start, vec_start, S, V, S, V, iters, st);
  if (status != nullptr) k_batch_status<<<blocks_for(n_sys), kBlock, 0, st>>>(n_sys, nullptr, sys, status);
  FEA_CHECK_LAUNCH("fea_pcg_batch_iterate");
  return FEA_OK;
}

extern "C" int fea_pcg_batch_iterate_dev(int mat_dtype, int vec_dtype, const int32_t *hdr,
                                         const fea_pcg_system *sys, const int32_t *spmv_start,
                                         const int32_t *vec_start, int spmv_grid, int vec_grid, int iters,
                                         double *status, void *stream) {
  FEA_CHECK_ARG(batch_combo(mat_dtype, vec_dtype), "fea_pcg_batch_iterate_dev: unsupported precision pair");
  FEA_CHECK_ARG(spmv_grid > 0 && vec_grid > 0 && iters >= 0, "fea_pcg_batch_iterate_dev: bad sizes");
  FEA_CHECK_ARG(hdr && sys && spmv_start && vec_start, "fea_pcg_batch_iterate_dev: null argument");
  hipStream_t st = as_stream(stream);
  const int G = spmv_grid, H = vec_grid;
  if (mat_dtype == FEA_F64)
    batch_iterate_t<double, double>(0, hdr, sys, spmv_start, vec_start, 0, 0, G, H, iters, st);
  else if (vec_dtype == FEA_F64)
    batch_iterate_t<float, double>(0, hdr, sys, spmv_start, vec_start, 0, 0, G, H, iters, st);
  else
    batch_iterate_t<float, float>(0, hdr, sys, spmv_start, vec_start, 0, 0, G, H, iters, st);
  if (status != nullptr)
    k_batch_status<<<blocks_for(FEA_PCG_BATCH_MAX), kBlock, 0, st>>>(0, hdr, sys, status);
  FEA_CHECK_LAUNCH("fea_pcg_batch_iterate_dev");
  return FEA_OK;
}
