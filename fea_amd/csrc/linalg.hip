// Assembly gathers, reductions, CSR SpMV, Jacobi-PCG with device-resident
// scalars, dof-mapped vector updates and the fused explicit update.
#include "common.h"

namespace fea {

// ------------------------------------------------------------------ gathers

// One thread per node-pair block: sum the block's staged 3x3 contributions in
// slot order (fixed => deterministic) and write the free x free entries.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_csr(int n_blocks, const int32_t *__restrict__ blk_a,
                                                       const int32_t *__restrict__ blk_b,
                                                       const int32_t *__restrict__ blk_colpos,
                                                       const int32_t *__restrict__ kseg,
                                                       const int32_t *__restrict__ dof_map,
                                                       const int32_t *__restrict__ rowptr,
                                                       const T *__restrict__ kstage,
                                                       T *__restrict__ val) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_blocks) return;
  T acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = T(0);
  const int s0 = kseg[q], s1 = kseg[q + 1];
  for (int s = s0; s < s1; ++s) {
    const T *src = kstage + (int64_t)s * 9;
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] += src[i];
  }
  const int A = blk_a[q], B = blk_b[q], cp = blk_colpos[q];
  int cf[3];
  int nb = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) cf[k] = (dof_map[3 * (int64_t)B + k] >= 0) ? nb++ : -1;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int r = dof_map[3 * (int64_t)A + j];
    if (r < 0) continue;
    const int64_t base = (int64_t)rowptr[r] + cp;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (cf[k] >= 0) val[base + cf[k]] = acc[j * 3 + k];
  }
}

__global__ __launch_bounds__(kBlock) void k_gather_residual(int n_nodes, const int32_t *__restrict__ rseg,
                                                            const int32_t *__restrict__ dof_map,
                                                            const double *__restrict__ rstage,
                                                            double *__restrict__ R) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= n_nodes) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int s = rseg[A]; s < rseg[A + 1]; ++s) {
    acc[0] += rstage[(int64_t)s * 3 + 0];
    acc[1] += rstage[(int64_t)s * 3 + 1];
    acc[2] += rstage[(int64_t)s * 3 + 2];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int r = dof_map[3 * (int64_t)A + j];
    if (r >= 0) R[r] = acc[j];
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce_sum(int n, const double *__restrict__ partial,
                                                       double *__restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[0] = s;
}

// ------------------------------------------------------------------ SpMV
// SW lanes cooperate on one row (SW divides 64).
template <typename T, int SW>
__device__ __forceinline__ T row_dot(int row, int lane, const int32_t *__restrict__ rowptr,
                                     const int32_t *__restrict__ col, const T *__restrict__ val,
                                     const T *__restrict__ x) {
  T acc = T(0);
  const int e = rowptr[row + 1];
  for (int j = rowptr[row] + lane; j < e; j += SW) acc += val[j] * x[col[j]];
#pragma unroll
  for (int off = SW / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, SW);
  return acc;
}

template <typename T, int SW>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ col,
                                                 const T *__restrict__ val, const T *__restrict__ x,
                                                 T *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / SW), lane = (int)(t % SW);
  if (row >= n) return;
  const T v = row_dot<T, SW>(row, lane, rowptr, col, val, x);
  if (lane == 0) y[row] = v;
}

// ------------------------------------------------------------------ PCG
// Device-resident Jacobi-PCG without grid-wide fences.  Every reduction
// writes one partial per workgroup; the NEXT kernel (kernel boundary =>
// visibility) sums the partials redundantly in every workgroup, in the same
// fixed order, so all workgroups derive identical scalars (deterministic, no
// atomics, no L2 writeback).  Scalars carried across iterations live in
// partial rings indexed by the iteration parity.
//
// ws layout (float64):
//   W_ITER  iteration counter (written only by workgroup 0 of k_pcg_spmv)
//   W_DONE  converged flag     (written only by workgroup 0 of k_pcg_dir)
//   W_RZ, W_BZ  last r.z and b.(dinv b) for the host status
//   W_PQ .. W_PQ+G      p.q partials
//   W_RZP + (i&1)*G ..  r.z partials of iteration i (ring of 2)
//   W_BZP ..            b.(dinv b) partials
enum : int {
  PCG_G = 512,  // workgroups of every PCG kernel (partials per reduction)
  W_ITER = 0, W_DONE = 1, W_RZ = 2, W_BZ = 3,
  W_PQ = 16, W_RZP = W_PQ + PCG_G, W_BZP = W_RZP + 2 * PCG_G, W_END = W_BZP + PCG_G
};

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

// r = b - A x
template <typename T, int SW>
__global__ __launch_bounds__(kBlock) void k_spmv_resid(int n, const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col,
                                                       const T *__restrict__ val,
                                                       const T *__restrict__ x,
                                                       const T *__restrict__ b, T *__restrict__ r) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / SW), lane = (int)(t % SW);
  if (row >= n) return;
  const T v = row_dot<T, SW>(row, lane, rowptr, col, val, x);
  if (lane == 0) r[row] = b[row] - v;
}

// z = dinv r, p = z; partials of r.z (ring slot 0) and b.(dinv b)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const T *__restrict__ dinv,
                                                     const T *__restrict__ b,
                                                     const T *__restrict__ r, T *__restrict__ z,
                                                     T *__restrict__ p, double *__restrict__ ws) {
  double rz = 0.0, bz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T zi = dinv[i] * r[i];
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

// iteration i = ITER+1: q = A p, partials of p.q; workgroup 0 advances ITER
template <typename T, int SW>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(int n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col,
                                                     const T *__restrict__ val,
                                                     const T *__restrict__ p, T *__restrict__ q,
                                                     double *__restrict__ ws) {
  if (ws[W_DONE] != 0.0) return;
  constexpr int RPB = kBlock / SW;  // rows per workgroup pass
  double pq = 0.0;
  const int lane = threadIdx.x % SW;
  for (int64_t row0 = (int64_t)blockIdx.x * RPB; row0 < n; row0 += (int64_t)gridDim.x * RPB) {
    const int row = (int)(row0 + threadIdx.x / SW);
    if (row < n) {
      const T a = row_dot<T, SW>(row, lane, rowptr, col, val, p);
      if (lane == 0) {
        q[row] = a;
        pq += (double)a * (double)p[row];
      }
    }
  }
  write_partial(ws + W_PQ, pq);
  if (blockIdx.x == 0 && threadIdx.x == 0) ws[W_ITER] += 1.0;
}

// alpha = rz_{i-1} / pq; x += alpha p; r -= alpha q; z = dinv r; partials of r.z (ring i&1)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, const T *__restrict__ dinv,
                                                       T *__restrict__ x, T *__restrict__ r,
                                                       T *__restrict__ z, const T *__restrict__ p,
                                                       const T *__restrict__ q, double *__restrict__ ws) {
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  const double rz_old = sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_G, gridDim.x);
  const double pq = sum_partials(ws + W_PQ, gridDim.x);
  const T alpha = (T)(rz_old / pq);
  double rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    z[i] = zi;
    rz += (double)ri * (double)zi;
  }
  write_partial(ws + W_RZP + (it & 1) * PCG_G, rz);
}

// convergence test (rz_i / bz < tol, _linear_solver.py:37) and p = z + beta p
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, const T *__restrict__ z, T *__restrict__ p,
                                                    double *__restrict__ ws, double tol, int max_iter) {
  if (ws[W_DONE] != 0.0) return;
  const int it = (int)ws[W_ITER];
  const double rz_new = sum_partials(ws + W_RZP + (it & 1) * PCG_G, gridDim.x);
  const double rz_old = sum_partials(ws + W_RZP + ((it - 1) & 1) * PCG_G, gridDim.x);
  const double bz = sum_partials(ws + W_BZP, gridDim.x);
  const bool done = (rz_new / bz < tol) || (it >= max_iter) || !(rz_new == rz_new);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws[W_RZ] = rz_new;
    ws[W_BZ] = bz;
    if (done) ws[W_DONE] = 1.0;
  }
  if (done) return;
  const T beta = (T)(rz_new / rz_old);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi_inv(int n, const int32_t *__restrict__ diag_slot,
                                                       const T *__restrict__ val, T *__restrict__ dinv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = diag_slot[r];
  const T d = (s >= 0) ? fabs(val[s]) : T(0);
  dinv[r] = (d == T(0)) ? T(1) : T(1) / d;
}

// r64 = b - A32 x64 (float64 arithmetic), per-block partial r.r (grid-stride rows)
template <int SW>
__global__ __launch_bounds__(kBlock) void k_spmv_resid_mixed(int n, const int32_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ col,
                                                             const float *__restrict__ val,
                                                             const double *__restrict__ x,
                                                             const double *__restrict__ b,
                                                             double *__restrict__ r,
                                                             double *__restrict__ partial) {
  constexpr int RPB = kBlock / SW;
  const int lane = threadIdx.x % SW;
  double rr = 0.0;
  for (int64_t row0 = (int64_t)blockIdx.x * RPB; row0 < n; row0 += (int64_t)gridDim.x * RPB) {
    const int row = (int)(row0 + threadIdx.x / SW);
    if (row >= n) continue;
    double acc = 0.0;
    const int e = rowptr[row + 1];
    for (int j = rowptr[row] + lane; j < e; j += SW) acc += (double)val[j] * x[col[j]];
#pragma unroll
    for (int off = SW / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, SW);
    if (lane == 0) {
      const double ri = b[row] - acc;
      r[row] = ri;
      rr += ri * ri;
    }
  }
  const double s = block_sum(rr);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ------------------------------------------------------------------ vectors
__global__ __launch_bounds__(kBlock) void k_dof_axpy(int64_t n3, const int32_t *__restrict__ dof_map,
                                                     const double *__restrict__ U0, double alpha,
                                                     const double *__restrict__ d,
                                                     double *__restrict__ U) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n3) return;
  const int i = dof_map[t];
  U[t] = U0[t] + ((i >= 0) ? alpha * d[i] : 0.0);
}

template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const A *__restrict__ x, B *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (B)x[i];
}

__global__ __launch_bounds__(kBlock) void k_dot_partial(int n, const double *__restrict__ a,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ partial) {
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v += a[i] * b[i];
  const double s = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_explicit_update(int n_nodes, const int32_t *__restrict__ rseg,
                                                            const int32_t *__restrict__ dof_map,
                                                            const double *__restrict__ rstage,
                                                            const double *__restrict__ minv,
                                                            double dt_v, double dt_u,
                                                            double *__restrict__ acc_out,
                                                            double *__restrict__ vh,
                                                            double *__restrict__ U) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= n_nodes) return;
  double f[3] = {0.0, 0.0, 0.0};
  for (int s = rseg[A]; s < rseg[A + 1]; ++s) {
    f[0] += rstage[(int64_t)s * 3 + 0];
    f[1] += rstage[(int64_t)s * 3 + 1];
    f[2] += rstage[(int64_t)s * 3 + 2];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int r = dof_map[3 * (int64_t)A + j];
    if (r < 0) continue;
    const double a = -minv[r] * f[j];
    acc_out[r] = a;
    const double v = vh[r] + dt_v * a;
    vh[r] = v;
    U[3 * (int64_t)A + j] += dt_u * v;
  }
}

// y(float64) += alpha * x(T)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(int n, double alpha, const T *__restrict__ x,
                                                 double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += alpha * (double)x[i];
}

// per block: (sum x^2, max |x|)
__global__ __launch_bounds__(kBlock) void k_vec_stats(int n, const double *__restrict__ x,
                                                      double *__restrict__ partial) {
  double ss = 0.0, mx = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    ss += v * v;
    mx = fmax(mx, fabs(v));
  }
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) m = fmax(m, red[w]);
  __syncthreads();
  const double s = block_sum(ss);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x + 0] = s;
    partial[2 * blockIdx.x + 1] = m;
  }
}

// SpMV lanes-per-row from the average row length
inline int pick_sw(int n, int64_t nnz) {
  const double avg = n > 0 ? (double)nnz / n : 0.0;
  if (avg > 48) return 32;
  if (avg > 24) return 16;
  if (avg > 12) return 8;
  return 4;
}

template <typename T>
static void launch_spmv(int sw, int n, const int32_t *rp, const int32_t *c, const T *v, const T *x,
                        T *y, hipStream_t st) {
  const int grid = blocks_for((int64_t)n * sw);
  switch (sw) {
    case 4: k_spmv<T, 4><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, y); break;
    case 8: k_spmv<T, 8><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, y); break;
    case 16: k_spmv<T, 16><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, y); break;
    default: k_spmv<T, 32><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, y); break;
  }
}

template <typename T>
static void launch_resid(int sw, int n, const int32_t *rp, const int32_t *c, const T *v, const T *x,
                         const T *b, T *r, hipStream_t st) {
  const int grid = blocks_for((int64_t)n * sw);
  switch (sw) {
    case 4: k_spmv_resid<T, 4><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, b, r); break;
    case 8: k_spmv_resid<T, 8><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, b, r); break;
    case 16: k_spmv_resid<T, 16><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, b, r); break;
    default: k_spmv_resid<T, 32><<<grid, kBlock, 0, st>>>(n, rp, c, v, x, b, r); break;
  }
}

// grid for reduction kernels: bounded so partial arrays stay small
inline int red_grid(int64_t n) {
  const int g = blocks_for(n);
  return g < 2048 ? (g > 0 ? g : 1) : 2048;
}

template <typename T>
static void launch_pcg_spmv(int sw, int n, const int32_t *rp, const int32_t *c, const T *v,
                            const T *p, T *q, double *ws, hipStream_t st) {
  switch (sw) {
    case 4: k_pcg_spmv<T, 4><<<PCG_G, kBlock, 0, st>>>(n, rp, c, v, p, q, ws); break;
    case 8: k_pcg_spmv<T, 8><<<PCG_G, kBlock, 0, st>>>(n, rp, c, v, p, q, ws); break;
    case 16: k_pcg_spmv<T, 16><<<PCG_G, kBlock, 0, st>>>(n, rp, c, v, p, q, ws); break;
    default: k_pcg_spmv<T, 32><<<PCG_G, kBlock, 0, st>>>(n, rp, c, v, p, q, ws); break;
  }
}

template <typename T>
static int pcg_start_t(int n, int64_t nnz, const int32_t *rp, const int32_t *c, const T *v,
                       const T *dinv, const T *b, T *x, T *r, T *z, T *p, double *ws, hipStream_t st) {
  const int sw = pick_sw(n, nnz);
  launch_resid<T>(sw, n, rp, c, v, x, b, r, st);
  k_pcg_init<T><<<PCG_G, kBlock, 0, st>>>(n, dinv, b, r, z, p, ws);
  FEA_CHECK_LAUNCH("fea_pcg_start");
  return FEA_OK;
}

template <typename T>
static int pcg_iterate_t(int n, int64_t nnz, const int32_t *rp, const int32_t *c, const T *v,
                         const T *dinv, T *x, T *r, T *z, T *p, T *q, double *ws, double tol,
                         int max_iter, int iters, double *status_host, hipStream_t st) {
  const int sw = pick_sw(n, nnz);
  for (int it = 0; it < iters; ++it) {
    launch_pcg_spmv<T>(sw, n, rp, c, v, p, q, ws, st);
    k_pcg_update<T><<<PCG_G, kBlock, 0, st>>>(n, dinv, x, r, z, p, q, ws);
    k_pcg_dir<T><<<PCG_G, kBlock, 0, st>>>(n, z, p, ws, tol, max_iter);
  }
  FEA_CHECK_LAUNCH("fea_pcg_iterate");
  if (status_host != nullptr) {
    double h[4];
    FEA_HIP(hipMemcpyAsync(h, ws, sizeof(h), hipMemcpyDeviceToHost, st), "fea_pcg_iterate");
    FEA_HIP(hipStreamSynchronize(st), "fea_pcg_iterate");
    status_host[0] = h[W_ITER];
    status_host[1] = h[W_DONE];
    status_host[2] = h[W_RZ];
    status_host[3] = h[W_BZ];
  }
  return FEA_OK;
}

}  // namespace fea

using namespace fea;

extern "C" int fea_gather_csr(int dtype, int n_blocks, const int32_t *blk_a, const int32_t *blk_b,
                              const int32_t *blk_colpos, const int32_t *kseg,
                              const int32_t *dof_map, const int32_t *rowptr, const void *kstage,
                              void *csr_val, void *stream) {
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_gather_csr: bad dtype");
  FEA_CHECK_ARG(n_blocks >= 0, "fea_gather_csr: bad n_blocks");
  if (n_blocks == 0) return FEA_OK;
  FEA_CHECK_ARG(blk_a && blk_b && blk_colpos && kseg && dof_map && rowptr && kstage && csr_val,
                "fea_gather_csr: null argument");
  hipStream_t st = as_stream(stream);
  if (dtype == FEA_F64)
    k_gather_csr<double><<<blocks_for(n_blocks), kBlock, 0, st>>>(
        n_blocks, blk_a, blk_b, blk_colpos, kseg, dof_map, rowptr, (const double *)kstage, (double *)csr_val);
  else
    k_gather_csr<float><<<blocks_for(n_blocks), kBlock, 0, st>>>(
        n_blocks, blk_a, blk_b, blk_colpos, kseg, dof_map, rowptr, (const float *)kstage, (float *)csr_val);
  FEA_CHECK_LAUNCH("fea_gather_csr");
  return FEA_OK;
}

extern "C" int fea_gather_residual(int n_nodes, const int32_t *rseg, const int32_t *dof_map,
                                   const double *rstage, double *R, void *stream) {
  FEA_CHECK_ARG(n_nodes >= 0, "fea_gather_residual: bad n_nodes");
  if (n_nodes == 0) return FEA_OK;
  FEA_CHECK_ARG(rseg && dof_map && rstage && R, "fea_gather_residual: null argument");
  k_gather_residual<<<blocks_for(n_nodes), kBlock, 0, as_stream(stream)>>>(n_nodes, rseg, dof_map, rstage, R);
  FEA_CHECK_LAUNCH("fea_gather_residual");
  return FEA_OK;
}

extern "C" int fea_reduce_sum(int n, const double *partial, double *out, void *stream) {
  FEA_CHECK_ARG(n >= 0 && out != nullptr && (n == 0 || partial != nullptr), "fea_reduce_sum: bad arguments");
  k_reduce_sum<<<1, kBlock, 0, as_stream(stream)>>>(n, partial, out);
  FEA_CHECK_LAUNCH("fea_reduce_sum");
  return FEA_OK;
}

extern "C" int fea_spmv(int dtype, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                        const void *val, const void *x, void *y, void *stream) {
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_spmv: bad dtype");
  FEA_CHECK_ARG(n >= 0, "fea_spmv: bad n");
  if (n == 0) return FEA_OK;
  FEA_CHECK_ARG(rowptr && col && val && x && y, "fea_spmv: null argument");
  hipStream_t st = as_stream(stream);
  const int sw = pick_sw(n, nnz);
  if (dtype == FEA_F64)
    launch_spmv<double>(sw, n, rowptr, col, (const double *)val, (const double *)x, (double *)y, st);
  else
    launch_spmv<float>(sw, n, rowptr, col, (const float *)val, (const float *)x, (float *)y, st);
  FEA_CHECK_LAUNCH("fea_spmv");
  return FEA_OK;
}

extern "C" int fea_jacobi_inv(int dtype, int n, const int32_t *diag_slot, const void *val, void *dinv,
                              void *stream) {
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_jacobi_inv: bad dtype");
  FEA_CHECK_ARG(n >= 0, "fea_jacobi_inv: bad n");
  if (n == 0) return FEA_OK;
  FEA_CHECK_ARG(diag_slot && val && dinv, "fea_jacobi_inv: null argument");
  hipStream_t st = as_stream(stream);
  if (dtype == FEA_F64)
    k_jacobi_inv<double><<<blocks_for(n), kBlock, 0, st>>>(n, diag_slot, (const double *)val, (double *)dinv);
  else
    k_jacobi_inv<float><<<blocks_for(n), kBlock, 0, st>>>(n, diag_slot, (const float *)val, (float *)dinv);
  FEA_CHECK_LAUNCH("fea_jacobi_inv");
  return FEA_OK;
}

extern "C" int64_t fea_pcg_ws_size(int) { return (int64_t)W_END; }

extern "C" int fea_pcg_start(int dtype, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                             const void *val, const void *dinv, const void *b, void *x, void *r,
                             void *z, void *p, double *ws, void *stream) {
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_pcg_start: bad dtype");
  FEA_CHECK_ARG(n > 0 && rowptr && col && val && dinv && b && x && r && z && p && ws,
                "fea_pcg_start: bad arguments");
  hipStream_t st = as_stream(stream);
  if (dtype == FEA_F64)
    return pcg_start_t<double>(n, nnz, rowptr, col, (const double *)val, (const double *)dinv,
                               (const double *)b, (double *)x, (double *)r, (double *)z,
                               (double *)p, ws, st);
  return pcg_start_t<float>(n, nnz, rowptr, col, (const float *)val, (const float *)dinv,
                            (const float *)b, (float *)x, (float *)r, (float *)z, (float *)p, ws, st);
}

extern "C" int fea_pcg_iterate(int dtype, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                               const void *val, const void *dinv, void *x, void *r, void *z,
                               void *p, void *q, double *ws, double tol, int max_iter, int iters,
                               double *status_host, void *stream) {
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_pcg_iterate: bad dtype");
  FEA_CHECK_ARG(n > 0 && iters >= 0 && max_iter >= 1 && rowptr && col && val && dinv && x && r &&
                    z && p && q && ws,
                "fea_pcg_iterate: bad arguments");
  hipStream_t st = as_stream(stream);
  if (dtype == FEA_F64)
    return pcg_iterate_t<double>(n, nnz, rowptr, col, (const double *)val, (const double *)dinv,
                                 (double *)x, (double *)r, (double *)z, (double *)p, (double *)q, ws,
                                 tol, max_iter, iters, status_host, st);
  return pcg_iterate_t<float>(n, nnz, rowptr, col, (const float *)val, (const float *)dinv, (float *)x,
                              (float *)r, (float *)z, (float *)p, (float *)q, ws, tol, max_iter,
                              iters, status_host, st);
}

extern "C" int fea_vec_blocks(int n) { return red_grid(n); }

extern "C" int fea_spmv_residual_mixed(int n, const int32_t *rowptr, const int32_t *col,
                                       const float *val, const double *x, const double *b,
                                       double *r, double *partial, void *stream) {
  FEA_CHECK_ARG(n > 0 && rowptr && col && val && x && b && r && partial,
                "fea_spmv_residual_mixed: bad arguments");
  k_spmv_resid_mixed<8><<<red_grid(n), kBlock, 0, as_stream(stream)>>>(n, rowptr, col, val, x, b, r, partial);
  FEA_CHECK_LAUNCH("fea_spmv_residual_mixed");
  return FEA_OK;
}

extern "C" int fea_dof_axpy(int n_nodes, const int32_t *dof_map, const double *U0, double alpha,
                            const double *d, double *U, void *stream) {
  FEA_CHECK_ARG(n_nodes >= 0 && dof_map && U0 && d && U, "fea_dof_axpy: bad arguments");
  const int64_t n3 = 3LL * n_nodes;
  if (n3 == 0) return FEA_OK;
  k_dof_axpy<<<blocks_for(n3), kBlock, 0, as_stream(stream)>>>(n3, dof_map, U0, alpha, d, U);
  FEA_CHECK_LAUNCH("fea_dof_axpy");
  return FEA_OK;
}

extern "C" int fea_convert(int n, int dtype_x, const void *x, int dtype_y, void *y, void *stream) {
  FEA_CHECK_ARG(n >= 0 && x && y, "fea_convert: bad arguments");
  if (n == 0) return FEA_OK;
  hipStream_t st = as_stream(stream);
  const int g = blocks_for(n);
  if (dtype_x == FEA_F64 && dtype_y == FEA_F32)
    k_convert<double, float><<<g, kBlock, 0, st>>>(n, (const double *)x, (float *)y);
  else if (dtype_x == FEA_F32 && dtype_y == FEA_F64)
    k_convert<float, double><<<g, kBlock, 0, st>>>(n, (const float *)x, (double *)y);
  else if (dtype_x == FEA_F64 && dtype_y == FEA_F64)
    k_convert<double, double><<<g, kBlock, 0, st>>>(n, (const double *)x, (double *)y);
  else if (dtype_x == FEA_F32 && dtype_y == FEA_F32)
    k_convert<float, float><<<g, kBlock, 0, st>>>(n, (const float *)x, (float *)y);
  else
    FEA_CHECK_ARG(false, "fea_convert: bad dtype");
  FEA_CHECK_LAUNCH("fea_convert");
  return FEA_OK;
}

extern "C" int fea_dot_partial(int n, const double *a, const double *b, double *partial, void *stream) {
  FEA_CHECK_ARG(n > 0 && a && b && partial, "fea_dot_partial: bad arguments");
  k_dot_partial<<<red_grid(n), kBlock, 0, as_stream(stream)>>>(n, a, b, partial);
  FEA_CHECK_LAUNCH("fea_dot_partial");
  return FEA_OK;
}

extern "C" int fea_explicit_update(int n_nodes, const int32_t *rseg, const int32_t *dof_map,
                                   const double *rstage, const double *minv, double dt_v,
                                   double dt_u, double *a, double *vh, double *U, void *stream) {
  FEA_CHECK_ARG(n_nodes >= 0 && rseg && dof_map && rstage && minv && a && vh && U,
                "fea_explicit_update: bad arguments");
  if (n_nodes == 0) return FEA_OK;
  k_explicit_update<<<blocks_for(n_nodes), kBlock, 0, as_stream(stream)>>>(n_nodes, rseg, dof_map, rstage,
                                                                            minv, dt_v, dt_u, a, vh, U);
  FEA_CHECK_LAUNCH("fea_explicit_update");
  return FEA_OK;
}

extern "C" int fea_axpy(int n, double alpha, int dtype_x, const void *x, double *y, void *stream) {
  FEA_CHECK_ARG(n >= 0 && x && y && (dtype_x == FEA_F32 || dtype_x == FEA_F64), "fea_axpy: bad arguments");
  if (n == 0) return FEA_OK;
  if (dtype_x == FEA_F64)
    k_axpy<double><<<blocks_for(n), kBlock, 0, as_stream(stream)>>>(n, alpha, (const double *)x, y);
  else
    k_axpy<float><<<blocks_for(n), kBlock, 0, as_stream(stream)>>>(n, alpha, (const float *)x, y);
  FEA_CHECK_LAUNCH("fea_axpy");
  return FEA_OK;
}

extern "C" int fea_vec_stats(int n, const double *x, double *partial, void *stream) {
  FEA_CHECK_ARG(n > 0 && x && partial, "fea_vec_stats: bad arguments");
  k_vec_stats<<<red_grid(n), kBlock, 0, as_stream(stream)>>>(n, x, partial);
  FEA_CHECK_LAUNCH("fea_vec_stats");
  return FEA_OK;
}

extern "C" int fea_pcg_spmv(int dtype, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                            const void *val, const void *p, void *q, double *ws, void *stream) {
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_pcg_spmv: bad dtype");
  FEA_CHECK_ARG(n > 0 && rowptr && col && val && p && q && ws, "fea_pcg_spmv: bad arguments");
  hipStream_t st = as_stream(stream);
  const int sw = pick_sw(n, nnz);
  if (dtype == FEA_F64)
    launch_pcg_spmv<double>(sw, n, rowptr, col, (const double *)val, (const double *)p, (double *)q, ws, st);
  else
    launch_pcg_spmv<float>(sw, n, rowptr, col, (const float *)val, (const float *)p, (float *)q, ws, st);
  FEA_CHECK_LAUNCH("fea_pcg_spmv");
  return FEA_OK;
}
