// Assembly gathers, reductions, CSR SpMV, Jacobi-PCG with device-resident
// scalars, dof-mapped vector updates and the fused explicit update.
#include "contact_device.h"

namespace fea {

// ------------------------------------------------------------------ gathers

// One thread per node-pair block: sum the block's staged 3x3 contributions in
// slot order (fixed => deterministic) in the staging precision TS and write the
// free x free entries, rounded once to the matrix precision TV.
template <typename TS, typename TV>
__global__ __launch_bounds__(kBlock) void k_gather_csr(int n_blocks, const int32_t *__restrict__ blk_a,
                                                       const int32_t *__restrict__ blk_b,
                                                       const int32_t *__restrict__ blk_colpos,
                                                       const int32_t *__restrict__ kseg,
                                                       const int32_t *__restrict__ dof_map,
                                                       const int32_t *__restrict__ rowptr,
                                                       const TS *__restrict__ kstage,
                                                       TV *__restrict__ val) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_blocks) return;
  TS acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = TS(0);
  const int s0 = kseg[q], s1 = kseg[q + 1];
  for (int s = s0; s < s1; ++s) {
    const TS *src = kstage + (int64_t)s * 9;
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
      if (cf[k] >= 0) val[base + cf[k]] = (TV)acc[j * 3 + k];
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

// out[3A + j] = sum of stage rows rseg[A] .. rseg[A+1] (every node, no DOF map)
__global__ __launch_bounds__(kBlock) void k_gather_nodal(int n_nodes, const int32_t *__restrict__ seg,
                                                         const double *__restrict__ stage,
                                                         double *__restrict__ out) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= n_nodes) return;
  double f[3] = {0.0, 0.0, 0.0};
  for (int s = seg[A]; s < seg[A + 1]; ++s) {
    f[0] += stage[(int64_t)s * 3 + 0];
    f[1] += stage[(int64_t)s * 3 + 1];
    f[2] += stage[(int64_t)s * 3 + 2];
  }
  out[3 * (int64_t)A + 0] = f[0];
  out[3 * (int64_t)A + 1] = f[1];
  out[3 * (int64_t)A + 2] = f[2];
}

// P = false: node A's forces are rstage[s*3 ..] for s in seg[A] .. seg[A+1] (node-major
// staging); P = true: they are partial[list[s]*3 ..] (per-tile partial sums).  Summed in
// that order in float64.
template <bool P>
__device__ __forceinline__ void node_force(int A, const int32_t *__restrict__ seg, const int32_t *__restrict__ list,
                                           const double *__restrict__ rstage, double f[3]) {
  f[0] = f[1] = f[2] = 0.0;
  for (int s = seg[A]; s < seg[A + 1]; ++s) {
    const int64_t q = P ? (int64_t)list[s] * 3 : (int64_t)s * 3;
    f[0] += rstage[q + 0];
    f[1] += rstage[q + 1];
    f[2] += rstage[q + 2];
  }
}

// A node's update inputs: per DOF j its free-DOF row r[j] (-1: fixed), minv, vh and U.
struct NodeIn {
  int r[3];
  double mi[3], v[3], u[3];
};

// NODE_MINV: minv holds one value per node (the node's free DOFs share their lumped mass),
// otherwise one per free DOF.
template <bool NODE_MINV = false>
__device__ __forceinline__ void node_load(int A, const int32_t *__restrict__ dof_map, const double *__restrict__ minv,
                                          const double *__restrict__ vh, const double *__restrict__ U, NodeIn &in) {
  const double mn = NODE_MINV ? minv[A] : 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int r = dof_map[3 * (int64_t)A + j];
    in.r[j] = r;
    in.mi[j] = NODE_MINV ? mn : (r >= 0 ? minv[r] : 0.0);
    in.v[j] = r >= 0 ? vh[r] : 0.0;
    in.u[j] = U[3 * (int64_t)A + j];
  }
}

// The central-difference update of node A from its force f: per free DOF r of the node,
// a = -minv[r] f_j, vh[r] += dt_v a, U[3A+j] += dt_u vh[r]; u gets the node's new U.
__device__ __forceinline__ void node_step(int A, const double f[3], const NodeIn &in, double dt_v, double dt_u,
                                          double *__restrict__ acc_out, double *__restrict__ vh,
                                          double *__restrict__ U, double u[3]) {
  #pragma clang fp contract(on)
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int r = in.r[j];
    if (r < 0) {
      u[j] = in.u[j];
      continue;
    }
    const double a = -in.mi[j] * f[j];
    if (acc_out != nullptr) acc_out[r] = a;
    const double v = in.v[j] + dt_v * a;
    vh[r] = v;
    u[j] = in.u[j] + dt_u * v;
    U[3 * (int64_t)A + j] = u[j];
  }
}

__device__ __forceinline__ void node_cd_update(int A, const double f[3], const int32_t *__restrict__ dof_map,
                                               const double *__restrict__ minv, double dt_v, double dt_u,
                                               double *__restrict__ acc_out, double *__restrict__ vh,
                                               double *__restrict__ U, double u[3]) {
  NodeIn in;
  node_load(A, dof_map, minv, vh, U, in);
  node_step(A, f, in, dt_v, dt_u, acc_out, vh, U, u);
}

template <bool P>
__global__ __launch_bounds__(kBlock) void k_explicit_update(int n_nodes, const int32_t *__restrict__ seg,
                                                            const int32_t *__restrict__ list,
                                                            const int32_t *__restrict__ dof_map,
                                                            const double *__restrict__ rstage,
                                                            const double *__restrict__ minv,
                                                            const double *__restrict__ fc,
                                                            double dt_v, double dt_u,
                                                            double *__restrict__ acc_out,
                                                            double *__restrict__ vh,
                                                            double *__restrict__ U) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= n_nodes) return;
  double f[3], u[3];
  node_force<P>(A, seg, list, rstage, f);
  if (fc != nullptr) {
    f[0] -= fc[3 * (int64_t)A + 0];
    f[1] -= fc[3 * (int64_t)A + 1];
    f[2] -= fc[3 * (int64_t)A + 2];
  }
  node_cd_update(A, f, dof_map, minv, dt_v, dt_u, acc_out, vh, U, u);
}

// The explicit loop's fused second launch (minv per node; acc_out may be null: the
// accelerations are then not stored).  Workgroups [0, node_blocks): one thread per node
// not marked in cmask (bit A % 32 of word A / 32), its partials summed and the update made
// (a node without contact force: the same bits as subtracting fc's zero).  The workgroups
// after them: one wave per contact node i, its contact force (reaction_node, with the list
// validity the first launch stored) written to fc, then on lane 0 the node's update with
// f - fc (the order of k_explicit_update) and the tracking of its displacement since the
// rebuild into dmax, which the next step's penalty reads.
__global__ __launch_bounds__(kBlock) void k_explicit_update_contact(
    int n_nodes, int node_blocks, const int32_t *__restrict__ seg, const int32_t *__restrict__ list,
    const int32_t *__restrict__ dof_map, const double *__restrict__ partial, const double *__restrict__ minv,
    const uint32_t *__restrict__ cmask, const double *__restrict__ X, double *__restrict__ fc, double dt_v,
    double dt_u, double *__restrict__ acc_out, double *__restrict__ vh, double *__restrict__ U, ContactDev cd) {
  double f[3], u[3];
  if ((int)blockIdx.x < node_blocks) {
    const int A = blockIdx.x * kBlock + threadIdx.x;
    if (A >= n_nodes || ((cmask[A >> 5] >> (A & 31)) & 1u)) return;
    NodeIn in;
    node_force<true>(A, seg, list, partial, f);
    node_load<true>(A, dof_map, minv, vh, U, in);
    node_step(A, f, in, dt_v, dt_u, acc_out, vh, U, u);
    return;
  }
  const int lane = threadIdx.x & 63;
  const int i = ((int)blockIdx.x - node_blocks) * (kBlock / 64) + (threadIdx.x >> 6);  // uniform over the wave
  if (i >= cd.n_cnode) return;
  // the node's own force and update inputs do not depend on its contact force: loaded first,
  // their latency overlaps the reaction's
  const int A = cd.cnode[i];
  NodeIn in;
  node_force<true>(A, seg, list, partial, f);
  node_load<true>(A, dof_map, minv, vh, U, in);
  double r[3];
  reaction_node(cd, i, lane, *cd.lists_ok != 0, r);
  if (lane != 0) return;
  fc[3 * (int64_t)A + 0] = r[0];
  fc[3 * (int64_t)A + 1] = r[1];
  fc[3 * (int64_t)A + 2] = r[2];
  f[0] -= r[0];
  f[1] -= r[1];
  f[2] -= r[2];
  node_step(A, f, in, dt_v, dt_u, acc_out, vh, U, u);
  double p[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) p[j] = dadd(X[3 * (int64_t)A + j], u[j]);
  const double d2 = track_d2(p, cd.x0, i);
  dmax_raise(cd.dmax, i, d2);
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

// grid for reduction kernels: bounded so partial arrays stay small
inline int red_grid(int64_t n) {
  const int g = blocks_for(n);
  return g < 2048 ? (g > 0 ? g : 1) : 2048;
}

}  // namespace fea

using namespace fea;

extern "C" int fea_gather_csr(int stage_dtype, int val_dtype, int n_blocks, const int32_t *blk_a,
                              const int32_t *blk_b, const int32_t *blk_colpos, const int32_t *kseg,
                              const int32_t *dof_map, const int32_t *rowptr, const void *kstage,
                              void *csr_val, void *stream) {
  FEA_CHECK_ARG((stage_dtype == FEA_F64 && (val_dtype == FEA_F64 || val_dtype == FEA_F32)) ||
                    (stage_dtype == FEA_F32 && val_dtype == FEA_F32),
                "fea_gather_csr: unsupported precision pair");
  FEA_CHECK_ARG(n_blocks >= 0, "fea_gather_csr: bad n_blocks");
  if (n_blocks == 0) return FEA_OK;
  FEA_CHECK_ARG(blk_a && blk_b && blk_colpos && kseg && dof_map && rowptr && kstage && csr_val,
                "fea_gather_csr: null argument");
  hipStream_t st = as_stream(stream);
  const int g = blocks_for(n_blocks);
  if (stage_dtype == FEA_F64 && val_dtype == FEA_F64)
    k_gather_csr<double, double><<<g, kBlock, 0, st>>>(n_blocks, blk_a, blk_b, blk_colpos, kseg, dof_map, rowptr,
                                                      (const double *)kstage, (double *)csr_val);
  else if (stage_dtype == FEA_F64)
    k_gather_csr<double, float><<<g, kBlock, 0, st>>>(n_blocks, blk_a, blk_b, blk_colpos, kseg, dof_map, rowptr,
                                                     (const double *)kstage, (float *)csr_val);
  else
    k_gather_csr<float, float><<<g, kBlock, 0, st>>>(n_blocks, blk_a, blk_b, blk_colpos, kseg, dof_map, rowptr,
                                                    (const float *)kstage, (float *)csr_val);
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
                                   const double *rstage, const double *minv, const double *fc,
                                   double dt_v, double dt_u, double *a, double *vh, double *U,
                                   void *stream) {
  FEA_CHECK_ARG(n_nodes >= 0 && rseg && dof_map && rstage && minv && a && vh && U,
                "fea_explicit_update: bad arguments");
  if (n_nodes == 0) return FEA_OK;
  k_explicit_update<false><<<blocks_for(n_nodes), kBlock, 0, as_stream(stream)>>>(
      n_nodes, rseg, nullptr, dof_map, rstage, minv, fc, dt_v, dt_u, a, vh, U);
  FEA_CHECK_LAUNCH("fea_explicit_update");
  return FEA_OK;
}

extern "C" int fea_explicit_update_tiles(int n_nodes, const int32_t *node_pptr, const int32_t *node_plist,
                                         const int32_t *dof_map, const double *partial, const double *minv,
                                         const double *fc, double dt_v, double dt_u, double *a, double *vh,
                                         double *U, void *stream) {
  FEA_CHECK_ARG(n_nodes >= 0 && node_pptr && node_plist && dof_map && partial && minv && a && vh && U,
                "fea_explicit_update_tiles: bad arguments");
  if (n_nodes == 0) return FEA_OK;
  k_explicit_update<true><<<blocks_for(n_nodes), kBlock, 0, as_stream(stream)>>>(
      n_nodes, node_pptr, node_plist, dof_map, partial, minv, fc, dt_v, dt_u, a, vh, U);
  FEA_CHECK_LAUNCH("fea_explicit_update_tiles");
  return FEA_OK;
}

extern "C" int fea_explicit_update_tiles_contact(int n_nodes, const int32_t *node_pptr, const int32_t *node_plist,
                                                 const int32_t *dof_map, const double *partial, const double *minv_node,
                                                 const uint32_t *cmask, const fea_contact_set *cs, const double *X,
                                                 double *fc, double dt_v, double dt_u, double *a, double *vh,
                                                 double *U, void *stream) {
  FEA_CHECK_ARG(n_nodes > 0 && node_pptr && node_plist && dof_map && partial && minv_node && cmask && X && fc && vh &&
                    U,
                "fea_explicit_update_tiles_contact: bad arguments");
  const int rc = contact_check_set(cs, "fea_explicit_update_tiles_contact");
  if (rc != FEA_OK) return rc;
  FEA_CHECK_ARG(cs->lists_ok != nullptr, "fea_explicit_update_tiles_contact: null lists_ok");
  const int node_blocks = blocks_for(n_nodes);
  const int grid = node_blocks + (cs->n_cnode + kBlock / 64 - 1) / (kBlock / 64);
  k_explicit_update_contact<<<grid, kBlock, 0, as_stream(stream)>>>(n_nodes, node_blocks, node_pptr, node_plist,
                                                                    dof_map, partial, minv_node, cmask, X, fc, dt_v, dt_u,
                                                                    a, vh, U, contact_dev(*cs));
  FEA_CHECK_LAUNCH("fea_explicit_update_tiles_contact");
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


extern "C" int fea_gather_nodal(int n_nodes, const int32_t *seg, const double *stage, double *out, void *stream) {
  FEA_CHECK_ARG(n_nodes >= 0, "fea_gather_nodal: bad arguments");
  if (n_nodes == 0) return FEA_OK;
  FEA_CHECK_ARG(seg && stage && out, "fea_gather_nodal: null argument");
  k_gather_nodal<<<blocks_for(n_nodes), kBlock, 0, as_stream(stream)>>>(n_nodes, seg, stage, out);
  FEA_CHECK_LAUNCH("fea_gather_nodal");
  return FEA_OK;
}
