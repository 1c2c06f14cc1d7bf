// Element kernels: residual + consistent tangent, residual only, strain energy,
// Gauss-point geometry export and lumped mass for C3D4 / C3D8 / C3D10.
#include "contact_device.h"
#include "element_traits.h"

namespace fea {

// Residual + tangent.  The stiffness is written with the first elasticity tensor of each
// Gauss point,
//   A_{jB,kD} = sum_{A,C} F_jA CC_ABCD F_kC + delta_jk S_BD   (times w det J),
// as K_ab[j][k] = sum_g sum_{B,D} dN_a,B A_{jB,kD} dN_b,D, which needs a third of the
// arithmetic of B^T D B formed per node pair.  A workgroup holds three waves and EPW
// elements (EPW nen <= 64): wave j serves displacement row j of every (element, node a)
// pair, one lane each, so j is uniform in a wave and each row's contraction is compiled
// with its own constant LDS offsets.
//   Phase 1: lane (e, g) of wave 0 forms the Gauss point's dN/dX, w P = w F S and w A in
// float64 (near incompressibility the deviatoric tangent is a small difference of
// bulk-sized terms) and stages them in LDS; A is symmetric under (jB) <-> (kD) and kept as
// its packed upper triangle (45 values).  (Splitting this phase over the three waves by
// entries of A made every part hold the registers of all three: 228-244 VGPRs.)
//   Phase 2: lane (e, a) of wave j forms v_kD = sum_B dN_a,B A_{jB,kD} (27 FMAs) and adds
// v . dN_b to K_ab[j][.] for every b (9 FMAs each), in T; its residual entry
// R_aj = sum_g sum_B (w P)_jB dN_a,B accumulates in float64.
template <int ET>
struct RkCfg;
template <>
struct RkCfg<4> {
  static constexpr int EPW = 16;
};
template <>
struct RkCfg<8> {
  static constexpr int EPW = 8;
};
template <>
struct RkCfg<10> {
  static constexpr int EPW = 6;
};
constexpr int kRkThreads = 192;  // three waves

// packed upper-triangle index of the symmetric 9 x 9 matrix A (I, J in 0..8)
__device__ __forceinline__ constexpr int sym9(int I, int J) {
  return I <= J ? I * 9 - (I * (I - 1)) / 2 + (J - I) : J * 9 - (J * (J - 1)) / 2 + (I - J);
}

template <typename T, int ET, int JR>
__device__ __forceinline__ void rk_contract(const double (*dn)[Elem<ET>::nen][3], const double (*wp)[9],
                                            const T *A, int a, T K[][3], double &Rj) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
#pragma unroll 1
  for (int g = 0; g < NGP; ++g) {
    const T *Ag = A + g * 45;
    const double d0 = dn[g][a][0], d1 = dn[g][a][1], d2 = dn[g][a][2];
    Rj += wp[g][JR * 3 + 0] * d0 + wp[g][JR * 3 + 1] * d1 + wp[g][JR * 3 + 2] * d2;
    const T t0 = T(d0), t1 = T(d1), t2 = T(d2);
    T v[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
      v[q] = t0 * Ag[sym9(3 * JR + 0, q)] + t1 * Ag[sym9(3 * JR + 1, q)] + t2 * Ag[sym9(3 * JR + 2, q)];
#pragma unroll
    for (int b = 0; b < NEN; ++b) {
      const T e0 = T(dn[g][b][0]), e1 = T(dn[g][b][1]), e2 = T(dn[g][b][2]);
#pragma unroll
      for (int k = 0; k < 3; ++k) K[b][k] += v[3 * k + 0] * e0 + v[3 * k + 1] * e1 + v[3 * k + 2] * e2;
    }
  }
}

// V is the type of the staged blocks (staged path) or of the CSR values (DIRECT).
// DIRECT: the elements are order[0 .. n_elem) (one colour: no two share a node) and each
// lane adds its row blocks into the CSR values (V) at rowptr[row] + colpos[(e,a,b)] + the
// free-column offset, instead of writing them to the staging; every CSR row has one
// writer per launch, and colours run one after another, so the sums are deterministic.
template <typename T, typename V, int ET, bool DIRECT>
__global__ __launch_bounds__(kRkThreads) void k_elem_rk(
    int n_elem, const int32_t *__restrict__ order, const int32_t *__restrict__ conn,
    const double *__restrict__ X, const double *__restrict__ U, MatParams mp, const int32_t *__restrict__ dst_k,
    V *__restrict__ kstage, const int32_t *__restrict__ dst_r, double *__restrict__ rstage,
    const int32_t *__restrict__ dof_map, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colpos,
    V *__restrict__ val, bool triple) {
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  constexpr int EPW = RkCfg<ET>::EPW;
  static_assert(EPW * NEN <= 64 && EPW * NGP <= 64, "one lane per (element, node) and per Gauss point");
  // per-element strides padded by one entry, so the elements of a wave start on
  // different banks
  constexpr int SA = NGP * 45 + 1;
  __shared__ double sx[EPW][NEN][3], su[EPW][NEN][3];
  __shared__ double sdn[EPW][NGP * NEN * 3 + 1];
  __shared__ double swp[EPW][NGP * 9 + 1];
  __shared__ T sA[EPW][SA];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // the row j of this wave
  const int r = threadIdx.x & 63;
  const int le = r / NEN, a = r - le * NEN;
  const int64_t i = (int64_t)blockIdx.x * EPW + le;
  const bool valid = r < EPW * NEN && i < n_elem;
  const int64_t e = !valid ? 0 : (DIRECT ? (int64_t)order[i] : i);
  int64_t n = 0;
  if (valid) {
    n = conn[e * NEN + a];
    sx[le][a][wv] = X[3 * n + wv];
    su[le][a][wv] = U[3 * n + wv];
  }
  // DIRECT: the CSR entries this lane adds to, and their values, loaded now so the
  // loads complete during the element arithmetic (no other lane of this launch writes
  // them: one colour, one writer per row)
  int64_t rb = -1;
  int cp[NEN];
  V old[NEN][3];
  if (DIRECT && valid) {
    const int row = dof_map[3 * n + wv];
    if (row >= 0) {
      rb = rowptr[row];
      const int64_t base = (e * NEN + a) * NEN;
#pragma unroll
      for (int b = 0; b < NEN; ++b) cp[b] = colpos[base + b];
#pragma unroll
      for (int b = 0; b < NEN; ++b)
#pragma unroll
        for (int k = 0; k < 3; ++k) old[b][k] = V(0);
      if (triple) {
#pragma unroll
        for (int b = 0; b < NEN; ++b)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            if (cp[b] >= 0) old[b][k] = val[rb + cp[b] + k];
      }
    }
  }
  __syncthreads();
  if (wv == 0 && r < EPW * NGP) {
    const int ge = r / NGP, g = r - ge * NGP;
    if ((int64_t)blockIdx.x * EPW + ge < n_elem) {
      double dNd[NEN][3];
      const double wd = gp_geometry<double, ET>(g, sx[ge], dNd);
      double Hd[9];
      disp_grad<double, NEN>(su[ge], dNd, Hd);
      GPState<double> sd;
      kinematics(Hd, sd);
      double Sd[6];
      stress(sd, mp, Sd);
      const TangentCoeffs tc = tangent_coeffs(sd, mp);
      double *dn = sdn[ge] + g * NEN * 3;
#pragma unroll
      for (int c = 0; c < NEN; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i) dn[c * 3 + i] = dNd[c][i];
      double *wp = swp[ge] + g * 9;
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
#pragma unroll
        for (int B = 0; B < 3; ++B) {
          double acc = 0.0;
#pragma unroll
          for (int A = 0; A < 3; ++A) acc += sd.F[jj * 3 + A] * Sd[vidx(A, B)];
          wp[jj * 3 + B] = wd * acc;
        }
      T *o = sA[ge] + g * 45;
      first_elasticity_packed<T>(sd, tc, Sd, wd, [&](int i, double v) { o[i] = (T)v; });
    }
  }
  __syncthreads();
  if (!valid) return;
  double Rj = 0.0;
  T K[NEN][3];
#pragma unroll
  for (int b = 0; b < NEN; ++b)
#pragma unroll
    for (int k = 0; k < 3; ++k) K[b][k] = T(0);
  const auto dn = reinterpret_cast<const double (*)[NEN][3]>(sdn[le]);
  const auto wp = reinterpret_cast<const double (*)[9]>(swp[le]);
  if (wv == 0)
    rk_contract<T, ET, 0>(dn, wp, sA[le], a, K, Rj);
  else if (wv == 1)
    rk_contract<T, ET, 1>(dn, wp, sA[le], a, K, Rj);
  else
    rk_contract<T, ET, 2>(dn, wp, sA[le], a, K, Rj);
  if (DIRECT) {
    if (rb >= 0) {
      if (triple) {
#pragma unroll
        for (int b = 0; b < NEN; ++b)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            if (cp[b] >= 0) val[rb + cp[b] + k] = (V)((T)old[b][k] + K[b][k]);
      } else {
        // partly fixed nodes: the free-column offset of each entry from dof_map
        int64_t at[NEN][3];
#pragma unroll
        for (int b = 0; b < NEN; ++b) {
          const int64_t nb = conn[e * NEN + b];
          int lk = 0;
#pragma unroll
          for (int k = 0; k < 3; ++k) at[b][k] = (cp[b] >= 0 && dof_map[3 * nb + k] >= 0) ? rb + cp[b] + lk++ : -1;
        }
#pragma unroll
        for (int b = 0; b < NEN; ++b)
#pragma unroll
          for (int k = 0; k < 3; ++k) old[b][k] = at[b][k] >= 0 ? val[at[b][k]] : V(0);
#pragma unroll
        for (int b = 0; b < NEN; ++b)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            if (at[b][k] >= 0) val[at[b][k]] = (V)((T)old[b][k] + K[b][k]);
      }
    }
  } else if (kstage != nullptr) {
    const int64_t base = (e * NEN + a) * NEN;
#pragma unroll
    for (int b = 0; b < NEN; ++b) {
      const int d = dst_k[base + b];
      if (d >= 0) {
        V *o = kstage + (int64_t)d * 9 + 3 * wv;
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = (V)K[b][k];
      }
    }
  }
  if (rstage != nullptr) {
    const int d = dst_r[e * NEN + a];
    if (d >= 0) rstage[(int64_t)d * 3 + wv] = Rj;
  }
}

// Nodal forces R_a = sum_g w P dN_a (P = F S) of element e, in T.
// In float, coordinates and displacements are taken relative to the element's first
// node, formed in double: the Jacobian and grad u only see differences (the shape
// function derivatives sum to zero), and a rigid translation of the element no longer
// costs float digits of the strain.  In double the absolute values are used.
// idx: the element's nen node indices into X and U (3 doubles per node: the global arrays, or
// a tile's nodes staged in LDS with tile-local indices).
template <typename T, int ET, typename I>
__device__ __forceinline__ void elem_force_at(const I *idx, const double *X, const double *U, const MatParams &mp,
                                              T R[Elem<ET>::nen][3]) {
  #pragma clang fp contract(on)
  constexpr int NEN = Elem<ET>::nen, NGP = Elem<ET>::ngp;
  constexpr bool kRel = sizeof(T) < sizeof(double);
  const int64_t n0 = idx[0];
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
    const int64_t n = idx[c];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[c][i] = (T)(X[3 * n + i] - x0[i]);
      u[c][i] = (T)(U[3 * n + i] - u0[i]);
    }
  }
  constexpr int C0 = kRel ? 1 : 0;  // node 0's relative x and u are exactly zero
#pragma unroll
  for (int c = 0; c < NEN; ++c) R[c][0] = R[c][1] = R[c][2] = T(-0.0);
  for (int g = 0; g < NGP; ++g) {
    T dN[NEN][3];
    const T w = gp_geometry<T, ET, C0>(g, x, dN);
    T H[9];
    disp_grad<T, NEN, C0>(u, dN, H);
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
}

template <typename T, int ET>
__device__ __forceinline__ void elem_force(int e, const int32_t *__restrict__ conn, const double *__restrict__ X,
                                           const double *__restrict__ U, const MatParams &mp,
                                           T R[Elem<ET>::nen][3]) {
  elem_force_at<T, ET, int32_t>(conn + (int64_t)e * Elem<ET>::nen, X, U, mp, R);
}

// One thread per element: R_a (float64) to rstage[dst_r[e*nen + a]*3 ..] (-1 skips).
template <typename T, int ET>
__global__ __launch_bounds__(kBlock) void k_elem_r(int n_elem, const int32_t *__restrict__ conn,
                                                   const double *__restrict__ X,
                                                   const double *__restrict__ U, MatParams mp,
                                                   const int32_t *__restrict__ dst_r,
                                                   double *__restrict__ rstage) {
  constexpr int NEN = Elem<ET>::nen;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  T R[NEN][3];
  elem_force<T, ET>(e, conn, X, U, mp, R);
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

// Element forces reduced per tile in LDS (the explicit loop's residual).  Workgroup t (TB
// threads: one element per thread for TE = TB) owns elements [t*TE, (t+1)*TE): their nodal forces (in T) go to LDS slot e_local*nen + a, then
// thread per local node p (partials tile_lbase[t] .. tile_lbase[t+1]) sums its slots
// pslot[pslot_ptr[p] .. pslot_ptr[p+1]) in that order in float64 into partial[3p ..].  The
// update sums a node's partials; no per-element staging reaches HBM.
template <int ET>
struct TileCfg;
template <>
struct TileCfg<4> {
  static constexpr int TE = 512, TB = 512;
};
template <>
struct TileCfg<8> {
  static constexpr int TE = 256, TB = 256;
};
template <>
struct TileCfg<10> {
  static constexpr int TE = 256, TB = 256;
};

// STAGE: the tile's nodes (tnode[tnode_ptr[t] .. tnode_ptr[t+1]), at most max_nodes) have
// their X and U copied to LDS first (coalesced), and lconn gives each element's nodes as
// tile-local indices; the element arithmetic is the same, on the same values.
template <typename T, int ET, bool STAGE>
__device__ __forceinline__ void tile_pass(int t, int n_elem, const int32_t *__restrict__ conn,
                                                         const double *__restrict__ X,
                                                         const double *__restrict__ U, MatParams mp,
                                                         const int32_t *__restrict__ tile_lbase,
                                                         const int32_t *__restrict__ pslot_ptr,
                                                         const uint16_t *__restrict__ pslot,
                                                         double *__restrict__ partial,
                                                         const int32_t *__restrict__ tnode_ptr,
                                                         const int32_t *__restrict__ tnode,
                                                         const uint16_t *__restrict__ lconn, int max_nodes) {
  constexpr int NEN = Elem<ET>::nen, TE = TileCfg<ET>::TE, TB = TileCfg<ET>::TB;
  constexpr int QT = TE * NEN, QPT = (QT + TB - 1) / TB;
  __shared__ T lds[TE * NEN * 3];
  __shared__ uint16_t sps[QT];          // the tile's pslot entries
  extern __shared__ double node_lds[];  // STAGE: X then U of the tile's nodes, 3 each
  const int e0 = t * TE;
  // the reduction's tables are loaded before the element work, so that their latency hides
  // behind it: the tile's pslot entries (at most TE * nen, one per free (element, node) slot)
  // and each thread's first partial range
  const int pb = tile_lbase[t], p1 = tile_lbase[t + 1];
  const int qb = pslot_ptr[pb], qn = min(pslot_ptr[p1] - qb, QT);
  uint16_t qv[QPT];
#pragma unroll
  for (int i = 0; i < QPT; ++i) {
    const int q = threadIdx.x + i * TB;
    qv[i] = q < qn ? pslot[qb + q] : uint16_t(0);
  }
  const int pf = pb + threadIdx.x;
  const int qf0 = pf < p1 ? pslot_ptr[pf] : 0, qf1 = pf < p1 ? pslot_ptr[pf + 1] : 0;
  if (STAGE) {
    const int n0 = tnode_ptr[t], nn = tnode_ptr[t + 1] - n0;
    double *xs = node_lds, *us = node_lds + 3 * max_nodes;
    for (int k = threadIdx.x; k < 3 * nn; k += TB) {
      const int64_t g = 3 * (int64_t)tnode[n0 + k / 3] + k % 3;
      xs[k] = X[g];
      us[k] = U[g];
    }
    __syncthreads();
  }
  for (int el = threadIdx.x; el < TE; el += TB) {
    const int e = e0 + el;
    if (e >= n_elem) break;
    T R[NEN][3];
    if (STAGE)
      elem_force_at<T, ET, uint16_t>(lconn + (int64_t)e * NEN, node_lds, node_lds + 3 * max_nodes, mp, R);
    else
      elem_force<T, ET>(e, conn, X, U, mp, R);
#pragma unroll
    for (int c = 0; c < NEN; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j) lds[(el * NEN + c) * 3 + j] = R[c][j];
  }
#pragma unroll
  for (int i = 0; i < QPT; ++i) {
    const int q = threadIdx.x + i * TB;
    if (q < qn) sps[q] = qv[i];
  }
  __syncthreads();
  for (int p = pf; p < p1; p += TB) {
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    const int q0 = (p == pf ? qf0 : pslot_ptr[p]) - qb, q1 = min((p == pf ? qf1 : pslot_ptr[p + 1]) - qb, qn);
    for (int q = max(q0, 0); q < q1; ++q) {
      const int sl = sps[q];
      f0 += (double)lds[sl * 3 + 0];
      f1 += (double)lds[sl * 3 + 1];
      f2 += (double)lds[sl * 3 + 2];
    }
    partial[3 * (int64_t)p + 0] = f0;
    partial[3 * (int64_t)p + 1] = f1;
    partial[3 * (int64_t)p + 2] = f2;
  }
}

// Workgroup b runs on XCD b % 8 (round-robin dispatch over the 8 XCDs); tile_of gives XCD x
// the contiguous tiles [x q + min(x, r), ...) (q = n / 8, r = n % 8) instead of every eighth
// tile, so that Morton-adjacent tiles, which share nodes, share an L2.
__device__ __forceinline__ int tile_of(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

template <typename T, int ET, bool STAGE>
__global__ __launch_bounds__(TileCfg<ET>::TB) void k_elem_r_tiles(int n_elem, const int32_t *__restrict__ conn,
                                                                  const double *__restrict__ X,
                                                                  const double *__restrict__ U, MatParams mp,
                                                                  const int32_t *__restrict__ tile_lbase,
                                                                  const int32_t *__restrict__ pslot_ptr,
                                                                  const uint16_t *__restrict__ pslot,
                                                                  double *__restrict__ partial,
                                                                  const int32_t *__restrict__ tnode_ptr,
                                                                  const int32_t *__restrict__ tnode,
                                                                  const uint16_t *__restrict__ lconn, int max_nodes) {
  tile_pass<T, ET, STAGE>(tile_of(blockIdx.x, gridDim.x), n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot,
                          partial, tnode_ptr, tnode, lconn, max_nodes);
}

// The explicit loop's fused first launch: workgroups [0, n_tiles) run the tile pass, the
// ones after them the contact penalty of one slave per wave (penalty_slave on the stored
// lists, or every facet when the lists are stale; it reads only X and U).  (Penalty
// workgroups first measured 1.4 % slower per step.)  The validity of the lists goes to
// lists_ok for the second launch's contact-node forces.
template <typename T, int ET>
__device__ __forceinline__ void tiles_pen_body(int n_tiles, int n_elem, const int32_t *__restrict__ conn,
                                               const double *__restrict__ X, const double *__restrict__ U,
                                               const MatParams &mp, const int32_t *__restrict__ tile_lbase,
                                               const int32_t *__restrict__ pslot_ptr,
                                               const uint16_t *__restrict__ pslot, double *__restrict__ partial,
                                               const ContactDev &cd) {
  if ((int)blockIdx.x < n_tiles) {
    tile_pass<T, ET, false>(tile_of(blockIdx.x, n_tiles), n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot,
                            partial, nullptr, nullptr, nullptr, 0);
    return;
  }
  constexpr int W = TileCfg<ET>::TB / 64;
  const int lane = threadIdx.x & 63;
  const int s = ((int)blockIdx.x - n_tiles) * W + (threadIdx.x >> 6);  // uniform over the wave
  if (s >= cd.n_slave) return;
  const bool valid = lists_valid(cd.dmax, cd.half_skin2, cd.sup_max, cd.kmax, lane);
  penalty_slave(cd, X, U, s, lane, valid);
  if (s == 0 && lane == 0) *cd.lists_ok = valid ? 1 : 0;
}

template <typename T, int ET>
__global__ __launch_bounds__(TileCfg<ET>::TB) void k_elem_r_tiles_pen(
    int n_tiles, int n_elem, const int32_t *__restrict__ conn, const double *__restrict__ X,
    const double *__restrict__ U, MatParams mp, const int32_t *__restrict__ tile_lbase,
    const int32_t *__restrict__ pslot_ptr, const uint16_t *__restrict__ pslot, double *__restrict__ partial,
    ContactDev cd) {
  tiles_pen_body<T, ET>(n_tiles, n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, cd);
}

// The float C3D4 form: the penalty code alone needs ~110 VGPRs, which would cut the tile
// pass (74) from 6 to 4 waves per SIMD; the register budget of 6 waves (80) is imposed and
// the penalty spills a few values instead (measured on the two-block step: 24.3k -> 27.1k
// steps/s).
__global__ __launch_bounds__(TileCfg<4>::TB) __attribute__((amdgpu_waves_per_eu(6))) void k_elem_r_tiles_pen_f4(
    int n_tiles, int n_elem, const int32_t *__restrict__ conn, const double *__restrict__ X,
    const double *__restrict__ U, MatParams mp, const int32_t *__restrict__ tile_lbase,
    const int32_t *__restrict__ pslot_ptr, const uint16_t *__restrict__ pslot, double *__restrict__ partial,
    ContactDev cd) {
  tiles_pen_body<float, 4>(n_tiles, n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, cd);
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

template <typename T, typename S, int ET>
static void launch_rk_t(int n, const int32_t *conn, const double *X, const double *U, MatParams mp,
                        const int32_t *dk, S *ks, const int32_t *dr, double *rs, hipStream_t st) {
  constexpr int EPW = RkCfg<ET>::EPW;
  const int grid = (int)(((int64_t)n + EPW - 1) / EPW);
  k_elem_rk<T, S, ET, false><<<grid, kRkThreads, 0, st>>>(n, nullptr, conn, X, U, mp, dk, ks, dr, rs, nullptr,
                                                          nullptr, nullptr, nullptr, false);
}

template <typename T, typename V, int ET>
static void launch_rk_direct_t(int n, const int32_t *order, const int32_t *conn, const double *X, const double *U,
                               MatParams mp, const int32_t *dr, double *rs, const int32_t *dof_map,
                               const int32_t *rowptr, const int32_t *colpos, V *val, bool triple, hipStream_t st) {
  constexpr int EPW = RkCfg<ET>::EPW;
  const int grid = (int)(((int64_t)n + EPW - 1) / EPW);
  k_elem_rk<T, V, ET, true><<<grid, kRkThreads, 0, st>>>(n, order, conn, X, U, mp, nullptr, nullptr, dr, rs,
                                                         dof_map, rowptr, colpos, val, triple);
}

template <typename T, typename V>
static int launch_rk_direct(int et, int n, const int32_t *order, const int32_t *conn, const double *X,
                            const double *U, MatParams mp, const int32_t *dr, double *rs, const int32_t *dof_map,
                            const int32_t *rowptr, const int32_t *colpos, V *val, bool triple, hipStream_t st) {
  switch (et) {
    case 4: launch_rk_direct_t<T, V, 4>(n, order, conn, X, U, mp, dr, rs, dof_map, rowptr, colpos, val, triple, st); break;
    case 8: launch_rk_direct_t<T, V, 8>(n, order, conn, X, U, mp, dr, rs, dof_map, rowptr, colpos, val, triple, st); break;
    case 10: launch_rk_direct_t<T, V, 10>(n, order, conn, X, U, mp, dr, rs, dof_map, rowptr, colpos, val, triple, st); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_tangent_direct");
  return FEA_OK;
}

template <typename T, typename S>
static int launch_rk(int et, int n, const int32_t *conn, const double *X, const double *U,
                     MatParams mp, const int32_t *dk, S *ks, const int32_t *dr, double *rs,
                     hipStream_t st) {
  switch (et) {
    case 4: launch_rk_t<T, S, 4>(n, conn, X, U, mp, dk, ks, dr, rs, st); break;
    case 8: launch_rk_t<T, S, 8>(n, conn, X, U, mp, dk, ks, dr, rs, st); break;
    case 10: launch_rk_t<T, S, 10>(n, conn, X, U, mp, dk, ks, dr, rs, st); break;
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

template <typename T, int ET>
static void launch_r_tiles_t(int n, const int32_t *conn, const double *X, const double *U, MatParams mp,
                             const int32_t *lb, const int32_t *pp, const uint16_t *ps, double *partial,
                             const int32_t *tnode_ptr, const int32_t *tnode, const uint16_t *lconn, int max_nodes,
                             hipStream_t st) {
  const int grid = (n + TileCfg<ET>::TE - 1) / TileCfg<ET>::TE;
  if (tnode_ptr != nullptr)
    k_elem_r_tiles<T, ET, true><<<grid, TileCfg<ET>::TB, (size_t)6 * max_nodes * sizeof(double), st>>>(
        n, conn, X, U, mp, lb, pp, ps, partial, tnode_ptr, tnode, lconn, max_nodes);
  else
    k_elem_r_tiles<T, ET, false><<<grid, TileCfg<ET>::TB, 0, st>>>(n, conn, X, U, mp, lb, pp, ps, partial, nullptr, nullptr,
                                                          nullptr, 0);
}

template <typename T>
static int launch_r_tiles(int et, int n, const int32_t *conn, const double *X, const double *U, MatParams mp,
                          const int32_t *tile_lbase, const int32_t *pslot_ptr, const uint16_t *pslot,
                          double *partial, const int32_t *tnode_ptr, const int32_t *tnode, const uint16_t *lconn,
                          int max_nodes, hipStream_t st) {
  switch (et) {
    case 4: launch_r_tiles_t<T, 4>(n, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, tnode_ptr, tnode, lconn, max_nodes, st); break;
    case 8: launch_r_tiles_t<T, 8>(n, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, tnode_ptr, tnode, lconn, max_nodes, st); break;
    case 10: launch_r_tiles_t<T, 10>(n, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, tnode_ptr, tnode, lconn, max_nodes, st); break;
  }
  FEA_CHECK_LAUNCH("fea_elem_residual_tiles");
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
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64 || dtype == FEA_F64_STAGE_F32,
                "fea_elem_residual_tangent: bad dtype %d", dtype);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr, "fea_elem_residual_tangent: bad arguments");
  FEA_CHECK_ARG((kstage == nullptr) == (dst_k == nullptr), "fea_elem_residual_tangent: kstage/dst_k mismatch");
  FEA_CHECK_ARG((rstage == nullptr) == (dst_r == nullptr), "fea_elem_residual_tangent: rstage/dst_r mismatch");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U, "fea_elem_residual_tangent: null input");
  const MatParams mp = mat_from(params_host);
  if (dtype == FEA_F64)
    return launch_rk<double, double>(elem_type, n_elem, conn, X, U, mp, dst_k, (double *)kstage, dst_r,
                                     rstage, as_stream(stream));
  if (dtype == FEA_F64_STAGE_F32)
    return launch_rk<double, float>(elem_type, n_elem, conn, X, U, mp, dst_k, (float *)kstage, dst_r,
                                    rstage, as_stream(stream));
  return launch_rk<float, float>(elem_type, n_elem, conn, X, U, mp, dst_k, (float *)kstage, dst_r, rstage,
                                 as_stream(stream));
}

extern "C" int fea_elem_tangent_direct(int elem_type, int dtype, int val_dtype, int n_elem, const int32_t *order,
                                       const int32_t *conn, const double *X, const double *U,
                                       const double *params_host, const int32_t *dof_map, const int32_t *rowptr,
                                       const int32_t *colpos, void *val, int triple, const int32_t *dst_r,
                                       double *rstage, void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_tangent_direct: bad element type %d", elem_type);
  FEA_CHECK_ARG((dtype == FEA_F64 && (val_dtype == FEA_F64 || val_dtype == FEA_F32)) ||
                    (dtype == FEA_F32 && val_dtype == FEA_F32),
                "fea_elem_tangent_direct: unsupported precision pair %d/%d", dtype, val_dtype);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr, "fea_elem_tangent_direct: bad arguments");
  FEA_CHECK_ARG((rstage == nullptr) == (dst_r == nullptr), "fea_elem_tangent_direct: rstage/dst_r mismatch");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(order && conn && X && U && dof_map && rowptr && colpos && val,
                "fea_elem_tangent_direct: null input");
  const MatParams mp = mat_from(params_host);
  hipStream_t st = as_stream(stream);
  if (dtype == FEA_F32)
    return launch_rk_direct<float, float>(elem_type, n_elem, order, conn, X, U, mp, dst_r, rstage, dof_map, rowptr,
                                          colpos, (float *)val, triple != 0, st);
  if (val_dtype == FEA_F32)
    return launch_rk_direct<double, float>(elem_type, n_elem, order, conn, X, U, mp, dst_r, rstage, dof_map, rowptr,
                                           colpos, (float *)val, triple != 0, st);
  return launch_rk_direct<double, double>(elem_type, n_elem, order, conn, X, U, mp, dst_r, rstage, dof_map, rowptr,
                                          colpos, (double *)val, triple != 0, st);
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

extern "C" int fea_elem_tile_size(int elem_type) {
  switch (elem_type) {
    case 4: return TileCfg<4>::TE;
    case 8: return TileCfg<8>::TE;
    case 10: return TileCfg<10>::TE;
  }
  return 0;
}

extern "C" int fea_elem_residual_tiles(int elem_type, int dtype, int n_elem, const int32_t *conn, const double *X,
                                       const double *U, const double *params_host, const int32_t *tile_lbase,
                                       const int32_t *pslot_ptr, const uint16_t *pslot, double *partial,
                                       const int32_t *tnode_ptr, const int32_t *tnode, const uint16_t *lconn,
                                       int max_tile_nodes, void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_residual_tiles: bad element type %d", elem_type);
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_elem_residual_tiles: bad dtype %d", dtype);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr, "fea_elem_residual_tiles: bad arguments");
  if (n_elem == 0) return FEA_OK;
  FEA_CHECK_ARG(conn && X && U && tile_lbase && pslot_ptr && pslot && partial, "fea_elem_residual_tiles: null input");
  FEA_CHECK_ARG(tnode_ptr == nullptr || (tnode && lconn && max_tile_nodes > 0 && max_tile_nodes <= 2048),
                "fea_elem_residual_tiles: tile node staging needs tnode, lconn and 1 <= max_tile_nodes <= 2048");
  const MatParams mp = mat_from(params_host);
  if (dtype == FEA_F64)
    return launch_r_tiles<double>(elem_type, n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, tnode_ptr,
                                  tnode, lconn, max_tile_nodes, as_stream(stream));
  return launch_r_tiles<float>(elem_type, n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, tnode_ptr,
                               tnode, lconn, max_tile_nodes, as_stream(stream));
}

template <typename T, int ET>
static void launch_r_tiles_pen_t(int n, const int32_t *conn, const double *X, const double *U, MatParams mp,
                                 const int32_t *lb, const int32_t *pp, const uint16_t *ps, double *partial,
                                 const ContactDev &cd, hipStream_t st) {
  constexpr int TE = TileCfg<ET>::TE, W = TileCfg<ET>::TB / 64;
  const int n_tiles = (n + TE - 1) / TE, n_pen = (cd.n_slave + W - 1) / W;
  if constexpr (ET == 4 && sizeof(T) == sizeof(float))
    k_elem_r_tiles_pen_f4<<<n_tiles + n_pen, TileCfg<ET>::TB, 0, st>>>(n_tiles, n, conn, X, U, mp, lb, pp, ps,
                                                                      partial, cd);
  else
    k_elem_r_tiles_pen<T, ET><<<n_tiles + n_pen, TileCfg<ET>::TB, 0, st>>>(n_tiles, n, conn, X, U, mp, lb, pp, ps,
                                                                            partial, cd);
}

template <typename T>
static void launch_r_tiles_pen(int et, int n, const int32_t *conn, const double *X, const double *U, MatParams mp,
                               const int32_t *lb, const int32_t *pp, const uint16_t *ps, double *partial,
                               const ContactDev &cd, hipStream_t st) {
  switch (et) {
    case 4: launch_r_tiles_pen_t<T, 4>(n, conn, X, U, mp, lb, pp, ps, partial, cd, st); break;
    case 8: launch_r_tiles_pen_t<T, 8>(n, conn, X, U, mp, lb, pp, ps, partial, cd, st); break;
    case 10: launch_r_tiles_pen_t<T, 10>(n, conn, X, U, mp, lb, pp, ps, partial, cd, st); break;
  }
}

extern "C" int fea_elem_residual_tiles_contact(int elem_type, int dtype, int n_elem, const int32_t *conn,
                                               const double *X, const double *U, const double *params_host,
                                               const int32_t *tile_lbase, const int32_t *pslot_ptr,
                                               const uint16_t *pslot, double *partial, const fea_contact_set *cs,
                                               void *stream) {
  FEA_CHECK_ARG(valid_elem(elem_type), "fea_elem_residual_tiles_contact: bad element type %d", elem_type);
  FEA_CHECK_ARG(dtype == FEA_F32 || dtype == FEA_F64, "fea_elem_residual_tiles_contact: bad dtype %d", dtype);
  FEA_CHECK_ARG(n_elem >= 0 && params_host != nullptr && X && U, "fea_elem_residual_tiles_contact: bad arguments");
  FEA_CHECK_ARG(n_elem == 0 || (conn && tile_lbase && pslot_ptr && pslot && partial),
                "fea_elem_residual_tiles_contact: null input");
  const int rc = contact_check_set(cs, "fea_elem_residual_tiles_contact");
  if (rc != FEA_OK) return rc;
  FEA_CHECK_ARG(cs->lists_ok != nullptr, "fea_elem_residual_tiles_contact: null lists_ok");
  const MatParams mp = mat_from(params_host);
  const ContactDev cd = contact_dev(*cs);
  if (dtype == FEA_F64)
    launch_r_tiles_pen<double>(elem_type, n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, cd,
                               as_stream(stream));
  else
    launch_r_tiles_pen<float>(elem_type, n_elem, conn, X, U, mp, tile_lbase, pslot_ptr, pslot, partial, cd,
                              as_stream(stream));
  FEA_CHECK_LAUNCH("fea_elem_residual_tiles_contact");
  return FEA_OK;
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
