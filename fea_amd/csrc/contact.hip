// Node-to-surface penalty contact with a uniform-grid broad phase (new law;
// the reference's contact is a Gauss-point-pair barrier with a scipy cKDTree
// broad phase, contact.py:86-122, :385-401).
//
// Candidate predicate (exact, for bit-exact pair sets):
//   c_f  = (x_a + x_b + x_c) / 3            (float64, per component)
//   d2   = dx*dx + dy*dy + dz*dz, dx = x_s - c_f   (left to right, no FMA)
//   pair (s, f) is a candidate iff d2 <= r2.
// Deterministic throughout: bucket sorts order items by id inside each bucket
// and every force is summed by one thread in a fixed order (no float atomics).
#include "contact_device.h"

namespace fea {

__device__ __forceinline__ unsigned int cell_key(int64_t ix, int64_t iy, int64_t iz, unsigned int mask) {
  const uint64_t h = ((uint64_t)ix * 73856093ull) ^ ((uint64_t)iy * 19349663ull) ^ ((uint64_t)iz * 83492791ull);
  return (unsigned int)(h & mask);
}

__device__ __forceinline__ int64_t cell_of(double x, double o, double inv_h) {
  return (int64_t)floor((x - o) * inv_h);
}

// centroids and hashed cell keys of the master facets
__global__ __launch_bounds__(kBlock) void k_facet_keys(int nf, const int32_t *__restrict__ conn,
                                                       const double *__restrict__ X,
                                                       const double *__restrict__ U, double ox, double oy,
                                                       double oz, double inv_h, unsigned int mask,
                                                       double *__restrict__ cent, int32_t *__restrict__ key) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double a[3], b[3], c[3];
  node_pos(X, U, conn[3 * (int64_t)f + 0], a);
  node_pos(X, U, conn[3 * (int64_t)f + 1], b);
  node_pos(X, U, conn[3 * (int64_t)f + 2], c);
  double m[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) m[i] = dadd(dadd(a[i], b[i]), c[i]) / 3.0;
  cent[3 * (int64_t)f + 0] = m[0];
  cent[3 * (int64_t)f + 1] = m[1];
  cent[3 * (int64_t)f + 2] = m[2];
  key[f] = (int32_t)cell_key(cell_of(m[0], ox, inv_h), cell_of(m[1], oy, inv_h), cell_of(m[2], oz, inv_h), mask);
}

// ---------------------------------------------------------------- bucket sort
__global__ __launch_bounds__(kBlock) void k_fill_int(int n, int32_t *__restrict__ a, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_bucket_count(int n, const int32_t *__restrict__ keys,
                                                         int32_t *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keys[i] >= 0) atomicAdd(&cnt[keys[i]], 1);
}

// exclusive scan of cnt[0..nb) into start[0..nb], one workgroup of 1024 threads; cnt is
// then cleared (it is reused as the fill cursor): each thread clears the chunk it alone read
__global__ __launch_bounds__(1024) void k_bucket_scan(int nb, int32_t *__restrict__ cnt,
                                                      int32_t *__restrict__ start) {
  scan_1024(nb, cnt, start);
  const int chunk = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * chunk, b1 = min(nb, b0 + chunk);
  for (int b = b0; b < b1; ++b) cnt[b] = 0;
}

__global__ __launch_bounds__(kBlock) void k_bucket_fill(int n, const int32_t *__restrict__ keys,
                                                        const int32_t *__restrict__ start,
                                                        int32_t *__restrict__ cursor,
                                                        int32_t *__restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || keys[i] < 0) return;
  const int k = keys[i];
  items[start[k] + atomicAdd(&cursor[k], 1)] = i;
}

// insertion sort of each bucket by item id (buckets are small)
__global__ __launch_bounds__(kBlock) void k_bucket_sort(int nb, const int32_t *__restrict__ start,
                                                        int32_t *__restrict__ items) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int s0 = start[b], s1 = start[b + 1];
  for (int i = s0 + 1; i < s1; ++i) {
    const int32_t v = items[i];
    int j = i - 1;
    while (j >= s0 && items[j] > v) {
      items[j + 1] = items[j];
      --j;
    }
    items[j + 1] = v;
  }
}

// ---------------------------------------------------------------- search
constexpr int kMaxCand = 64;

// One wave64 per slave node.  Lane c < 27 takes neighbour cell c of the slave's
// cell and scans its bucket unless a lower lane hashed to the same bucket (so
// every bucket is scanned once); a wave prefix sum of the lanes' match counts
// places the candidates in LDS (at most kMaxCand kept, all counted), and a
// 64-lane bitonic sort orders them by facet id.  Same pair set and output as a
// serial scan of the 27 cells followed by a sort.
constexpr int kSearchWaves = kBlock / 64;

// The candidates of the wave's slave position p (wave-uniform): returns this lane's entry of
// the ascending candidate list (INT32_MAX past the kept ones) and the total count in `total`.
// buf is the wave's kMaxCand-entry LDS row.
__device__ __forceinline__ int32_t wave_candidates(const double p[3], const double *__restrict__ cent,
                                                   const int32_t *__restrict__ bstart,
                                                   const int32_t *__restrict__ bitems, double ox, double oy,
                                                   double oz, double inv_h, unsigned int mask, double r2,
                                                   int32_t *buf, int lane, int &total) {
  const int64_t cx = cell_of(p[0], ox, inv_h), cy = cell_of(p[1], oy, inv_h), cz = cell_of(p[2], oz, inv_h);
  const bool active = lane < 27;
  unsigned int k = 0xffffffffu;
  if (active) k = cell_key(cx + (lane % 3) - 1, cy + (lane / 3) % 3 - 1, cz + lane / 9 - 1, mask);
  bool dup = false;
  for (int j = 0; j < 27; ++j) {
    const unsigned int kj = __shfl(k, j);
    dup = dup || (j < lane && kj == k);
  }
  const bool scan = active && !dup;
  int b0 = 0, b1 = 0;
  if (scan) {
    b0 = bstart[k];
    b1 = bstart[k + 1];
  }
  auto hit = [&](int f) {
    const double ddx = p[0] - cent[3 * (int64_t)f + 0];
    const double ddy = p[1] - cent[3 * (int64_t)f + 1];
    const double ddz = p[2] - cent[3 * (int64_t)f + 2];
    return dadd(dadd(dmul(ddx, ddx), dmul(ddy, ddy)), dmul(ddz, ddz)) <= r2;
  };
  int cnt = 0;
  for (int it = b0; it < b1; ++it) cnt += hit(bitems[it]) ? 1 : 0;
  int incl = cnt;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  total = __shfl(incl, 63);
  int pos = incl - cnt;
  for (int it = b0; it < b1; ++it) {
    const int f = bitems[it];
    if (hit(f)) {
      if (pos < kMaxCand) buf[pos] = f;
      ++pos;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int nloc = total < kMaxCand ? total : kMaxCand;
  int32_t v = lane < nloc ? buf[lane] : INT32_MAX;
  for (int size = 2; size <= 64; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int32_t o = __shfl_xor(v, stride);
      const bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0);
      v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
    }
  return v;
}

// the slave's candidate list and count (and the overflow maximum)
__device__ __forceinline__ void store_candidates(int s, int lane, int32_t v, int total, int kmax,
                                                 int32_t *__restrict__ cand, int32_t *__restrict__ count,
                                                 int32_t *__restrict__ max_count) {
  const int nloc = total < kMaxCand ? total : kMaxCand;
  const int nout = nloc < kmax ? nloc : kmax;
  if (lane < nout) cand[(int64_t)s * kmax + lane] = v;
  if (lane == 0) {
    count[s] = total;
    if (max_count != nullptr && total > kmax) atomicMax(max_count, total);
  }
}

__global__ __launch_bounds__(kBlock) void k_contact_search(
    int ns, const int32_t *__restrict__ slave, const double *__restrict__ X, const double *__restrict__ U,
    const double *__restrict__ cent, const int32_t *__restrict__ bstart, const int32_t *__restrict__ bitems,
    double ox, double oy, double oz, double inv_h, unsigned int mask, double r2, int kmax,
    int32_t *__restrict__ cand, int32_t *__restrict__ count, int32_t *__restrict__ max_count) {
  __shared__ int32_t buf[kSearchWaves][kMaxCand];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * kSearchWaves + wave;  // uniform across the wave
  if (s >= ns) return;
  double p[3];
  node_pos(X, U, slave[s], p);
  int total = 0;
  const int32_t v = wave_candidates(p, cent, bstart, bitems, ox, oy, oz, inv_h, mask, r2, buf[wave], lane, total);
  store_candidates(s, lane, v, total, kmax, cand, count, max_count);
}

// One wave per slave node: lane i evaluates candidates i, i + 64, ... of the slave's sorted
// candidate list and keeps its best by (tier ascending, key descending, index ascending); a
// butterfly over the wave picks the best of the lanes by the same order, which is the
// candidate a serial scan of the list keeps (the first of equal ones).
constexpr int kPenWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_contact_penalty(
    int ns, const int32_t *__restrict__ slave, const double *__restrict__ X, const double *__restrict__ U,
    const int32_t *__restrict__ conn, const int32_t *__restrict__ cand, const int32_t *__restrict__ count,
    int kmax, double k_pen, const double *__restrict__ area, double depth_max, int32_t *__restrict__ chosen,
    double *__restrict__ force,
    double *__restrict__ weight, double *__restrict__ energy_partial) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * kPenWaves + (threadIdx.x >> 6);
  double e = 0.0;
  if (s < ns) {  // uniform over the wave
    double p[3];
    node_pos(X, U, slave[s], p);
    int bf = -1, bi = INT_MAX, btier = 2;
    double bg = 0.0, bkey = 0.0, bn[3] = {0, 0, 0}, bw[3] = {0, 0, 0};
    const int nc = min(count[s], kmax);
    for (int i = lane; i < nc; i += 64) {
      const int f = cand[(int64_t)s * kmax + i];
      PenCand c;
      if (!pen_eval(p, X, U, conn, f, depth_max, c)) continue;
      if (c.tier < btier || (c.tier == btier && c.key > bkey)) {
        bf = f;
        bi = i;
        btier = c.tier;
        bkey = c.key;
        bg = c.g;
        bn[0] = c.n[0];
        bn[1] = c.n[1];
        bn[2] = c.n[2];
        bw[0] = c.w[0];
        bw[1] = c.w[1];
        bw[2] = c.w[2];
      }
    }
    e = pen_finish(s, lane, bf, bi, btier, bkey, bg, bn, bw, k_pen, area, chosen, force, weight);
  }
  if (energy_partial != nullptr) {
    const double t = block_sum(e);
    if (threadIdx.x == 0) energy_partial[blockIdx.x] = t;
  }
}

// master-node reactions: for master node m, over its facets (static adjacency,
// with m's local vertex in each) and the slaves that chose each facet (bucket
// sorted by slave id): fc[m] -= F_s * w_s(local vertex).
__global__ __launch_bounds__(kBlock) void k_contact_master(
    int nm, const int32_t *__restrict__ mnode, const int32_t *__restrict__ nf_ptr,
    const int32_t *__restrict__ nf_facet, const int32_t *__restrict__ nf_vert,
    const int32_t *__restrict__ sb_start, const int32_t *__restrict__ sb_items,
    const double *__restrict__ force, const double *__restrict__ weight, double *__restrict__ fc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = nf_ptr[i]; t < nf_ptr[i + 1]; ++t) {
    const int f = nf_facet[t], v = nf_vert[t];
    for (int it = sb_start[f]; it < sb_start[f + 1]; ++it) {
      const int s = sb_items[it];
      const double w = weight[3 * (int64_t)s + v];
      acc[0] -= force[3 * (int64_t)s + 0] * w;
      acc[1] -= force[3 * (int64_t)s + 1] * w;
      acc[2] -= force[3 * (int64_t)s + 2] * w;
    }
  }
  const int64_t m = mnode[i];
  fc[3 * m + 0] += acc[0];
  fc[3 * m + 1] += acc[1];
  fc[3 * m + 2] += acc[2];
}

__global__ __launch_bounds__(kBlock) void k_contact_slave(int ns, const int32_t *__restrict__ slave,
                                                          const double *__restrict__ force,
                                                          double *__restrict__ fc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int64_t n = slave[s];
  fc[3 * n + 0] += force[3 * (int64_t)s + 0];
  fc[3 * n + 1] += force[3 * (int64_t)s + 1];
  fc[3 * n + 2] += force[3 * (int64_t)s + 2];
}

__global__ __launch_bounds__(kBlock) void k_fill_dbl(int64_t n, double *__restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// dmax = max over contact nodes of |x - x0|^2
__global__ __launch_bounds__(kBlock) void k_contact_track(int nc, const int32_t *__restrict__ cnode,
                                                          const double *__restrict__ X,
                                                          const double *__restrict__ U,
                                                          const double *__restrict__ x0,
                                                          unsigned long long *__restrict__ dmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < nc) {
    double p[3];
    node_pos(X, U, cnode[i], p);
    d2 = track_d2(p, x0, i);
  }
  for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_down(d2, off, 64));
  if ((threadIdx.x & 63) == 0) dmax_raise(dmax, (int)(i >> 6), d2);
}

// x0 = current positions of the contact nodes, dmax = 0 (after a rebuild's search)
__global__ __launch_bounds__(kBlock) void k_contact_snapshot(int nc, const int32_t *__restrict__ cnode,
                                                             const double *__restrict__ X,
                                                             const double *__restrict__ U, double *__restrict__ x0,
                                                             unsigned long long *__restrict__ dmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kDmaxSlots) dmax[(int64_t)i * kDmaxStride] = 0ull;  // the launch has >= 256 threads
  if (i >= nc) return;
  double p[3];
  node_pos(X, U, cnode[i], p);
  x0[3 * (int64_t)i + 0] = p[0];
  x0[3 * (int64_t)i + 1] = p[1];
  x0[3 * (int64_t)i + 2] = p[2];
}

// One wave per slave (penalty_slave).  lists_ok (when set) gets the validity of the stored
// lists at this launch (the reactions of a later launch read it).
__global__ __launch_bounds__(kBlock) void k_contact_penalty_v(ContactDev c, const double *__restrict__ X,
                                                              const double *__restrict__ U,
                                                              double *__restrict__ energy_partial) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * kPenWaves + (threadIdx.x >> 6);
  double e = 0.0;
  if (s < c.n_slave) {
    const bool valid = lists_valid(c.dmax, c.half_skin2, c.sup_max, c.kmax, lane);
    e = penalty_slave(c, X, U, s, lane, valid);
    if (s == 0 && lane == 0 && c.lists_ok != nullptr) *c.lists_ok = valid ? 1 : 0;
  }
  if (energy_partial != nullptr) {
    const double t = block_sum(e);
    if (threadIdx.x == 0) energy_partial[blockIdx.x] = t;
  }
}

// Contact node forces, one wave per contact node (reaction_node); fc[node] is written once.
constexpr int kReactWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_contact_reactions_v(ContactDev c, double *__restrict__ fc) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kReactWaves + (threadIdx.x >> 6);  // uniform over the wave
  if (i >= c.n_cnode) return;
  const bool valid = lists_valid(c.dmax, c.half_skin2, c.sup_max, c.kmax, lane);
  double r[3];
  reaction_node(c, i, lane, valid, r);
  if (lane == 0) {
    const int64_t n = c.cnode[i];
    fc[3 * n + 0] = r[0];
    fc[3 * n + 1] = r[1];
    fc[3 * n + 2] = r[2];
  }
}

// per master node: the number of stored (slave, facet) pairs over its facets
__global__ __launch_bounds__(kBlock) void k_mlist_count(int nm, const int32_t *__restrict__ nf_ptr,
                                                        const int32_t *__restrict__ nf_facet,
                                                        const int32_t *__restrict__ inv_start,
                                                        int32_t *__restrict__ m_ptr) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  int c = 0;
  for (int t = nf_ptr[m]; t < nf_ptr[m + 1]; ++t) {
    const int f = nf_facet[t];
    c += inv_start[f + 1] - inv_start[f];
  }
  m_ptr[m] = c;
}

__global__ __launch_bounds__(1024) void k_scan_inplace(int n, int32_t *__restrict__ a) { scan_1024(n, a, a); }

// per master node: its items (slave, facet * 4 + local vertex) in facet order, slaves ascending
__global__ __launch_bounds__(kBlock) void k_mlist_fill(int nm, const int32_t *__restrict__ nf_ptr,
                                                       const int32_t *__restrict__ nf_facet,
                                                       const int32_t *__restrict__ nf_vert,
                                                       const int32_t *__restrict__ inv_start,
                                                       const int32_t *__restrict__ inv_items,
                                                       const int32_t *__restrict__ m_ptr,
                                                       int32_t *__restrict__ m_items) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  int64_t pos = m_ptr[m];
  for (int t = nf_ptr[m]; t < nf_ptr[m + 1]; ++t) {
    const int f = nf_facet[t], v = nf_vert[t];
    for (int it = inv_start[f]; it < inv_start[f + 1]; ++it) {
      m_items[2 * pos] = inv_items[it];
      m_items[2 * pos + 1] = f * 4 + v;
      ++pos;
    }
  }
}

// facet -> slave pair counts of the stored candidate lists
__global__ __launch_bounds__(kBlock) void k_inv_count(int ns, const int32_t *__restrict__ cand,
                                                      const int32_t *__restrict__ count, int kmax,
                                                      int32_t *__restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int n = min(count[s], kmax);
  for (int i = 0; i < n; ++i) atomicAdd(&cnt[cand[(int64_t)s * kmax + i]], 1);
}

__global__ __launch_bounds__(kBlock) void k_inv_fill(int ns, const int32_t *__restrict__ cand,
                                                     const int32_t *__restrict__ count, int kmax,
                                                     const int32_t *__restrict__ start, int32_t *__restrict__ cursor,
                                                     int32_t *__restrict__ items) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int n = min(count[s], kmax);
  for (int i = 0; i < n; ++i) {
    const int f = cand[(int64_t)s * kmax + i];
    items[start[f] + atomicAdd(&cursor[f], 1)] = s;
  }
}

static int bucket_sort(int n, const int32_t *keys, int nb, int32_t *start, int32_t *items, int32_t *work,
                       hipStream_t st) {
  k_fill_int<<<blocks_for(nb), kBlock, 0, st>>>(nb, work, 0);
  k_bucket_count<<<blocks_for(n > 0 ? n : 1), kBlock, 0, st>>>(n, keys, work);
  k_bucket_scan<<<1, 1024, 0, st>>>(nb, work, start);  // leaves work zeroed for the fill cursor
  k_bucket_fill<<<blocks_for(n > 0 ? n : 1), kBlock, 0, st>>>(n, keys, start, work, items);
  k_bucket_sort<<<blocks_for(nb), kBlock, 0, st>>>(nb, start, items);
  FEA_CHECK_LAUNCH("bucket sort");
  return FEA_OK;
}

}  // namespace fea

using namespace fea;

extern "C" int fea_contact_grid_build(int n_facets, const int32_t *conn, const double *X, const double *U,
                                      const double *origin_host, double cell, int table_bits,
                                      double *centroid, int32_t *key, int32_t *bucket_start,
                                      int32_t *bucket_items, int32_t *work, void *stream) {
  FEA_CHECK_ARG(n_facets >= 0 && origin_host && cell > 0.0 && table_bits >= 1 && table_bits <= 28,
                "fea_contact_grid_build: bad arguments");
  FEA_CHECK_ARG(n_facets == 0 || (conn && X && U && centroid && key && bucket_start && bucket_items && work),
                "fea_contact_grid_build: null argument");
  hipStream_t st = as_stream(stream);
  const unsigned int nb = 1u << table_bits;
  if (n_facets > 0)
    k_facet_keys<<<blocks_for(n_facets), kBlock, 0, st>>>(n_facets, conn, X, U, origin_host[0], origin_host[1],
                                                         origin_host[2], 1.0 / cell, nb - 1, centroid, key);
  FEA_CHECK_LAUNCH("fea_contact_grid_build");
  return bucket_sort(n_facets, key, (int)nb, bucket_start, bucket_items, work, st);
}

extern "C" int fea_contact_search(int n_slave, const int32_t *slave, const double *X, const double *U,
                                  const double *centroid, const int32_t *bucket_start,
                                  const int32_t *bucket_items, const double *origin_host, double cell,
                                  int table_bits, double radius, int kmax, int32_t *cand, int32_t *count,
                                  int32_t *max_count, void *stream) {
  FEA_CHECK_ARG(n_slave >= 0 && origin_host && cell > 0.0 && radius > 0.0 && radius <= cell &&
                    kmax >= 1 && kmax <= kMaxCand && table_bits >= 1 && table_bits <= 28,
                "fea_contact_search: bad arguments (radius must not exceed the cell size, kmax <= 64)");
  if (n_slave == 0) return FEA_OK;
  FEA_CHECK_ARG(slave && X && U && centroid && bucket_start && bucket_items && cand && count,
                "fea_contact_search: null argument");
  const unsigned int nb = 1u << table_bits;
  k_contact_search<<<(n_slave + kSearchWaves - 1) / kSearchWaves, kBlock, 0, as_stream(stream)>>>(
      n_slave, slave, X, U, centroid, bucket_start, bucket_items, origin_host[0], origin_host[1], origin_host[2],
      1.0 / cell, nb - 1, radius * radius, kmax, cand, count, max_count);
  FEA_CHECK_LAUNCH("fea_contact_search");
  return FEA_OK;
}

extern "C" int fea_contact_blocks(int n_slave) {
  return n_slave > 0 ? (n_slave + kPenWaves - 1) / kPenWaves : 0;  // one wave per slave
}

extern "C" int fea_contact_forces(int n_slave, const int32_t *slave, const double *X, const double *U,
                                  const int32_t *conn, int n_facets, const int32_t *cand,
                                  const int32_t *count, int kmax, double penalty, const double *slave_area,
                                  double depth_max, int n_master, const int32_t *master_nodes,
                                  const int32_t *nf_ptr,
                                  const int32_t *nf_facet, const int32_t *nf_vert, int32_t *chosen,
                                  double *force, double *weight, int32_t *sb_start, int32_t *sb_items,
                                  int32_t *work, int64_t n_dof, double *fc, double *energy_partial,
                                  void *stream) {
  FEA_CHECK_ARG(n_slave >= 0 && n_facets >= 0 && n_master >= 0 && kmax >= 1 && kmax <= kMaxCand && n_dof >= 0,
                "fea_contact_forces: bad arguments");
  FEA_CHECK_ARG(fc != nullptr, "fea_contact_forces: null fc");
  hipStream_t st = as_stream(stream);
  if (n_dof > 0) k_fill_dbl<<<blocks_for(n_dof), kBlock, 0, st>>>(n_dof, fc);
  if (n_slave == 0) {
    FEA_CHECK_LAUNCH("fea_contact_forces");
    return FEA_OK;
  }
  FEA_CHECK_ARG(slave && X && U && conn && cand && count && chosen && force && weight && sb_start && sb_items &&
                    work && (n_master == 0 || (master_nodes && nf_ptr && nf_facet && nf_vert)),
                "fea_contact_forces: null argument");
  k_contact_penalty<<<fea_contact_blocks(n_slave), kBlock, 0, st>>>(n_slave, slave, X, U, conn, cand, count, kmax,
                                                            penalty, slave_area, depth_max, chosen, force,
                                                            weight,
                                                            energy_partial);
  FEA_CHECK_LAUNCH("fea_contact_forces");
  const int rc = bucket_sort(n_slave, chosen, n_facets > 0 ? n_facets : 1, sb_start, sb_items, work, st);
  if (rc != FEA_OK) return rc;
  if (n_master > 0)
    k_contact_master<<<blocks_for(n_master), kBlock, 0, st>>>(n_master, master_nodes, nf_ptr, nf_facet, nf_vert,
                                                              sb_start, sb_items, force, weight, fc);
  k_contact_slave<<<blocks_for(n_slave), kBlock, 0, st>>>(n_slave, slave, force, fc);
  FEA_CHECK_LAUNCH("fea_contact_forces");
  return FEA_OK;
}

extern "C" int fea_contact_set_size(void) { return (int)sizeof(fea_contact_set); }

namespace fea {
int contact_check_set(const fea_contact_set *cs, const char *name) {
  FEA_CHECK_ARG(cs != nullptr, "%s: null set", name);
  const fea_contact_set &c = *cs;
  FEA_CHECK_ARG(c.n_slave > 0 && c.n_facets > 0 && c.n_cnode > 0 && c.n_master > 0 && c.kmax >= 1 &&
                    c.kmax <= kMaxCand && c.n_facets < (1 << 29) &&
                    c.table_bits >= 1 && c.table_bits <= 28 && c.radius > 0.0 && c.skin >= 0.0 &&
                    c.radius + c.skin <= c.cell,
                "%s: bad sizes (radius + skin must not exceed the cell size, kmax <= 64)", name);
  FEA_CHECK_ARG(c.slave && c.conn && c.cnode && c.cn_slave && c.cn_master && c.nf_ptr && c.nf_facet && c.nf_vert &&
                    c.centroid && c.force && c.weight && c.key && c.bucket_start && c.bucket_items && c.work &&
                    c.cand && c.count && c.chosen && c.x0 && c.dmax && c.sup_max && c.inv_start && c.inv_items &&
                    c.m_ptr && c.m_items,
                "%s: null buffer", name);
  return FEA_OK;
}
}  // namespace fea

extern "C" int fea_contact_rebuild(const fea_contact_set *cs, const double *X, const double *U, void *stream) {
  const int rc0 = contact_check_set(cs, "fea_contact_rebuild");
  if (rc0 != FEA_OK) return rc0;
  FEA_CHECK_ARG(X && U, "fea_contact_rebuild: null X or U");
  const fea_contact_set &c = *cs;
  hipStream_t st = as_stream(stream);
  const int nb = 1 << c.table_bits;
  const double inv_h = 1.0 / c.cell;
  const double rs = c.radius + c.skin;
  k_facet_keys<<<blocks_for(c.n_facets), kBlock, 0, st>>>(c.n_facets, c.conn, X, U, c.origin[0], c.origin[1],
                                                          c.origin[2], inv_h, (unsigned int)nb - 1u, c.centroid, c.key);
  int rc = bucket_sort(c.n_facets, c.key, nb, c.bucket_start, c.bucket_items, c.work, st);
  if (rc != FEA_OK) return rc;
  k_fill_int<<<1, kBlock, 0, st>>>(1, c.sup_max, 0);
  k_contact_search<<<(c.n_slave + kSearchWaves - 1) / kSearchWaves, kBlock, 0, st>>>(
      c.n_slave, c.slave, X, U, c.centroid, c.bucket_start, c.bucket_items, c.origin[0], c.origin[1], c.origin[2],
      inv_h, (unsigned int)nb - 1u, rs * rs, c.kmax, c.cand, c.count, c.sup_max);
  // facet -> slave lists of the stored pairs, slaves ascending in each
  k_fill_int<<<blocks_for(c.n_facets), kBlock, 0, st>>>(c.n_facets, c.work, 0);
  k_inv_count<<<blocks_for(c.n_slave), kBlock, 0, st>>>(c.n_slave, c.cand, c.count, c.kmax, c.work);
  k_bucket_scan<<<1, 1024, 0, st>>>(c.n_facets, c.work, c.inv_start);
  k_inv_fill<<<blocks_for(c.n_slave), kBlock, 0, st>>>(c.n_slave, c.cand, c.count, c.kmax, c.inv_start, c.work,
                                                       c.inv_items);
  k_bucket_sort<<<blocks_for(c.n_facets), kBlock, 0, st>>>(c.n_facets, c.inv_start, c.inv_items);
  // per master node lists (items in facet order, slaves ascending)
  k_mlist_count<<<blocks_for(c.n_master), kBlock, 0, st>>>(c.n_master, c.nf_ptr, c.nf_facet, c.inv_start, c.m_ptr);
  k_scan_inplace<<<1, 1024, 0, st>>>(c.n_master, c.m_ptr);
  k_mlist_fill<<<blocks_for(c.n_master), kBlock, 0, st>>>(c.n_master, c.nf_ptr, c.nf_facet, c.nf_vert, c.inv_start,
                                                          c.inv_items, c.m_ptr, c.m_items);
  k_contact_snapshot<<<blocks_for(c.n_cnode), kBlock, 0, st>>>(c.n_cnode, c.cnode, X, U, c.x0, (unsigned long long *)c.dmax);
  FEA_CHECK_LAUNCH("fea_contact_rebuild");
  return FEA_OK;
}

extern "C" int fea_contact_track(const fea_contact_set *cs, const double *X, const double *U, void *stream) {
  const int rc0 = contact_check_set(cs, "fea_contact_track");
  if (rc0 != FEA_OK) return rc0;
  FEA_CHECK_ARG(X && U, "fea_contact_track: null X or U");
  const fea_contact_set &c = *cs;
  k_contact_track<<<blocks_for(c.n_cnode), kBlock, 0, as_stream(stream)>>>(c.n_cnode, c.cnode, X, U, c.x0,
                                                                            (unsigned long long *)c.dmax);
  FEA_CHECK_LAUNCH("fea_contact_track");
  return FEA_OK;
}

extern "C" int fea_contact_eval(const fea_contact_set *cs, const double *X, const double *U, double *fc,
                                double *energy_partial, void *stream) {
  const int rc0 = contact_check_set(cs, "fea_contact_eval");
  if (rc0 != FEA_OK) return rc0;
  FEA_CHECK_ARG(X && U && fc, "fea_contact_eval: null X, U or fc");
  const fea_contact_set &c = *cs;
  hipStream_t st = as_stream(stream);
  const ContactDev d = contact_dev(c);
  k_contact_track<<<blocks_for(c.n_cnode), kBlock, 0, st>>>(c.n_cnode, c.cnode, X, U, c.x0, d.dmax);
  k_contact_penalty_v<<<fea_contact_blocks(c.n_slave), kBlock, 0, st>>>(d, X, U, energy_partial);
  k_contact_reactions_v<<<(c.n_cnode + kReactWaves - 1) / kReactWaves, kBlock, 0, st>>>(d, fc);
  FEA_CHECK_LAUNCH("fea_contact_eval");
  return FEA_OK;
}
