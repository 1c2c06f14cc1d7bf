// Per-step device code of the node-to-surface contact with reused candidate lists: the
// penalty of one slave, the force of one contact node and the displacement tracking,
// shared by the contact kernels and by the explicit loop's fused launches (element pass +
// penalty; node update + contact-node forces + tracking), which give the same bits.
// Arithmetic outside the explicit no-fma predicate fuses a * b + c only inside one
// expression (`#pragma clang fp contract(on)`), so every kernel inlining it rounds alike.
#pragma once

#include <climits>

#include "common.h"

namespace fea {

__device__ __forceinline__ double dadd(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double dmul(double a, double b) { return __dmul_rn(a, b); }

__device__ __forceinline__ void node_pos(const double *X, const double *U, int64_t n, double p[3]) {
  p[0] = dadd(X[3 * n + 0], U[3 * n + 0]);
  p[1] = dadd(X[3 * n + 1], U[3 * n + 1]);
  p[2] = dadd(X[3 * n + 2], U[3 * n + 2]);
}

// ---------------------------------------------------------------- penalty
// Facet choice per slave, over the candidates whose plane it penetrates
// (-depth_max < g < 0), with v = the smallest barycentric coordinate of the
// slave's projection on the facet plane:
//   tier 0  v >= -kInsideTol (inside; edges and vertices count, whatever the
//           rounding, since conforming meshes put slaves right over master
//           nodes): the largest g wins (least penetration);
//   tier 1  -kExtendTol <= v < -kInsideTol, used only when no facet is in
//           tier 0: the largest v wins.  This closes the dead zone under a
//           concave crease (a valley, such as the dent a slave presses into the
//           master surface), where a penetrating slave's projections on both
//           adjacent planes fall just outside their facets and the slave would
//           otherwise drop out of contact and re-enter deeper (an energy jump).
// Ties keep the smaller facet id.  Force F = -k A_s g n on the slave (A_s its
// tributary area, 1 when area is null).
constexpr double kInsideTol = 1e-10;
constexpr double kExtendTol = 0.1;

// The penalty terms of slave position p against facet f: false when the slave does not
// penetrate the facet's plane within depth_max or projects outside the extended facet.
struct PenCand {
  int tier;
  double key, g, n[3], w[3];
};

__device__ __forceinline__ bool pen_eval(const double p[3], const double *__restrict__ X,
                                         const double *__restrict__ U, const int32_t *__restrict__ conn, int f,
                                         double depth_max, PenCand &out) {
  #pragma clang fp contract(on)
  double a[3], b[3], c[3];
  node_pos(X, U, conn[3 * (int64_t)f + 0], a);
  node_pos(X, U, conn[3 * (int64_t)f + 1], b);
  node_pos(X, U, conn[3 * (int64_t)f + 2], c);
  double e1[3], e2[3], ap[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    e1[j] = b[j] - a[j];
    e2[j] = c[j] - a[j];
    ap[j] = p[j] - a[j];
  }
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (!(nn > 0.0)) return false;
  n[0] /= nn;
  n[1] /= nn;
  n[2] /= nn;
  const double g = ap[0] * n[0] + ap[1] * n[1] + ap[2] * n[2];
  if (!(g < 0.0) || !(g > -depth_max)) return false;
  double q[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) q[j] = ap[j] - g * n[j];
  const double d00 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  const double d01 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  const double d11 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  const double d20 = q[0] * e1[0] + q[1] * e1[1] + q[2] * e1[2];
  const double d21 = q[0] * e2[0] + q[1] * e2[1] + q[2] * e2[2];
  const double den = d00 * d11 - d01 * d01;
  const double xi = (d11 * d20 - d01 * d21) / den;
  const double eta = (d00 * d21 - d01 * d20) / den;
  const double w0 = 1.0 - xi - eta;
  const double v = fmin(w0, fmin(xi, eta));
  if (!(v >= -kExtendTol)) return false;
  out.tier = (v >= -kInsideTol) ? 0 : 1;
  out.key = (out.tier == 0) ? g : v;
  out.g = g;
  out.n[0] = n[0];
  out.n[1] = n[1];
  out.n[2] = n[2];
  out.w[0] = w0;
  out.w[1] = xi;
  out.w[2] = eta;
  return true;
}

// The wave's best lane by (tier ascending, key descending, index ascending); lane 0 writes
// the slave's choice, force and weights and returns its penalty energy (0 on other lanes).
__device__ __forceinline__ double pen_finish(int s, int lane, int bf, int bi, int btier, double bkey, double bg,
                                             double bn[3], double bw[3], double k_pen,
                                             const double *__restrict__ area, int32_t *__restrict__ chosen,
                                             double *__restrict__ force, double *__restrict__ weight) {
  #pragma clang fp contract(on)
  int wl = lane;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int ot = __shfl_xor(btier, off, 64);
    const double ok = __shfl_xor(bkey, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    const int ol = __shfl_xor(wl, off, 64);
    if (ot < btier || (ot == btier && (ok > bkey || (ok == bkey && oi < bi)))) {
      btier = ot;
      bkey = ok;
      bi = oi;
      wl = ol;
    }
  }
  bf = __shfl(bf, wl, 64);
  bg = __shfl(bg, wl, 64);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    bn[j] = __shfl(bn[j], wl, 64);
    bw[j] = __shfl(bw[j], wl, 64);
  }
  double e = 0.0;
  if (lane == 0) {
    const int best = btier < 2 ? bf : -1;
    chosen[s] = best;
    const double ks = (area != nullptr) ? k_pen * area[s] : k_pen;
    const double fm = (best >= 0) ? -ks * bg : 0.0;
    force[3 * (int64_t)s + 0] = fm * bn[0];
    force[3 * (int64_t)s + 1] = fm * bn[1];
    force[3 * (int64_t)s + 2] = fm * bn[2];
    weight[3 * (int64_t)s + 0] = best >= 0 ? bw[0] : 0.0;
    weight[3 * (int64_t)s + 1] = best >= 0 ? bw[1] : 0.0;
    weight[3 * (int64_t)s + 2] = best >= 0 ? bw[2] : 0.0;
    if (best >= 0) e = 0.5 * ks * bg * bg;
  }
  return e;
}

// ---------------------------------------------------------------- reused candidate lists
// Between rebuilds each slave keeps the candidates of a search with the enlarged radius
// r + skin, and facet -> slave lists of those pairs.  While no contact node has moved more
// than skin/2 since the rebuild, every pair within r now was within r + skin then, so the
// stored lists hold every exact candidate; each step tests the exact predicate on them.
// Otherwise (the host rebuilds late) the kernels search every facet / slave instead.
// dmax holds the largest squared displacement of a contact node since the rebuild, as the
// bits of a non-negative double (ordered like the values).

__device__ __forceinline__ void facet_centroid(const double *__restrict__ X, const double *__restrict__ U,
                                               const int32_t *__restrict__ conn, int f, double m[3]) {
  double a[3], b[3], c[3];
  node_pos(X, U, conn[3 * (int64_t)f + 0], a);
  node_pos(X, U, conn[3 * (int64_t)f + 1], b);
  node_pos(X, U, conn[3 * (int64_t)f + 2], c);
#pragma unroll
  for (int i = 0; i < 3; ++i) m[i] = dadd(dadd(a[i], b[i]), c[i]) / 3.0;
}

// the exact candidate predicate of k_contact_search at the current centroid
__device__ __forceinline__ bool exact_candidate(const double p[3], const double m[3], double r2) {
  const double ddx = p[0] - m[0], ddy = p[1] - m[1], ddz = p[2] - m[2];
  return dadd(dadd(dmul(ddx, ddx), dmul(ddy, ddy)), dmul(ddz, ddz)) <= r2;
}

constexpr int kDmaxSlots = FEA_CONTACT_DMAX_SLOTS, kDmaxStride = FEA_CONTACT_DMAX_STRIDE;
static_assert(kDmaxSlots == 64, "one dmax slot per lane");

// raise slot (k % 64) of dmax to the bits of d2 (>= 0; doubles of one sign order like their bits)
__device__ __forceinline__ void dmax_raise(unsigned long long *dmax, int k, double d2) {
  if (d2 > 0.0)
    atomicMax(dmax + (int64_t)(k & (kDmaxSlots - 1)) * kDmaxStride, (unsigned long long)__double_as_longlong(d2));
}

// the tracked squared displacement: the largest slot (one wave, every lane gets it)
__device__ __forceinline__ double dmax_wave(const unsigned long long *dmax, int lane) {
  unsigned long long v = dmax[(int64_t)lane * kDmaxStride];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return __longlong_as_double((long long)v);
}

// whether the stored lists hold every exact candidate (called by a whole wave)
__device__ __forceinline__ bool lists_valid(const unsigned long long *dmax, double half_skin2, const int32_t *sup_max,
                                            int kmax, int lane) {
  return dmax_wave(dmax, lane) <= half_skin2 && *sup_max <= kmax;
}

// Contact node forces, one wave per contact node.  Its master items (slave s, facet f,
// local vertex v), in facet order and ascending slaves within a facet, come from the stored
// per-master-node lists (valid lists: every slave that chose one of the node's facets is
// there) or, otherwise, from a scan of every slave per facet.  Lanes test items chosen[s] == f
// in chunks of 64, a ballot marks the matches, and the wave subtracts F_s w_s(v) in item
// order: the sums of the bucket sort of slaves by chosen facet.  Then the node's own slave
// force is added; fc[node] is written once.

__device__ __forceinline__ void react_chunk(bool match, int s, int v, const double *__restrict__ force,
                                            const double *__restrict__ weight, double acc[3]) {
  #pragma clang fp contract(on)
  unsigned long long mask = __ballot(match);
  double f0 = 0.0, f1 = 0.0, f2 = 0.0, w = 0.0;
  if (match) {
    f0 = force[3 * (int64_t)s + 0];
    f1 = force[3 * (int64_t)s + 1];
    f2 = force[3 * (int64_t)s + 2];
    w = weight[3 * (int64_t)s + v];
  }
  while (mask) {
    const int b = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const double wb = __shfl(w, b, 64);
    acc[0] -= __shfl(f0, b, 64) * wb;
    acc[1] -= __shfl(f1, b, 64) * wb;
    acc[2] -= __shfl(f2, b, 64) * wb;
  }
}

// The contact set's fields that the per-step kernels read (passed by value).
struct ContactDev {
  int n_slave, n_facets, n_cnode, kmax;
  double r2, half_skin2, penalty, depth_max;
  const int32_t *slave, *conn, *cnode, *cn_slave, *cn_master, *nf_ptr, *nf_facet, *nf_vert;
  const int32_t *m_ptr, *m_items, *cand, *count, *sup_max;
  const double *area, *x0;
  unsigned long long *dmax;
  int32_t *max_count, *chosen, *lists_ok;
  double *force, *weight;
};

inline ContactDev contact_dev(const fea_contact_set &c) {
  ContactDev d;
  d.n_slave = c.n_slave;
  d.n_facets = c.n_facets;
  d.n_cnode = c.n_cnode;
  d.kmax = c.kmax;
  d.r2 = c.radius * c.radius;
  d.half_skin2 = 0.25 * c.skin * c.skin;
  d.penalty = c.penalty;
  d.depth_max = c.depth_max;
  d.slave = c.slave;
  d.conn = c.conn;
  d.cnode = c.cnode;
  d.cn_slave = c.cn_slave;
  d.cn_master = c.cn_master;
  d.nf_ptr = c.nf_ptr;
  d.nf_facet = c.nf_facet;
  d.nf_vert = c.nf_vert;
  d.m_ptr = c.m_ptr;
  d.m_items = c.m_items;
  d.cand = c.cand;
  d.count = c.count;
  d.sup_max = c.sup_max;
  d.area = c.slave_area;
  d.x0 = c.x0;
  d.dmax = reinterpret_cast<unsigned long long *>(c.dmax);
  d.max_count = c.max_count;
  d.chosen = c.chosen;
  d.lists_ok = c.lists_ok;
  d.force = c.force;
  d.weight = c.weight;
  return d;
}

// Validates a fea_contact_set (sizes and buffers); FEA_OK or FEA_ERR_ARG with the message set.
int contact_check_set(const fea_contact_set *cs, const char *name);

// Slave s (one wave; lane = lane id): candidates from its stored list (valid) or every
// facet, the exact predicate, then the penalty choice with ties broken by facet id (the
// order of the exact search's sorted list).  Writes chosen / force / weight of s, raises
// max_count by exact candidate counts above kmax; returns the penalty energy on lane 0.
__device__ __forceinline__ double penalty_slave(const ContactDev &c, const double *__restrict__ X,
                                                const double *__restrict__ U, int s, int lane, bool valid) {
  double p[3];
  node_pos(X, U, c.slave[s], p);
  int bf = -1, bi = INT_MAX, btier = 2, hits = 0;
  double bg = 0.0, bkey = 0.0, bn[3] = {0, 0, 0}, bw[3] = {0, 0, 0};
  const int n = valid ? min(c.count[s], c.kmax) : c.n_facets;
  for (int i = lane; i < n; i += 64) {
    const int f = valid ? c.cand[(int64_t)s * c.kmax + i] : i;
    double m[3];
    facet_centroid(X, U, c.conn, f, m);
    if (!exact_candidate(p, m, c.r2)) continue;
    ++hits;
    PenCand pc;
    if (!pen_eval(p, X, U, c.conn, f, c.depth_max, pc)) continue;
    if (pc.tier < btier || (pc.tier == btier && pc.key > bkey)) {
      bf = f;
      bi = f;
      btier = pc.tier;
      bkey = pc.key;
      bg = pc.g;
      bn[0] = pc.n[0];
      bn[1] = pc.n[1];
      bn[2] = pc.n[2];
      bw[0] = pc.w[0];
      bw[1] = pc.w[1];
      bw[2] = pc.w[2];
    }
  }
  for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off, 64);
  if (lane == 0 && c.max_count != nullptr && hits > c.kmax) atomicMax(c.max_count, hits);
  return pen_finish(s, lane, bf, bi, btier, bkey, bg, bn, bw, c.penalty, c.area, c.chosen, c.force, c.weight);
}

// The force on contact node i (one wave): its master reactions over the stored
// per-master-node lists (valid) or a scan of every slave per facet, then its own slave
// force.  The result is valid on lane 0.
__device__ __forceinline__ void reaction_node(const ContactDev &c, int i, int lane, bool valid, double r[3]) {
  #pragma clang fp contract(on)
  r[0] = r[1] = r[2] = 0.0;
  const int m = c.cn_master[i];
  if (m >= 0) {
    double acc[3] = {0.0, 0.0, 0.0};
    if (valid) {
      const int k0 = c.m_ptr[m], k1 = c.m_ptr[m + 1];
      for (int b = k0; b < k1; b += 64) {
        const int k = b + lane;
        int s = 0, fv = -1;
        if (k < k1) {
          s = c.m_items[2 * (int64_t)k];
          fv = c.m_items[2 * (int64_t)k + 1];
        }
        const bool match = fv >= 0 && c.chosen[s] == (fv >> 2);
        react_chunk(match, s, fv & 3, c.force, c.weight, acc);
      }
    } else {
      for (int t = c.nf_ptr[m]; t < c.nf_ptr[m + 1]; ++t) {
        const int f = c.nf_facet[t], v = c.nf_vert[t];
        for (int b = 0; b < c.n_slave; b += 64) {
          const int s = b + lane;
          const bool match = s < c.n_slave && c.chosen[s] == f;
          react_chunk(match, s, v, c.force, c.weight, acc);
        }
      }
    }
    r[0] += acc[0];
    r[1] += acc[1];
    r[2] += acc[2];
  }
  const int sl = c.cn_slave[i];
  if (sl >= 0) {
    r[0] += c.force[3 * (int64_t)sl + 0];
    r[1] += c.force[3 * (int64_t)sl + 1];
    r[2] += c.force[3 * (int64_t)sl + 2];
  }
}

// |p - x0_i|^2 without fused multiply-add (the tracked displacement of contact node i)
__device__ __forceinline__ double track_d2(const double p[3], const double *__restrict__ x0, int i) {
  const double dx = p[0] - x0[3 * (int64_t)i], dy = p[1] - x0[3 * (int64_t)i + 1], dz = p[2] - x0[3 * (int64_t)i + 2];
  return dadd(dadd(dmul(dx, dx), dmul(dy, dy)), dmul(dz, dz));
}

}  // namespace fea
