// Host-side sparsity pattern: free-DOF CSR + deterministic staging maps.
//
// Replaces the reference's per-group COO pattern with argsort/unique
// coalescing (C3base.py:109-153), the free-DOF masking and per-axis unique()
// compaction (assembly.py:371-380) and the COO->CSR coalesce of
// _linear_solver.py:18-19.  The CSR produced here is the one those steps
// produce (rows = free DOFs in GC order, columns ascending, duplicates
// merged); on top of it every element/facet node-pair block gets a fixed,
// contiguous slot range in a staging array so the device gather sums
// contributions in a fixed order without atomics.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/fea_amd.h"

namespace fea {
void set_error(const char *fmt, ...);
}

struct fea_pattern {
  int n_nodes = 0;
  int64_t n_free = 0, nnz = 0, n_blocks = 0, n_kstage = 0, n_rstage = 0;
  std::vector<int32_t> rowptr, col, diag_slot;
  std::vector<int32_t> blk_a, blk_b, blk_colpos, kseg, rseg;
  std::vector<std::vector<int32_t>> dst_k, dst_r;
};

namespace {

struct GroupView {
  const int32_t *conn;
  int n_elem, nen, in_k;
};

int build(fea_pattern &P, int n_nodes, const int32_t *dof_map, const std::vector<GroupView> &G) {
  P.n_nodes = n_nodes;
  // ---- free DOF bookkeeping: GC order requires dof_map increasing over free DOFs
  std::vector<int8_t> nfree(n_nodes, 0);
  int64_t expect = 0;
  for (int64_t i = 0; i < 3LL * n_nodes; ++i) {
    const int32_t d = dof_map[i];
    if (d >= 0) {
      if (d != expect) {
        fea::set_error("fea_pattern_create: dof_map must number free DOFs 0..n-1 in node-major order");
        return FEA_ERR_ARG;
      }
      ++expect;
      ++nfree[i / 3];
    } else if (d != -1) {
      fea::set_error("fea_pattern_create: dof_map entries must be >= -1");
      return FEA_ERR_ARG;
    }
  }
  P.n_free = expect;

  // ---- node -> (group, element) incidence over in_k groups
  std::vector<int64_t> inc_ptr(n_nodes + 1, 0);
  for (const auto &g : G) {
    if (!g.in_k) continue;
    for (int64_t t = 0; t < (int64_t)g.n_elem * g.nen; ++t) {
      const int32_t n = g.conn[t];
      if (n < 0 || n >= n_nodes) {
        fea::set_error("fea_pattern_create: node id %d out of range", (int)n);
        return FEA_ERR_ARG;
      }
      ++inc_ptr[n + 1];
    }
  }
  for (int i = 0; i < n_nodes; ++i) inc_ptr[i + 1] += inc_ptr[i];
  std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
  std::vector<std::pair<int32_t, int32_t>> inc(inc_ptr[n_nodes]);  // (group, element)
  for (int gi = 0; gi < (int)G.size(); ++gi) {
    const auto &g = G[gi];
    if (!g.in_k) continue;
    for (int e = 0; e < g.n_elem; ++e)
      for (int a = 0; a < g.nen; ++a) inc[fill[g.conn[(int64_t)e * g.nen + a]]++] = {gi, e};
  }

  // ---- node adjacency (sorted, unique)
  std::vector<int64_t> adj_ptr(n_nodes + 1, 0);
  std::vector<int32_t> adj;
  adj.reserve((size_t)inc_ptr[n_nodes] * 4);
  std::vector<int32_t> tmp;
  std::vector<int32_t> stamp(n_nodes, -1);  // stamp[B] == A: B already listed for A
  for (int A = 0; A < n_nodes; ++A) {
    tmp.clear();
    for (int64_t s = inc_ptr[A]; s < inc_ptr[A + 1]; ++s) {
      const auto &g = G[inc[s].first];
      const int32_t *c = g.conn + (int64_t)inc[s].second * g.nen;
      for (int b = 0; b < g.nen; ++b)
        if (stamp[c[b]] != A) {
          stamp[c[b]] = A;
          tmp.push_back(c[b]);
        }
    }
    std::sort(tmp.begin(), tmp.end());
    adj.insert(adj.end(), tmp.begin(), tmp.end());
    adj_ptr[A + 1] = (int64_t)adj.size();
  }

  // ---- blocks (A,B) with free rows and free columns, and the CSR
  std::vector<int64_t> blk_start(n_nodes + 1, 0);  // first block of node A (-1 if none)
  P.rowptr.assign(P.n_free + 1, 0);
  int64_t nb = 0, nnz = 0;
  for (int A = 0; A < n_nodes; ++A) {
    blk_start[A] = nb;
    if (nfree[A] == 0) continue;
    int64_t rowlen = 0;
    for (int64_t s = adj_ptr[A]; s < adj_ptr[A + 1]; ++s) {
      const int B = adj[s];
      if (nfree[B] == 0) continue;
      P.blk_a.push_back(A);
      P.blk_b.push_back(B);
      P.blk_colpos.push_back((int32_t)rowlen);
      rowlen += nfree[B];
      ++nb;
    }
    for (int j = 0; j < 3; ++j) {
      const int32_t r = dof_map[3LL * A + j];
      if (r < 0) continue;
      P.rowptr[r + 1] = (int32_t)rowlen;
      nnz += rowlen;
    }
  }
  blk_start[n_nodes] = nb;
  if (nnz > INT32_MAX) {
    fea::set_error("fea_pattern_create: nnz %lld exceeds int32", (long long)nnz);
    return FEA_ERR_UNSUPPORTED;
  }
  for (int64_t r = 0; r < P.n_free; ++r) P.rowptr[r + 1] += P.rowptr[r];
  P.n_blocks = nb;
  P.nnz = nnz;
  P.col.resize(nnz);
  P.diag_slot.assign(P.n_free, -1);
  for (int A = 0; A < n_nodes; ++A) {
    if (nfree[A] == 0) continue;
    for (int j = 0; j < 3; ++j) {
      const int32_t r = dof_map[3LL * A + j];
      if (r < 0) continue;
      int64_t p = P.rowptr[r];
      for (int64_t s = adj_ptr[A]; s < adj_ptr[A + 1]; ++s) {
        const int B = adj[s];
        for (int k = 0; k < 3; ++k) {
          const int32_t c = dof_map[3LL * B + k];
          if (c < 0) continue;
          if (c == r) P.diag_slot[r] = (int32_t)p;
          P.col[p++] = c;
        }
      }
    }
  }

  // ---- K staging: count contributions per block, then place in (group, e, a, b) order.
  // The block of (c[a], c[b]) is looked up node by node: while node A is current, bid[B]
  // holds the block (A, B) for every free neighbour B (bstamp[B] == A marks it valid).
  std::vector<int64_t> cnt(nb + 1, 0);
  P.dst_k.assign(G.size(), {});
  P.dst_r.assign(G.size(), {});
  for (int gi = 0; gi < (int)G.size(); ++gi)
    if (G[gi].in_k) P.dst_k[gi].assign((size_t)G[gi].n_elem * G[gi].nen * G[gi].nen, -1);
  {
    std::vector<int32_t> bstamp(n_nodes, -1);
    std::vector<int64_t> bid(n_nodes, -1);
    for (int A = 0; A < n_nodes; ++A) {
      if (nfree[A] == 0) continue;  // rows of a fixed node stay -1
      for (int64_t q = blk_start[A]; q < blk_start[A + 1]; ++q) {
        bstamp[P.blk_b[q]] = A;
        bid[P.blk_b[q]] = q;
      }
      for (int64_t s = inc_ptr[A]; s < inc_ptr[A + 1]; ++s) {
        const int gi = inc[s].first;
        const auto &g = G[gi];
        const int e = inc[s].second;
        const int32_t *c = g.conn + (int64_t)e * g.nen;
        auto &dk = P.dst_k[gi];
        for (int a = 0; a < g.nen; ++a) {
          if (c[a] != A) continue;
          for (int b = 0; b < g.nen; ++b) {
            const int32_t B = c[b];
            const int64_t q = bstamp[B] == A ? bid[B] : -1;
            dk[((size_t)e * g.nen + a) * g.nen + b] = (int32_t)q;   // temporarily the block id
            if (q >= 0) ++cnt[q + 1];
          }
        }
      }
    }
  }
  for (int64_t q = 0; q < nb; ++q) cnt[q + 1] += cnt[q];
  if (cnt[nb] > INT32_MAX) {
    fea::set_error("fea_pattern_create: staging size exceeds int32");
    return FEA_ERR_UNSUPPORTED;
  }
  P.kseg.assign(cnt.begin(), cnt.end());
  P.n_kstage = cnt[nb];
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int gi = 0; gi < (int)G.size(); ++gi) {
      if (!G[gi].in_k) continue;
      for (auto &d : P.dst_k[gi])
        if (d >= 0) d = (int32_t)(pos[d]++);
    }
  }

  // ---- R staging: per node, contributions of every group in (group, e, a) order
  std::vector<int64_t> rc(n_nodes + 1, 0);
  for (const auto &g : G)
    for (int64_t t = 0; t < (int64_t)g.n_elem * g.nen; ++t) {
      const int32_t n = g.conn[t];
      if (n < 0 || n >= n_nodes) {
        fea::set_error("fea_pattern_create: node id %d out of range", (int)n);
        return FEA_ERR_ARG;
      }
      if (nfree[n] > 0) ++rc[n + 1];
    }
  for (int i = 0; i < n_nodes; ++i) rc[i + 1] += rc[i];
  P.rseg.assign(rc.begin(), rc.end());
  P.n_rstage = rc[n_nodes];
  {
    std::vector<int64_t> pos(rc.begin(), rc.end() - 1);
    for (int gi = 0; gi < (int)G.size(); ++gi) {
      const auto &g = G[gi];
      auto &dr = P.dst_r[gi];
      dr.assign((size_t)g.n_elem * g.nen, -1);
      for (int64_t t = 0; t < (int64_t)g.n_elem * g.nen; ++t) {
        const int32_t n = g.conn[t];
        if (nfree[n] > 0) dr[t] = (int32_t)(pos[n]++);
      }
    }
  }
  return FEA_OK;
}

}  // namespace

extern "C" int fea_pattern_create(int n_nodes, const int32_t *dof_map_host, int n_groups,
                                  const int32_t *const *conn_host, const int32_t *n_elem_host,
                                  const int32_t *nen_host, const int32_t *in_k_host,
                                  fea_pattern **out) {
  if (out == nullptr || n_nodes < 0 || n_groups < 0 || (n_nodes > 0 && dof_map_host == nullptr)) {
    fea::set_error("fea_pattern_create: bad arguments");
    return FEA_ERR_ARG;
  }
  *out = nullptr;
  std::vector<GroupView> G;
  for (int i = 0; i < n_groups; ++i) {
    if (n_elem_host[i] < 0 || nen_host[i] <= 0 || (n_elem_host[i] > 0 && conn_host[i] == nullptr)) {
      fea::set_error("fea_pattern_create: bad group %d", i);
      return FEA_ERR_ARG;
    }
    G.push_back({conn_host[i], n_elem_host[i], nen_host[i], in_k_host[i]});
  }
  fea_pattern *P = new (std::nothrow) fea_pattern();
  if (P == nullptr) {
    fea::set_error("fea_pattern_create: out of host memory");
    return FEA_ERR_STATE;
  }
  int rc;
  try {
    rc = build(*P, n_nodes, dof_map_host, G);
  } catch (const std::bad_alloc &) {
    fea::set_error("fea_pattern_create: out of host memory");
    rc = FEA_ERR_STATE;
  }
  if (rc != FEA_OK) {
    delete P;
    return rc;
  }
  *out = P;
  return FEA_OK;
}

extern "C" int fea_pattern_sizes(const fea_pattern *p, int64_t sizes[5]) {
  if (p == nullptr || sizes == nullptr) {
    fea::set_error("fea_pattern_sizes: null argument");
    return FEA_ERR_ARG;
  }
  sizes[0] = p->n_free;
  sizes[1] = p->nnz;
  sizes[2] = p->n_blocks;
  sizes[3] = p->n_kstage;
  sizes[4] = p->n_rstage;
  return FEA_OK;
}

template <typename V>
static void copy_out(const V &v, int32_t *dst) {
  if (dst != nullptr && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int32_t));
}

extern "C" int fea_pattern_export_csr(const fea_pattern *p, int32_t *rowptr_host, int32_t *col_host,
                                      int32_t *diag_slot_host) {
  if (p == nullptr) {
    fea::set_error("fea_pattern_export_csr: null pattern");
    return FEA_ERR_ARG;
  }
  copy_out(p->rowptr, rowptr_host);
  copy_out(p->col, col_host);
  copy_out(p->diag_slot, diag_slot_host);
  return FEA_OK;
}

extern "C" int fea_pattern_export_blocks(const fea_pattern *p, int32_t *blk_a_host,
                                         int32_t *blk_b_host, int32_t *blk_colpos_host,
                                         int32_t *kseg_host, int32_t *rseg_host) {
  if (p == nullptr) {
    fea::set_error("fea_pattern_export_blocks: null pattern");
    return FEA_ERR_ARG;
  }
  copy_out(p->blk_a, blk_a_host);
  copy_out(p->blk_b, blk_b_host);
  copy_out(p->blk_colpos, blk_colpos_host);
  copy_out(p->kseg, kseg_host);
  copy_out(p->rseg, rseg_host);
  return FEA_OK;
}

extern "C" int fea_pattern_export_group(const fea_pattern *p, int group, int32_t *dst_k_host,
                                        int32_t *dst_r_host) {
  if (p == nullptr || group < 0 || group >= (int)p->dst_r.size()) {
    fea::set_error("fea_pattern_export_group: bad group");
    return FEA_ERR_ARG;
  }
  copy_out(p->dst_k[group], dst_k_host);
  copy_out(p->dst_r[group], dst_r_host);
  return FEA_OK;
}

// Column position of every (element, a, b) block of a group in its CSR rows: the block's
// blk_colpos, found from the staging slot through kseg; -1 where the staging map is -1.
extern "C" int fea_pattern_export_group_colpos(const fea_pattern *p, int group, int32_t *colpos_host) {
  if (p == nullptr || group < 0 || group >= (int)p->dst_k.size() || colpos_host == nullptr) {
    fea::set_error("fea_pattern_export_group_colpos: bad argument");
    return FEA_ERR_ARG;
  }
  const auto &dk = p->dst_k[group];
  const auto &ks = p->kseg;
  for (size_t t = 0; t < dk.size(); ++t) {
    const int32_t d = dk[t];
    if (d < 0) {
      colpos_host[t] = -1;
      continue;
    }
    // the block q with kseg[q] <= d < kseg[q + 1]
    const int64_t q = (int64_t)(std::upper_bound(ks.begin(), ks.end(), d) - ks.begin()) - 1;
    colpos_host[t] = p->blk_colpos[q];
  }
  return FEA_OK;
}

extern "C" void fea_pattern_destroy(fea_pattern *p) { delete p; }

// Greedy colouring of elements: no two elements of one colour share a node; elements
// are taken in order and get the smallest colour none of their already coloured
// neighbours holds (a structured brick grid in natural order gets 8).
extern "C" int fea_color_elements(int n_nodes, int n_elem, int nen, const int32_t *conn, int32_t *color,
                                  int32_t *n_colors) {
  if (n_nodes < 0 || n_elem < 0 || nen <= 0 || (n_elem > 0 && (conn == nullptr || color == nullptr)) ||
      n_colors == nullptr) {
    fea::set_error("fea_color_elements: bad arguments");
    return FEA_ERR_ARG;
  }
  try {
    std::vector<int64_t> ptr(n_nodes + 1, 0);
    for (int64_t t = 0; t < (int64_t)n_elem * nen; ++t) {
      const int32_t n = conn[t];
      if (n < 0 || n >= n_nodes) {
        fea::set_error("fea_color_elements: node id %d out of range", (int)n);
        return FEA_ERR_ARG;
      }
      ++ptr[n + 1];
    }
    for (int i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
    std::vector<int32_t> inc(ptr[n_nodes]);
    {
      std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
      for (int e = 0; e < n_elem; ++e)
        for (int a = 0; a < nen; ++a) inc[fill[conn[(int64_t)e * nen + a]]++] = e;
    }
    std::vector<int32_t> stamp;  // stamp[c] == e: colour c is taken by a neighbour of e
    int32_t nc = 0;
    for (int e = 0; e < n_elem; ++e) color[e] = -1;
    for (int e = 0; e < n_elem; ++e) {
      for (int a = 0; a < nen; ++a) {
        const int32_t n = conn[(int64_t)e * nen + a];
        for (int64_t s = ptr[n]; s < ptr[n + 1]; ++s) {
          const int32_t c = color[inc[s]];
          if (c >= 0) stamp[c] = e;
        }
      }
      int32_t c = 0;
      while (c < nc && stamp[c] == e) ++c;
      if (c == nc) {
        stamp.push_back(-1);
        ++nc;
      }
      color[e] = c;
    }
    *n_colors = nc;
  } catch (const std::bad_alloc &) {
    fea::set_error("fea_color_elements: out of host memory");
    return FEA_ERR_STATE;
  }
  return FEA_OK;
}
