// PyTorch-ROCm operators over the fea_amd C ABI: TORCH_LIBRARY(fea, ...).
//
// Each operator checks its tensors (device, dtype, contiguity, sizes), launches the
// C-ABI entry point on the current HIP stream of the tensors' device, and turns a
// non-zero status into a Python RuntimeError carrying fea_last_error().  The
// operators are the ones the drop-in classes call on their hot path: the element
// residual / tangent of Element_3D.structural_Force (C3base.py:413-452) and its
// strain energy (C3base.py:387-411), the assembly gathers that replace the COO
// scatter-add of Assembly._assemble_generalized_Matrix (assembly.py:311-387), the
// SpMV, and the device PCG of conjugate_gradient (_linear_solver.py:4-42).
#include <ATen/ATen.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <string>

#include "../../include/fea_amd.h"

namespace {

using at::Tensor;

void *stream_of(const Tensor &t) {
  return reinterpret_cast<void *>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

void check_status(int rc, const char *what) {
  TORCH_CHECK(rc == 0, "fea::", what, " failed (status ", rc, "): ", fea_last_error());
}

// a device tensor of the given dtype, contiguous, on the reference tensor's device
void check_dev(const Tensor &t, at::ScalarType dt, const Tensor &ref, const char *name) {
  TORCH_CHECK(t.defined(), name, ": tensor required");
  TORCH_CHECK(t.is_cuda(), name, ": must be a ROCm device tensor");
  TORCH_CHECK(t.device() == ref.device(), name, ": must be on ", ref.device(), ", not ", t.device());
  TORCH_CHECK(t.scalar_type() == dt, name, ": expected dtype ", dt, ", got ", t.scalar_type());
  TORCH_CHECK(t.is_contiguous(), name, ": must be contiguous");
}

at::ScalarType float_type(int64_t prec) {
  TORCH_CHECK(prec == FEA_F32 || prec == FEA_F64, "precision enum must be 0 (float32) or 1 (float64)");
  return prec == FEA_F64 ? at::kDouble : at::kFloat;
}

int nen_of(int64_t elem_type) {
  TORCH_CHECK(elem_type == 4 || elem_type == 8 || elem_type == 10, "elem_type must be 4, 8 or 10 (C3D4/C3D8/C3D10)");
  return (int)elem_type;
}

int64_t elems_of(const Tensor &conn, int nen) {
  TORCH_CHECK(conn.numel() % nen == 0, "conn: ", conn.numel(), " entries is not a multiple of ", nen);
  return conn.numel() / nen;
}

void check_params(c10::ArrayRef<double> params) {
  TORCH_CHECK(params.size() == 3, "params: 3 material constants expected (w1, w2, kv)");
}

void check_nodes(const Tensor &X, const Tensor &U) {
  TORCH_CHECK(X.numel() % 3 == 0 && U.numel() == X.numel(), "X and U: [3 n_nodes] each");
}

void elem_residual_tangent(int64_t elem_type, int64_t dtype, const Tensor &conn, const Tensor &X, const Tensor &U,
                           c10::ArrayRef<double> params, const Tensor &dst_k, const Tensor &kstage,
                           const c10::optional<Tensor> &dst_r, const c10::optional<Tensor> &rstage) {
  const int nen = nen_of(elem_type);
  check_dev(conn, at::kInt, X, "conn");
  check_dev(X, at::kDouble, X, "X");
  check_dev(U, at::kDouble, X, "U");
  check_dev(dst_k, at::kInt, X, "dst_k");
  check_dev(kstage, dtype == FEA_F64_STAGE_F32 ? at::kFloat : float_type(dtype), X, "kstage");
  check_nodes(X, U);
  check_params(params);
  const int64_t n = elems_of(conn, nen);
  TORCH_CHECK(dst_k.numel() == n * nen * nen, "dst_k: one slot per element node pair");
  const bool with_r = dst_r.has_value() && dst_r->defined();
  TORCH_CHECK(with_r == (rstage.has_value() && rstage->defined()), "dst_r and rstage: both or neither");
  if (with_r) {
    check_dev(*dst_r, at::kInt, X, "dst_r");
    check_dev(*rstage, at::kDouble, X, "rstage");
    TORCH_CHECK(dst_r->numel() == n * nen, "dst_r: one slot per element node");
  }
  check_status(fea_elem_residual_tangent((int)elem_type, (int)dtype, (int)n, conn.data_ptr<int32_t>(),
                                         X.data_ptr<double>(), U.data_ptr<double>(), params.data(),
                                         dst_k.data_ptr<int32_t>(), kstage.data_ptr(),
                                         with_r ? dst_r->data_ptr<int32_t>() : nullptr,
                                         with_r ? rstage->data_ptr<double>() : nullptr, stream_of(X)),
               "elem_residual_tangent");
}

void elem_tangent_direct(int64_t elem_type, int64_t dtype, int64_t val_dtype, const Tensor &order, const Tensor &conn,
                         const Tensor &X, const Tensor &U, c10::ArrayRef<double> params, const Tensor &dof_map,
                         const Tensor &rowptr, const Tensor &colpos, const Tensor &val, bool triple,
                         const c10::optional<Tensor> &dst_r, const c10::optional<Tensor> &rstage) {
  const int nen = nen_of(elem_type);
  float_type(dtype);
  check_dev(order, at::kInt, X, "order");
  check_dev(conn, at::kInt, X, "conn");
  check_dev(X, at::kDouble, X, "X");
  check_dev(U, at::kDouble, X, "U");
  check_dev(dof_map, at::kInt, X, "dof_map");
  check_dev(rowptr, at::kInt, X, "rowptr");
  check_dev(colpos, at::kInt, X, "colpos");
  check_dev(val, float_type(val_dtype), X, "val");
  check_nodes(X, U);
  check_params(params);
  const int64_t n_all = elems_of(conn, nen);
  TORCH_CHECK(colpos.numel() == n_all * nen * nen, "colpos: one entry per element node pair");
  TORCH_CHECK(dof_map.numel() == X.numel(), "dof_map: one entry per nodal DOF");
  TORCH_CHECK(order.numel() <= n_all, "order: at most one entry per element");
  const bool with_r = dst_r.has_value() && dst_r->defined();
  TORCH_CHECK(with_r == (rstage.has_value() && rstage->defined()), "dst_r and rstage: both or neither");
  if (with_r) {
    check_dev(*dst_r, at::kInt, X, "dst_r");
    check_dev(*rstage, at::kDouble, X, "rstage");
    TORCH_CHECK(dst_r->numel() == n_all * nen, "dst_r: one slot per element node");
  }
  check_status(fea_elem_tangent_direct((int)elem_type, (int)dtype, (int)val_dtype, (int)order.numel(),
                                       order.data_ptr<int32_t>(), conn.data_ptr<int32_t>(), X.data_ptr<double>(),
                                       U.data_ptr<double>(), params.data(), dof_map.data_ptr<int32_t>(),
                                       rowptr.data_ptr<int32_t>(), colpos.data_ptr<int32_t>(), val.data_ptr(),
                                       triple ? 1 : 0, with_r ? dst_r->data_ptr<int32_t>() : nullptr,
                                       with_r ? rstage->data_ptr<double>() : nullptr, stream_of(X)),
               "elem_tangent_direct");
}

void elem_residual(int64_t elem_type, int64_t dtype, const Tensor &conn, const Tensor &X, const Tensor &U,
                   c10::ArrayRef<double> params, const Tensor &dst_r, const Tensor &rstage) {
  const int nen = nen_of(elem_type);
  float_type(dtype);
  check_dev(conn, at::kInt, X, "conn");
  check_dev(X, at::kDouble, X, "X");
  check_dev(U, at::kDouble, X, "U");
  check_dev(dst_r, at::kInt, X, "dst_r");
  check_dev(rstage, at::kDouble, X, "rstage");
  check_nodes(X, U);
  check_params(params);
  const int64_t n = elems_of(conn, nen);
  TORCH_CHECK(dst_r.numel() == n * nen, "dst_r: one slot per element node");
  check_status(fea_elem_residual((int)elem_type, (int)dtype, (int)n, conn.data_ptr<int32_t>(), X.data_ptr<double>(),
                                 U.data_ptr<double>(), params.data(), dst_r.data_ptr<int32_t>(),
                                 rstage.data_ptr<double>(), stream_of(X)),
               "elem_residual");
}

void elem_energy(int64_t elem_type, const Tensor &conn, const Tensor &X, const Tensor &U, c10::ArrayRef<double> params,
                 const Tensor &partial) {
  const int nen = nen_of(elem_type);
  check_dev(conn, at::kInt, X, "conn");
  check_dev(X, at::kDouble, X, "X");
  check_dev(U, at::kDouble, X, "U");
  check_dev(partial, at::kDouble, X, "partial");
  check_nodes(X, U);
  check_params(params);
  const int64_t n = elems_of(conn, nen);
  TORCH_CHECK(partial.numel() >= fea_elem_energy_blocks((int)n), "partial: fea_elem_energy_blocks(n) entries needed");
  check_status(fea_elem_energy((int)elem_type, (int)n, conn.data_ptr<int32_t>(), X.data_ptr<double>(),
                               U.data_ptr<double>(), params.data(), partial.data_ptr<double>(), stream_of(X)),
               "elem_energy");
}

void gather_csr(int64_t stage_dtype, int64_t val_dtype, const Tensor &blk_a, const Tensor &blk_b,
                const Tensor &blk_colpos, const Tensor &kseg, const Tensor &dof_map, const Tensor &rowptr,
                const Tensor &kstage, const Tensor &csr_val) {
  check_dev(blk_a, at::kInt, kstage, "blk_a");
  check_dev(blk_b, at::kInt, kstage, "blk_b");
  check_dev(blk_colpos, at::kInt, kstage, "blk_colpos");
  check_dev(kseg, at::kInt, kstage, "kseg");
  check_dev(dof_map, at::kInt, kstage, "dof_map");
  check_dev(rowptr, at::kInt, kstage, "rowptr");
  check_dev(kstage, float_type(stage_dtype), kstage, "kstage");
  check_dev(csr_val, float_type(val_dtype), kstage, "csr_val");
  const int64_t nb = blk_a.numel();
  TORCH_CHECK(blk_b.numel() == nb && blk_colpos.numel() == nb && kseg.numel() == nb + 1,
              "block table: blk_a, blk_b, blk_colpos of one length, kseg one longer");
  check_status(fea_gather_csr((int)stage_dtype, (int)val_dtype, (int)nb, blk_a.data_ptr<int32_t>(),
                              blk_b.data_ptr<int32_t>(), blk_colpos.data_ptr<int32_t>(), kseg.data_ptr<int32_t>(),
                              dof_map.data_ptr<int32_t>(), rowptr.data_ptr<int32_t>(), kstage.data_ptr(),
                              csr_val.data_ptr(), stream_of(kstage)),
               "gather_csr");
}

void gather_residual(const Tensor &rseg, const Tensor &dof_map, const Tensor &rstage, const Tensor &R) {
  check_dev(rseg, at::kInt, rstage, "rseg");
  check_dev(dof_map, at::kInt, rstage, "dof_map");
  check_dev(rstage, at::kDouble, rstage, "rstage");
  check_dev(R, at::kDouble, rstage, "R");
  const int64_t n_nodes = rseg.numel() - 1;
  TORCH_CHECK(n_nodes >= 0 && dof_map.numel() == 3 * n_nodes, "rseg [n_nodes + 1] and dof_map [3 n_nodes]");
  check_status(fea_gather_residual((int)n_nodes, rseg.data_ptr<int32_t>(), dof_map.data_ptr<int32_t>(),
                                   rstage.data_ptr<double>(), R.data_ptr<double>(), stream_of(rstage)),
               "gather_residual");
}

bool valid_pair(int64_t m, int64_t v) {
  return (m == FEA_F64 && v == FEA_F64) || (m == FEA_F32 && v == FEA_F64) || (m == FEA_F32 && v == FEA_F32);
}

struct Csr {
  int n;
  int64_t nnz;
  const int32_t *grp;
};

// the CSR operands shared by the SpMV and PCG operators; returns the row count
Csr check_csr(int64_t mat_dtype, int64_t n_groups, const c10::optional<Tensor> &grp, const Tensor &rowptr,
              const Tensor &col, const Tensor &val) {
  check_dev(rowptr, at::kInt, val, "rowptr");
  check_dev(col, at::kInt, val, "col");
  check_dev(val, float_type(mat_dtype), val, "val");
  TORCH_CHECK(rowptr.numel() >= 2, "rowptr: n + 1 entries, n >= 1");
  Csr c{(int)(rowptr.numel() - 1), val.numel(), nullptr};
  if (grp.has_value() && grp->defined()) {
    check_dev(*grp, at::kInt, val, "grp");
    const int64_t ng = n_groups < 0 ? -n_groups : n_groups;
    TORCH_CHECK(grp->numel() == ng + 1, "grp: |n_groups| + 1 entries");
    c.grp = grp->data_ptr<int32_t>();
  } else {
    TORCH_CHECK(n_groups == 0, "n_groups != 0 needs grp");
  }
  return c;
}

void check_vec(const Tensor &t, int64_t vec_dtype, int n, const Tensor &ref, const char *name) {
  check_dev(t, float_type(vec_dtype), ref, name);
  TORCH_CHECK(t.numel() == n, name, ": length ", n, " expected, got ", t.numel());
}

void check_dinv(const Tensor &dinv, int64_t precond, int64_t vec_dtype, int n, int64_t n_groups, const Tensor &ref) {
  TORCH_CHECK(precond == 0 || precond == 1, "precond must be 0 (point Jacobi) or 1 (node-block Jacobi)");
  if (precond == 0) {
    check_vec(dinv, vec_dtype, n, ref, "dinv");
  } else {
    check_dev(dinv, at::kFloat, ref, "dinv");
    const int64_t ng = n_groups < 0 ? -n_groups : n_groups;
    TORCH_CHECK(ng > 0 && dinv.numel() == 6 * ng, "dinv: 6 floats per row group for block Jacobi");
  }
}

void spmv(int64_t mat_dtype, int64_t vec_dtype, int64_t n_groups, const c10::optional<Tensor> &grp,
          const Tensor &rowptr, const Tensor &col, const Tensor &val, const Tensor &x, const Tensor &y) {
  TORCH_CHECK(valid_pair(mat_dtype, vec_dtype), "unsupported (matrix, vector) precision pair");
  const Csr c = check_csr(mat_dtype, n_groups, grp, rowptr, col, val);
  check_vec(x, vec_dtype, c.n, val, "x");
  check_vec(y, vec_dtype, c.n, val, "y");
  check_status(fea_spmv((int)mat_dtype, (int)vec_dtype, c.n, c.nnz, (int)n_groups, c.grp, rowptr.data_ptr<int32_t>(),
                        col.data_ptr<int32_t>(), val.data_ptr(), x.data_ptr(), y.data_ptr(), stream_of(val)),
               "spmv");
}

void pcg_start(int64_t mat_dtype, int64_t vec_dtype, int64_t n_groups, const c10::optional<Tensor> &grp,
               const Tensor &rowptr, const Tensor &col, const Tensor &val, int64_t precond, const Tensor &dinv,
               const Tensor &b, const Tensor &x, const Tensor &r, const Tensor &z, const Tensor &p, const Tensor &ws) {
  TORCH_CHECK(valid_pair(mat_dtype, vec_dtype), "unsupported (matrix, vector) precision pair");
  const Csr c = check_csr(mat_dtype, n_groups, grp, rowptr, col, val);
  check_dinv(dinv, precond, vec_dtype, c.n, n_groups, val);
  check_vec(b, vec_dtype, c.n, val, "b");
  check_vec(x, vec_dtype, c.n, val, "x");
  check_vec(r, vec_dtype, c.n, val, "r");
  check_vec(z, vec_dtype, c.n, val, "z");
  check_vec(p, vec_dtype, c.n, val, "p");
  check_dev(ws, at::kDouble, val, "ws");
  TORCH_CHECK(ws.numel() >= fea_pcg_ws_size(c.n), "ws: fea_pcg_ws_size(n) doubles needed");
  check_status(fea_pcg_start((int)mat_dtype, (int)vec_dtype, c.n, c.nnz, (int)n_groups, c.grp,
                             rowptr.data_ptr<int32_t>(), col.data_ptr<int32_t>(), val.data_ptr(), (int)precond,
                             dinv.data_ptr(), b.data_ptr(), x.data_ptr(), r.data_ptr(), z.data_ptr(), p.data_ptr(),
                             ws.data_ptr<double>(), stream_of(val)),
               "pcg_start");
}

void pcg_iterate(int64_t mat_dtype, int64_t vec_dtype, int64_t n_groups, const c10::optional<Tensor> &grp,
                 const Tensor &rowptr, const Tensor &col, const Tensor &val, int64_t precond, const Tensor &dinv,
                 const Tensor &x, const Tensor &r, const Tensor &z, const Tensor &p, const Tensor &q, const Tensor &ws,
                 double tol, int64_t max_iter, int64_t iters, const c10::optional<Tensor> &status) {
  TORCH_CHECK(valid_pair(mat_dtype, vec_dtype), "unsupported (matrix, vector) precision pair");
  TORCH_CHECK(max_iter >= 1 && iters >= 0, "max_iter >= 1 and iters >= 0 required");
  const Csr c = check_csr(mat_dtype, n_groups, grp, rowptr, col, val);
  check_dinv(dinv, precond, vec_dtype, c.n, n_groups, val);
  check_vec(x, vec_dtype, c.n, val, "x");
  check_vec(r, vec_dtype, c.n, val, "r");
  check_vec(z, vec_dtype, c.n, val, "z");
  check_vec(p, vec_dtype, c.n, val, "p");
  check_vec(q, vec_dtype, c.n, val, "q");
  check_dev(ws, at::kDouble, val, "ws");
  TORCH_CHECK(ws.numel() >= fea_pcg_ws_size(c.n), "ws: fea_pcg_ws_size(n) doubles needed");
  double *st = nullptr;
  if (status.has_value() && status->defined()) {
    TORCH_CHECK(status->device().is_cpu() && status->scalar_type() == at::kDouble && status->is_contiguous() &&
                    status->numel() >= 4,
                "status: a contiguous float64 CPU tensor of 4 entries");
    st = status->data_ptr<double>();
  }
  check_status(fea_pcg_iterate((int)mat_dtype, (int)vec_dtype, c.n, c.nnz, (int)n_groups, c.grp,
                               rowptr.data_ptr<int32_t>(), col.data_ptr<int32_t>(), val.data_ptr(), (int)precond,
                               dinv.data_ptr(), x.data_ptr(), r.data_ptr(), z.data_ptr(), p.data_ptr(), q.data_ptr(),
                               ws.data_ptr<double>(), tol, (int)max_iter, (int)iters, st, stream_of(val)),
               "pcg_iterate");
}

void jacobi_inv(int64_t mat_dtype, int64_t vec_dtype, const Tensor &diag_slot, const Tensor &val, const Tensor &dinv) {
  TORCH_CHECK(valid_pair(mat_dtype, vec_dtype), "unsupported (matrix, vector) precision pair");
  check_dev(diag_slot, at::kInt, val, "diag_slot");
  check_dev(val, float_type(mat_dtype), val, "val");
  const int n = (int)diag_slot.numel();
  check_vec(dinv, vec_dtype, n, val, "dinv");
  check_status(fea_jacobi_inv((int)mat_dtype, (int)vec_dtype, n, diag_slot.data_ptr<int32_t>(), val.data_ptr(),
                              dinv.data_ptr(), stream_of(val)),
               "jacobi_inv");
}

void block_jacobi_inv(int64_t mat_dtype, int64_t n_groups, const Tensor &grp, const Tensor &rowptr, const Tensor &col,
                      const Tensor &diag_slot, const Tensor &val, const Tensor &binv) {
  const Csr c = check_csr(mat_dtype, n_groups, grp, rowptr, col, val);
  check_dev(diag_slot, at::kInt, val, "diag_slot");
  TORCH_CHECK(diag_slot.numel() == c.n, "diag_slot: one entry per row");
  check_dev(binv, at::kFloat, val, "binv");
  const int64_t ng = n_groups < 0 ? -n_groups : n_groups;
  TORCH_CHECK(binv.numel() == 6 * ng, "binv: 6 floats per row group");
  check_status(fea_block_jacobi_inv((int)mat_dtype, c.n, (int)n_groups, c.grp, rowptr.data_ptr<int32_t>(),
                                    col.data_ptr<int32_t>(), diag_slot.data_ptr<int32_t>(), val.data_ptr(),
                                    binv.data_ptr<float>(), stream_of(val)),
               "block_jacobi_inv");
}

}  // namespace

TORCH_LIBRARY(fea, m) {
  m.def("elem_residual_tangent(int elem_type, int dtype, Tensor conn, Tensor X, Tensor U, float[] params, "
        "Tensor dst_k, Tensor(a!) kstage, Tensor? dst_r, Tensor(b!)? rstage) -> ()");
  m.def("elem_tangent_direct(int elem_type, int dtype, int val_dtype, Tensor order, Tensor conn, Tensor X, "
        "Tensor U, float[] params, Tensor dof_map, Tensor rowptr, Tensor colpos, Tensor(a!) val, bool triple, "
        "Tensor? dst_r, Tensor(b!)? rstage) -> ()");
  m.def("elem_residual(int elem_type, int dtype, Tensor conn, Tensor X, Tensor U, float[] params, Tensor dst_r, "
        "Tensor(a!) rstage) -> ()");
  m.def("elem_energy(int elem_type, Tensor conn, Tensor X, Tensor U, float[] params, Tensor(a!) partial) -> ()");
  m.def("gather_csr(int stage_dtype, int val_dtype, Tensor blk_a, Tensor blk_b, Tensor blk_colpos, Tensor kseg, "
        "Tensor dof_map, Tensor rowptr, Tensor kstage, Tensor(a!) csr_val) -> ()");
  m.def("gather_residual(Tensor rseg, Tensor dof_map, Tensor rstage, Tensor(a!) R) -> ()");
  m.def("spmv(int mat_dtype, int vec_dtype, int n_groups, Tensor? grp, Tensor rowptr, Tensor col, Tensor val, "
        "Tensor x, Tensor(a!) y) -> ()");
  m.def("pcg_start(int mat_dtype, int vec_dtype, int n_groups, Tensor? grp, Tensor rowptr, Tensor col, Tensor val, "
        "int precond, Tensor dinv, Tensor b, Tensor(a!) x, Tensor(b!) r, Tensor(c!) z, Tensor(d!) p, "
        "Tensor(e!) ws) -> ()");
  m.def("pcg_iterate(int mat_dtype, int vec_dtype, int n_groups, Tensor? grp, Tensor rowptr, Tensor col, "
        "Tensor val, int precond, Tensor dinv, Tensor(a!) x, Tensor(b!) r, Tensor(c!) z, Tensor(d!) p, "
        "Tensor(e!) q, Tensor(f!) ws, float tol, int max_iter, int iters, Tensor(g!)? status) -> ()");
  m.def("jacobi_inv(int mat_dtype, int vec_dtype, Tensor diag_slot, Tensor val, Tensor(a!) dinv) -> ()");
  m.def("block_jacobi_inv(int mat_dtype, int n_groups, Tensor grp, Tensor rowptr, Tensor col, Tensor diag_slot, "
        "Tensor val, Tensor(a!) binv) -> ()");
}

TORCH_LIBRARY_IMPL(fea, CUDA, m) {
  m.impl("elem_residual_tangent", &elem_residual_tangent);
  m.impl("elem_residual", &elem_residual);
  m.impl("elem_tangent_direct", &elem_tangent_direct);
  m.impl("elem_energy", &elem_energy);
  m.impl("gather_csr", &gather_csr);
  m.impl("gather_residual", &gather_residual);
  m.impl("spmv", &spmv);
  m.impl("pcg_start", &pcg_start);
  m.impl("pcg_iterate", &pcg_iterate);
  m.impl("jacobi_inv", &jacobi_inv);
  m.impl("block_jacobi_inv", &block_jacobi_inv);
}
