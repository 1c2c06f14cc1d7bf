/*
 * fea_amd — C ABI of the MI355X (gfx950) nonlinear solid-mechanics hot path.
 *
 * Conventions (every entry point):
 *   - returns int status: FEA_OK (0) or a negative FEA_ERR_* code; the message
 *     of the last failure on the calling thread is fea_last_error();
 *   - arrays are DEVICE pointers unless the parameter name ends in `_host`;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *   - no entry point allocates device memory except fea_pattern_* (host) and
 *     the explicit workspace helpers; kernels run asynchronously on `stream`.
 *
 * Layouts: nodal coordinates / displacements are AoS float64 [n_nodes][3]
 * (row-major, the reference's RGC block layout, assembly.py:97-128);
 * connectivity is int32 [n_elem][nen] in the reference node order
 * (tetrahedral.py:6-22 C3D4, brick.py:13-22 C3D8, tetrahedral.py:66-86 C3D10).
 * `dof_map[3*node+comp]` is the free-DOF index or -1 for a constrained DOF
 * (the GC numbering of assembly.py:170-183 / _RGC2GC :494-499).
 *
 * The reference (sssnnn7119/FEA) is pure PyTorch; each function below names the
 * reference Python function whose device work it replaces.
 */
#ifndef FEA_AMD_H
#define FEA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEA_OK 0
#define FEA_ERR_ARG (-1)
#define FEA_ERR_HIP (-2)
#define FEA_ERR_UNSUPPORTED (-3)
#define FEA_ERR_STATE (-4)

/* precision of a kernel's arithmetic and of its floating outputs */
#define FEA_F32 0
#define FEA_F64 1
/* fea_elem_residual_tangent only: float64 element arithmetic, each staged block
 * rounded once to float32 (kstage is float32) */
#define FEA_F64_STAGE_F32 2

/* element types = nodes per element */
#define FEA_C3D4 4
#define FEA_C3D8 8
#define FEA_C3D10 10

/* surface facet types = nodes per facet (T3, T6, Q4) */
#define FEA_T3 3
#define FEA_T6 6
#define FEA_Q4 4

/*
 * Material: isochoric two-invariant hyperelasticity with quadratic volumetric part
 *   W = w1 (I1bar - 3) + w2 (I2bar - 3) + kv/2 (J - 1)^2,  params = {w1, w2, kv}.
 * NeoHookean(mu, kappa) of hyperelastic.py:6-84 is {mu/2, 0, kappa};
 * Mooney-Rivlin(C10, C01, D1) (absent in the reference) is {C10, C01, 2/D1}.
 */

const char *fea_last_error(void);
int fea_version(void);

/* ---------------------------------------------------------------- pattern --
 * Host-side integer work (replaces Element_3D.initialize's COO pattern and
 * coalesce map, C3base.py:109-153, and the free-DOF compaction + COO->CSR
 * coalesce of assembly.py:371-380 / _linear_solver.py:18-19).
 *
 * Groups are element groups or surface-facet groups; all connectivity uses
 * global node ids.  For every group with in_k != 0 its (e,a,b) node-pair
 * blocks enter the CSR sparsity; every group gets residual staging slots.
 * The result is the free-DOF CSR (rows sorted, columns ascending, duplicates
 * merged) plus the staging maps the assembly kernels write through.
 */
typedef struct fea_pattern fea_pattern;

int fea_pattern_create(int n_nodes, const int32_t *dof_map_host, int n_groups,
                       const int32_t *const *conn_host, const int32_t *n_elem_host,
                       const int32_t *nen_host, const int32_t *in_k_host,
                       fea_pattern **out);
/* sizes[0]=n_free, [1]=nnz, [2]=n_blocks, [3]=n_kstage (blocks of 9),
 * [4]=n_rstage (blocks of 3) */
int fea_pattern_sizes(const fea_pattern *p, int64_t sizes[5]);
/* host outputs; arrays sized from fea_pattern_sizes:
 *   rowptr[n_free+1], col[nnz], diag_slot[n_free],
 *   blk_a[n_blocks], blk_b[n_blocks], blk_colpos[n_blocks], kseg[n_blocks+1],
 *   rseg[n_nodes+1] */
int fea_pattern_export_csr(const fea_pattern *p, int32_t *rowptr_host, int32_t *col_host,
                           int32_t *diag_slot_host);
int fea_pattern_export_blocks(const fea_pattern *p, int32_t *blk_a_host, int32_t *blk_b_host,
                              int32_t *blk_colpos_host, int32_t *kseg_host, int32_t *rseg_host);
/* group maps: dst_k[n_elem*nen*nen] (only for in_k groups; -1 = block with no
 * free x free entry), dst_r[n_elem*nen] (-1 = node with no free DOF) */
int fea_pattern_export_group(const fea_pattern *p, int group, int32_t *dst_k_host,
                             int32_t *dst_r_host);
/* colpos[n_elem*nen*nen] of an in_k group: the column position of its (element, a, b)
 * block inside the CSR rows of node a (blk_colpos of the block), -1 where dst_k is -1;
 * the map of fea_elem_tangent_direct.  Replaces, with the colouring below, the
 * scatter-add of element blocks into the global matrix (C3base.py:449,
 * assembly.py:347-380). */
int fea_pattern_export_group_colpos(const fea_pattern *p, int group, int32_t *colpos_host);
void fea_pattern_destroy(fea_pattern *p);
/* Greedy element colouring: color[e] in 0..n_colors-1, no two elements of one colour
 * share a node (elements in order, smallest colour free of coloured neighbours). */
int fea_color_elements(int n_nodes, int n_elem, int nen, const int32_t *conn, int32_t *color,
                       int32_t *n_colors);

/* ---------------------------------------------------------------- elements --
 * fea_elem_residual_tangent replaces Element_3D.structural_Force
 * (C3base.py:413-452) + components_Solid (:454-480) +
 * NeoHookean.material_Constitutive_C3 (hyperelastic.py:35-84):
 * per element and Gauss point it forms F = I + grad u from X and U, the
 * stress and consistent tangent, and writes the 3x3 blocks K_ab (dtype; float32
 * for FEA_F64_STAGE_F32) to
 * kstage[dst_k[..]*9 ..] and the nodal forces R_a (float64) to
 * rstage[dst_r[..]*3 ..].  kstage/dst_k may be NULL (residual only);
 * rstage/dst_r may be NULL (tangent only).  Gauss-point geometry (dN/dX,
 * w*detJ, C3base.py:155-216) is recomputed from X in registers.
 */
int fea_elem_residual_tangent(int elem_type, int dtype, int n_elem, const int32_t *conn,
                              const double *X, const double *U, const double *params_host,
                              const int32_t *dst_k, void *kstage, const int32_t *dst_r,
                              double *rstage, void *stream);

/* fea_elem_residual_tangent for the elements order[0 .. n_elem) of one colour (no
 * two share a node), adding each 3x3 block K_ab into the CSR values val (val_dtype)
 * at rowptr[row] + colpos[(e*nen + a)*nen + b] + free-column offset instead of
 * staging it; the residual goes to rstage as there.  Colours launched one after
 * another on one stream give deterministic sums.  dtype/val_dtype: F64/F64, F64/F32
 * (float64 contraction, each block added into float32 values) or F32/F32.  triple != 0
 * asserts every node is fully free or fully fixed (the free-column offset of a block
 * entry is then its component; dof_map is read for the rows only).  Replaces
 * structural_Force + the scatter-add of C3base.py:413-452 and assembly.py:347-380. */
int fea_elem_tangent_direct(int elem_type, int dtype, int val_dtype, int n_elem, const int32_t *order,
                            const int32_t *conn, const double *X, const double *U,
                            const double *params_host, const int32_t *dof_map, const int32_t *rowptr,
                            const int32_t *colpos, void *val, int triple, const int32_t *dst_r,
                            double *rstage, void *stream);

/* Residual only (force branch of structural_Force, C3base.py:428-431), float64 or
 * float32 arithmetic, float64 output into rstage. */
int fea_elem_residual(int elem_type, int dtype, int n_elem, const int32_t *conn,
                      const double *X, const double *U, const double *params_host,
                      const int32_t *dst_r, double *rstage, void *stream);

/* The explicit loop's residual with a per-tile reduction (the force branch of
 * structural_Force, C3base.py:428-431, and its scatter-add, assembly.py:347-380):
 * workgroup t takes elements [t*TE, (t+1)*TE) (TE = fea_elem_tile_size(elem_type)),
 * keeps their nodal forces (arithmetic in dtype) in LDS, and writes for every partial
 * p in tile_lbase[t] .. tile_lbase[t+1] the float64 sum, in order, of the local slots
 * pslot[pslot_ptr[p] .. pslot_ptr[p+1]) (slot = e_local*nen + a) to partial[3p ..].
 * One partial per (tile, node): fea_explicit_update_tiles adds a node's partials.
 * Optional (tnode_ptr non-NULL): tile t's nodes tnode[tnode_ptr[t] .. tnode_ptr[t+1])
 * (at most max_tile_nodes <= 2048) have their X and U staged in LDS first, and
 * lconn[e*nen + a] is element e's node a as an index into its tile's list: the same
 * arithmetic on the same values (bit-identical partials), without per-element gathers. */
int fea_elem_tile_size(int elem_type);
int fea_elem_residual_tiles(int elem_type, int dtype, int n_elem, const int32_t *conn,
                            const double *X, const double *U, const double *params_host,
                            const int32_t *tile_lbase, const int32_t *pslot_ptr, const uint16_t *pslot,
                            double *partial, const int32_t *tnode_ptr, const int32_t *tnode,
                            const uint16_t *lconn, int max_tile_nodes, void *stream);

/* Strain energy per element block (Element_3D.potential_Energy,
 * C3base.py:387-411): writes one float64 partial sum per 256-element block to
 * partial[0 .. ceil(n_elem/256)). */
int fea_elem_energy(int elem_type, int n_elem, const int32_t *conn, const double *X,
                    const double *U, const double *params_host, double *partial, void *stream);
int fea_elem_energy_blocks(int n_elem);

/* Gauss-point geometry export (Element_3D._pre_load_gaussian, C3base.py:155-216):
 * dNdX [g][n_elem][3][nen], wdet [g][n_elem], N [g][n_elem][nen] (float64). */
int fea_elem_prepare(int elem_type, int n_elem, const int32_t *conn, const double *X,
                     double *dNdX, double *wdet, double *N, void *stream);

/* Row-sum lumped mass restricted to free DOFs (Element_3D.get_mass_matrix
 * C3base.py:348-384 -> Assembly.assemble_mass_matrix assembly.py:421-443 ->
 * DynamicExplicitSolver.get_lumped_mass_inv dynamic_explicit.py:147-166):
 * writes per (e,a) the 3 component masses sum_b(free) M_ab into rstage. */
int fea_elem_lumped_mass(int elem_type, int n_elem, const int32_t *conn, const double *X,
                         const int32_t *dof_map, double rho, const int32_t *dst_r,
                         double *rstage, void *stream);

/* Consistent mass blocks (Element_3D.get_mass_matrix, C3base.py:348-384, assembled
 * by Assembly.assemble_mass_matrix, assembly.py:421-443): M_ab = rho sum_g N_a N_b
 * w detJ on the diagonal of each 3x3 block, float64, written to kstage through
 * the tangent's dst_k map (fea_gather_csr then forms the CSR in K's pattern;
 * surface groups' slots must hold zeros). */
int fea_elem_mass(int elem_type, int n_elem, const int32_t *conn, const double *X, double rho,
                  const int32_t *dst_k, double *kstage, void *stream);

/* ---------------------------------------------------------------- surfaces --
 * Follower pressure on T3/T6/Q4 facets (Pressure.get_stiffness,
 * pressure.py:80-132, with the assembly's load sign, assembly.py:337-341):
 * writes scale*dV/dx to rstage and scale*d2V/dx2 blocks to kstage, where
 * V = 1/3 sum_g w_g det[x, x_xi, x_eta] and scale = +pressure. */
int fea_surface_pressure(int surf_type, int dtype, int n_facets, const int32_t *conn,
                         const double *X, const double *U, double scale,
                         const int32_t *dst_k, void *kstage, const int32_t *dst_r,
                         double *rstage, void *stream);
/* Enclosed-volume partials V per 256-facet block (pressure.py:134-159). */
int fea_surface_volume(int surf_type, int n_facets, const int32_t *conn, const double *X,
                       const double *U, double *partial, void *stream);
/* Dead-load traction (new; no reference counterpart): writes -scale * f_a with
 * f_a = sum_g N_a t w_g detJ0 into rstage (the residual sign of a load force,
 * assembly.py:337-341); `work_partial` (optional) receives per-block sums of
 * f . u for the load potential. */
int fea_surface_traction(int surf_type, int n_facets, const int32_t *conn, const double *X,
                         const double *U, const double *traction_host, double scale,
                         const int32_t *dst_r, double *rstage, double *work_partial,
                         void *stream);
int fea_surface_blocks(int n_facets);

/* ---------------------------------------------------------------- assembly --
 * Deterministic, atomic-free gathers replacing the scatter_add calls of
 * C3base.py:449 and assembly.py:347-354 plus the reduction of :371-384. */
/* stage_dtype: precision of kstage (the element kernels' dtype); val_dtype: of
 * the CSR values.  The block sums run in stage_dtype and are rounded once to
 * val_dtype (pairs F64->F64, F64->F32, F32->F32). */
int fea_gather_csr(int stage_dtype, int val_dtype, int n_blocks, const int32_t *blk_a,
                   const int32_t *blk_b, const int32_t *blk_colpos, const int32_t *kseg,
                   const int32_t *dof_map, const int32_t *rowptr, const void *kstage, void *csr_val,
                   void *stream);
/* R[dof] = sum of the node's staged contributions, for free DOFs */
int fea_gather_residual(int n_nodes, const int32_t *rseg, const int32_t *dof_map,
                        const double *rstage, double *R, void *stream);

/* Deterministic sum of n float64 partials into out[0] (single workgroup). */
int fea_reduce_sum(int n, const double *partial, double *out, void *stream);

/* ---------------------------------------------------------------- linear algebra --
 * Precision pairs (matrix values, vectors): (FEA_F64, FEA_F64) plain FP64,
 * (FEA_F32, FEA_F64) FP32 matrix storage with FP64 vectors / accumulation,
 * (FEA_F32, FEA_F32) FP32.
 * Row groups: grp[g] (int32 [|n_groups|+1]) is the first row of group g; the rows
 * of a group (at most 3, consecutive) must share one column list, as the free
 * rows of one node do in the fea_pattern CSR.  grp == NULL: one row per group.
 * n_groups < 0 additionally asserts that every group has 3 rows and a column
 * list made of whole neighbour-node triples (3 consecutive column indices),
 * which selects the triple-blocked SpMV; `col` is then the compact triple-column
 * array, not the CSR's: one entry (the first column) per neighbour triple of each
 * group, group g's at rowptr[grp[g]] / 9 + j (nnz / 9 entries).  This also holds
 * for the PCG entry points and the batched systems below.
 *
 * CSR SpMV y = A x (the CSR `A @ p` of _linear_solver.py:27,32). */
int fea_spmv(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
             const int32_t *rowptr, const int32_t *col, const void *val, const void *x, void *y,
             void *stream);
/* Node-block Jacobi (an addition to the reference's point Jacobi): binv[6 g .. 6 g + 5]
 * = the symmetric inverse (00, 01, 02, 11, 12, 22) of the diagonal block of row group g
 * (the free rows of one node, consecutive columns), in float; a block that is not positive
 * definite falls back to the point-Jacobi inverse of its rows. */
int fea_block_jacobi_inv(int mat_dtype, int n, int n_groups, const int32_t *grp, const int32_t *rowptr,
                         const int32_t *col, const int32_t *diag_slot, const void *val, float *binv,
                         void *stream);
/* Jacobi diagonal of static_implicit.py:314-317: dinv[r] = 1/|A_rr| (1 if 0). */
int fea_jacobi_inv(int mat_dtype, int vec_dtype, int n, const int32_t *diag_slot, const void *val,
                   void *dinv, void *stream);

/* Jacobi-preconditioned CG, the conjugate_gradient of _linear_solver.py:4-42 on
 * the symmetrically scaled system of static_implicit.py:314-321 (identical
 * iterates), with all scalars on the device and no host sync per iteration.
 * ws: float64 workspace of fea_pcg_ws_size(n) doubles; vectors (vec_dtype) x, b,
 * r, z, p, q of length n.  fea_pcg_start sets r = b - A x, z = dinv r, p = z;
 * fea_pcg_iterate enqueues `iters` iterations (each a no-op once converged:
 * r.z / b.(dinv b) < tol, or max_iter reached) and, if status_host != NULL,
 * copies {iterations, done, r.z, b.(dinv b)} to it (synchronises the stream).
 * precond 0: point Jacobi, dinv = vec_dtype[n] (fea_jacobi_inv; the reference's
 * scaling).  precond 1: node-block Jacobi over the row groups (grp required), dinv =
 * float[6 * |n_groups|] from fea_block_jacobi_inv; an addition that changes the
 * iterates, not the solution. */
int64_t fea_pcg_ws_size(int n);
int fea_pcg_start(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                  const int32_t *rowptr, const int32_t *col, const void *val, int precond,
                  const void *dinv, const void *b, void *x, void *r, void *z, void *p, double *ws,
                  void *stream);
int fea_pcg_iterate(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups,
                    const int32_t *grp, const int32_t *rowptr, const int32_t *col, const void *val,
                    int precond, const void *dinv, void *x, void *r, void *z, void *p, void *q,
                    double *ws, double tol, int max_iter, int iters, double *status_host, void *stream);
/* Allocation-free PCG: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): fea_pcg_start / fea_pcg_iterate with the float64 workspace
 * taken from a static device array of the library (one per GPU, 65 664 B, held until
 * the process ends) instead of a `ws` argument.  b != NULL starts a solve (the start
 * stage of fea_pcg_start, then `iters` iterations); b == NULL continues the solve the
 * previous call on this GPU started.  Vectors, matrix and preconditioner are the
 * caller's, as for fea_pcg_iterate; status_host as there (NULL: no host sync, so the
 * call can be captured).  Calls that overlap on one GPU share the workspace and race;
 * a size whose workspace exceeds the array is rejected with FEA_ERR_ARG.  Nothing is
 * allocated or freed. */
int fea_pcg_run_noalloc(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                        const int32_t *rowptr, const int32_t *col, const void *val, int precond,
                        const void *dinv, const void *b, void *x, void *r, void *z, void *p, void *q,
                        double tol, int max_iter, int iters, double *status_host, void *stream);
/* ---- smoothed-aggregation multigrid preconditioner (an addition: the reference's CG uses
 * point-Jacobi scaling, static_implicit.py:314-321; same solution, far fewer iterations).
 * Host set-up, once per model: fea_amg_build aggregates the node groups of the free-DOF
 * CSR (host arrays; rows grouped by node as `grp`, 1..3 rows each), builds the tentative
 * prolongator from the six rigid-body modes (dof_xyz: coordinates of each row's node,
 * dof_comp: its component 0..2), smooths it with one damped block-Jacobi step, forms the
 * Galerkin coarse operators (6 DOFs per aggregate) until at most max_coarse DOFs remain
 * and inverts the coarsest densely.  fea_amg_level_sizes / fea_amg_level_export copy a
 * level out (sizes: rows, aggregates of the next level, prolongator blocks, operator
 * nonzeros, row groups, dense order); any NULL output is skipped. */
typedef struct fea_amg_host fea_amg_host;
int fea_amg_build(int n, const int32_t *rowptr, const int32_t *col, const double *val, int n_groups,
                  const int32_t *grp, const double *dof_xyz, const int32_t *dof_comp, int max_coarse,
                  int max_levels, int n_threads, fea_amg_host **out);
int fea_amg_num_levels(const fea_amg_host *h);
int fea_amg_level_sizes(const fea_amg_host *h, int level, int64_t *sizes);
int fea_amg_level_export(const fea_amg_host *h, int level, int32_t *prow, int32_t *pagg, float *pval,
                         int32_t *pt_ptr, int32_t *pt_blk, int32_t *pt_row, int32_t *rowptr, int32_t *col,
                         double *val, float *dinv36, double *rho, float *dense_inv);
void fea_amg_destroy(fea_amg_host *h);

/* A level of the hierarchy on the device (caller-owned buffers; a host array of these is
 * passed to the calls below). */
typedef struct fea_amg_level {
  int32_t n;                  /* rows of this level */
  int32_t nagg;               /* aggregates of the next level (its rows / 6); 0 on the coarsest */
  double rho;                 /* lambda_max(D^-1 A) of the level's smoother (levels >= 1) */
  const int32_t *rowptr;      /* levels >= 1: operator rows, in groups of 6 sharing one column list */
  const int32_t *col;
  const float *val;
  const float *dinv;          /* levels >= 1: 6x6 inverse of each group's diagonal block */
  const int32_t *prow;        /* prolongator to the next level: per row, blocks of 6 columns */
  const int32_t *pagg;        /* aggregate of each block */
  const uint16_t *pval;       /* 6 values per block, IEEE binary16 */
  const int32_t *pt_ptr;      /* transpose: per aggregate, its entries (row, 6 binary16 values) */
  const int32_t *pt_row;
  const uint16_t *ptval;
  const float *dense_inv;     /* coarsest level: n x n inverse */
  double *b, *x, *x2, *d, *r; /* work vectors of length n (b unused on level 0) */
  /* level 0 only, may be NULL: the level-0 matrix values in IEEE binary16, times a power of
   * two (fea_amg_half_values), with val16_scale[1] the power that undoes it.  When set, the
   * V-cycle's level-0 smoothing and residual read these (half the value bytes of a float
   * matrix); the PCG product keeps the matrix itself. */
  const uint16_t *val16;
  const double *val16_scale;
} fea_amg_level;
int fea_amg_level_size(void);
/* z = V-cycle(b): level 0 is the matrix (node-grouped CSR as for fea_spmv, current values)
 * with its node-block Jacobi inverses binv (fea_block_jacobi_inv) and lambda_max estimate
 * rho0 (fea_amg_power); Chebyshev smoothing of the given degree on [rho/30, 1.1 rho]. */
int fea_amg_vcycle(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                   const int32_t *col, const void *val, const float *binv, double rho0,
                   const fea_amg_level *levels, int n_levels, int degree, const double *b, double *z,
                   void *stream);
/* One Chebyshev smoothing step of the V-cycle's level 0: d = c1 d + c2 M^-1 (b - A x),
 * xo = x + d with M the node-block Jacobi inverses binv; x == NULL is the first step from
 * x = 0 (no product, c1 ignored).  xo must not alias x. */
int fea_amg_smooth0(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                    const int32_t *col, const void *val, const float *binv, const double *b, const double *x,
                    double *xo, double *d, double c1, double c2, void *stream);
/* val16 = binary16(2^k val) for the V-cycle's level 0, k chosen from max |val| so that the
 * largest value lands in [2^14, 2^15); scale[1] = 2^-k, scale[0] is work.  The matrix must be
 * triple-structured (n_groups < 0 in the SpMV calls; nnz a multiple of 3): each triple of
 * values t becomes the four binary16 val16[4t .. 4t+3] (the fourth zero; val16 holds 4 nnz / 3
 * entries), read with one 8-byte load.  Values below 2^-24 of the scaled range flush to zero
 * (no counterpart in the reference: an operand of the multigrid preconditioner, whose
 * result the outer PCG still tests against the matrix). */
int fea_amg_half_values(int mat_dtype, int64_t nnz, const void *val, uint16_t *val16, double *scale,
                        void *stream);
/* fea_amg_smooth0 on the binary16 values of fea_amg_half_values (scale as written there). */
int fea_amg_smooth0_half(int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                         const int32_t *col, const uint16_t *val16, const double *scale, const float *binv,
                         const double *b, const double *x, double *xo, double *d, double c1, double c2,
                         void *stream);
/* lambda_max(D^-1 A) of the level-0 matrix by `iters` power steps from x (unit 2-norm;
 * x, y float work vectors, partial: float64 scratch of the SpMV grid); synchronises. */
int fea_amg_power(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                  const int32_t *col, const void *val, const float *binv, float *x, float *y, double *partial,
                  int iters, double *rho_out, void *stream);
/* PCG preconditioned by the V-cycle: as fea_pcg_start / fea_pcg_iterate with float64
 * vectors; the stopping test is r.z / b.(M^-1 b) < tol as for the Jacobi PCG.  The
 * V-cycle stores operators in float and keeps every vector and sum in float64. */
int fea_pcg_start_amg(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp, const int32_t *rowptr,
                      const int32_t *col, const void *val, const float *binv, double rho0,
                      const fea_amg_level *levels, int n_levels, int degree, const double *b, double *x, double *r,
                      double *z, double *p, double *ws, void *stream);
int fea_pcg_iterate_amg(int mat_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                        const int32_t *rowptr, const int32_t *col, const void *val, const float *binv, double rho0,
                        const fea_amg_level *levels, int n_levels, int degree, double *x, double *r, double *z,
                        double *p, double *q, double *ws, double tol, int max_iter, int iters,
                        double *status_host, void *stream);

/* One SpMV stage of the PCG iteration alone: q = A p and the p.q partials
 * (the CSR `A @ p` of _linear_solver.py:32 fused with the dot of :33); a no-op
 * when ws marks the solve converged. */
int fea_pcg_spmv(int mat_dtype, int vec_dtype, int n, int64_t nnz, int n_groups, const int32_t *grp,
                 const int32_t *rowptr, const int32_t *col, const void *val, const void *p, void *q,
                 double *ws, void *stream);

/* ---------------------------------------------------------------- batched PCG --
 * The Jacobi-PCG above over a batch of independent systems, one launch per CG
 * stage for the whole batch (the solve server's many concurrent static solves;
 * the reference solves one system at a time, _linear_solver.py:4-42).  No
 * counterpart in the reference: it replaces running conjugate_gradient once per
 * request.  Each system brings its own CSR (node row groups required, n_groups
 * as for fea_spmv), diag_slot, vectors (vec_dtype) and ws (fea_pcg_ws_size
 * doubles); dinv is written by fea_pcg_batch_start.  A system's iterates do not
 * depend on the other members of the batch (bit-identical alone and batched).
 *
 * fea_pcg_batch_layout (host): fills spmv_blocks / vec_blocks of every system and
 * the prefix tables spmv_start[n_sys+1], vec_start[n_sys+1]; the caller copies
 * the system array and both tables to the device and passes the totals
 * spmv_start[n_sys], vec_start[n_sys].  The same device copies serve every
 * start / iterate call until the batch changes.
 * fea_pcg_batch_start: dinv = 1/|diag A| (or the node-block inverses, precond 1),
 * r = b - A x (x = warm start), z, p,
 * reset of the iteration count and convergence flag.
 * fea_pcg_batch_iterate: `iters` iterations of every unconverged system; with
 * status != NULL (device, 2*n_sys doubles) writes {iterations, done} per system.
 * Enqueues kernels only (no host synchronisation, no memset): capturable in a
 * HIP graph. */
#define FEA_PCG_BATCH_MAX 1024
typedef struct fea_pcg_system {
  int32_t n;           /* rows */
  int32_t n_groups;    /* node row groups (< 0: triple-structured, see fea_spmv) */
  int32_t spmv_blocks; /* set by fea_pcg_batch_layout */
  int32_t vec_blocks;  /* set by fea_pcg_batch_layout */
  int32_t max_iter;
  int32_t precond;     /* 0: point Jacobi (dinv: vec_dtype[n]); 1: node-block Jacobi (dinv:
                          float[6 |n_groups|], written by fea_pcg_batch_start) */
  int64_t nnz;
  double tol;          /* stop when r.z / b.(M^-1 b) < tol */
  const int32_t *grp, *rowptr, *col, *diag_slot;
  const void *val;     /* mat_dtype */
  void *dinv;
  const void *b;
  void *x, *r, *z, *p, *q;
  double *ws;
} fea_pcg_system;
/* sizeof(fea_pcg_system), for bindings that mirror the struct */
int fea_pcg_system_size(void);
int fea_pcg_batch_layout(int n_sys, fea_pcg_system *sys_host, int32_t *spmv_start_host,
                         int32_t *vec_start_host);
int fea_pcg_batch_start(int mat_dtype, int vec_dtype, int n_sys, const fea_pcg_system *sys,
                        const int32_t *spmv_start, const int32_t *vec_start, int spmv_total,
                        int vec_total, void *stream);
int fea_pcg_batch_iterate(int mat_dtype, int vec_dtype, int n_sys, const fea_pcg_system *sys,
                          const int32_t *spmv_start, const int32_t *vec_start, int spmv_total,
                          int vec_total, int iters, double *status, void *stream);
/* fea_pcg_batch_iterate with the batch's counts read on the device: hdr[3] =
 * {n_sys, spmv_start[n_sys], vec_start[n_sys]} (device int32, written with the
 * tables), and fixed grids of spmv_grid / vec_grid workgroups, each looping over the
 * stage's virtual workgroups (same results as the exact grids).  Its launches do not
 * depend on the batch layout, so one captured HIP graph of a chunk serves every batch
 * the tables describe; status receives 2*n_sys doubles. */
int fea_pcg_batch_iterate_dev(int mat_dtype, int vec_dtype, const int32_t *hdr,
                              const fea_pcg_system *sys, const int32_t *spmv_start,
                              const int32_t *vec_start, int spmv_grid, int vec_grid, int iters,
                              double *status, void *stream);
/* The SpMV stage of one batched iteration alone (q = A p and the p.q partials of
 * every unconverged system), for timing. */
int fea_pcg_batch_spmv(int mat_dtype, int vec_dtype, int n_sys, const fea_pcg_system *sys,
                       const int32_t *spmv_start, int spmv_total, void *stream);

/* FP64 residual of an FP32-stored matrix: r64 = b64 - A32 x64 (iterative
 * refinement outer loop), plus r64 . r64 into `partial` (fea_vec_blocks(n)
 * per-block sums). */
int fea_spmv_residual_mixed(int n, const int32_t *rowptr, const int32_t *col, const float *val,
                            const double *x, const double *b, double *r, double *partial,
                            void *stream);
int fea_vec_blocks(int n);

/* Vector helpers on dof-mapped state.
 * fea_dof_axpy: U[node][c] = U0[node][c] + alpha * d[dof_map[3 node + c]] for
 * free DOFs, U[..] = U0[..] otherwise (_GC2RGC of assembly.py:470-492 fused
 * with the line-search update GC0 + alpha dGC, static_implicit.py:126). */
int fea_dof_axpy(int n_nodes, const int32_t *dof_map, const double *U0, double alpha,
                 const double *d, double *U, void *stream);
/* y(dtype_y) = x(dtype_x) elementwise conversion */
int fea_convert(int n, int dtype_x, const void *x, int dtype_y, void *y, void *stream);
/* y(float64) += alpha * x(dtype_x) */
int fea_axpy(int n, double alpha, int dtype_x, const void *x, double *y, void *stream);
/* per block (sum x^2, max |x|) pairs into partial[2*fea_vec_blocks(n)]: the
 * norms of the residual used by the Newton convergence tests
 * (static_implicit.py:231, :290). */
int fea_vec_stats(int n, const double *x, double *partial, void *stream);
/* partial[blk] = sum a.b (float64 vectors), fea_vec_blocks(n) partials */
int fea_dot_partial(int n, const double *a, const double *b, double *partial, void *stream);

/* ---------------------------------------------------------------- explicit --
 * Fused force gather + central-difference update (the loop body of
 * DynamicExplicitSolver.solve, dynamic_explicit.py:100-114): per free DOF
 * R = sum of staged forces - fc (an optional nodal force [3N], e.g. contact),
 * a = -minv R, vh += dt_v a, U += dt_u vh (nodal).  The first step uses
 * dt_v = dt/2 (v_{1/2} = v_0 + dt/2 a_0), later ones dt_v = dt. */
int fea_explicit_update(int n_nodes, const int32_t *rseg, const int32_t *dof_map,
                        const double *rstage, const double *minv, const double *fc, double dt_v,
                        double dt_u, double *a, double *vh, double *U, void *stream);

/* fea_explicit_update with the forces of fea_elem_residual_tiles: node A sums
 * partial[node_plist[s]*3 ..] for s in node_pptr[A] .. node_pptr[A+1] in float64, in
 * that order.  Replaces the same loop body of dynamic_explicit.py:100-114. */
int fea_explicit_update_tiles(int n_nodes, const int32_t *node_pptr, const int32_t *node_plist,
                              const int32_t *dof_map, const double *partial, const double *minv,
                              const double *fc, double dt_v, double dt_u, double *a, double *vh,
                              double *U, void *stream);

/* ---------------------------------------------------------------- contact --
 * Node-to-surface penalty contact on T3 master facets with a uniform-grid broad
 * phase (new law; the reference's is the Gauss-point-pair barrier with a
 * cKDTree broad phase of contact.py:86-122).  Candidate predicate, exactly:
 * c_f = (x_a + x_b + x_c)/3 and |x_s - c_f|^2 <= radius^2 in float64 without
 * fused multiply-add, so pair sets are reproducible bit for bit.
 *
 * fea_contact_grid_build: facet centroids and hashed cell keys (cell size
 * `cell`, 2^table_bits buckets, grid origin origin_host[3]) and a bucket sort
 * of the facets (ids ascending in each bucket): centroid[3F], key[F],
 * bucket_start[2^bits + 1], bucket_items[F], work[max(2^bits, F)] int32. */
int fea_contact_grid_build(int n_facets, const int32_t *conn, const double *X, const double *U,
                           const double *origin_host, double cell, int table_bits, double *centroid,
                           int32_t *key, int32_t *bucket_start, int32_t *bucket_items, int32_t *work,
                           void *stream);
/* Candidate facets of every slave node (radius <= cell, kmax <= 64): cand[S*kmax]
 * ascending facet ids, count[S] = number of candidates (may exceed kmax: the
 * caller must treat that as an error).  max_count (optional, one int32 the
 * caller zeroes) is raised to every count above kmax, so overflow can be
 * checked once after many searches. */
int fea_contact_search(int n_slave, const int32_t *slave, const double *X, const double *U,
                       const double *centroid, const int32_t *bucket_start,
                       const int32_t *bucket_items, const double *origin_host, double cell,
                       int table_bits, double radius, int kmax, int32_t *cand, int32_t *count,
                       int32_t *max_count, void *stream);
int fea_contact_blocks(int n_slave);
/* Penalty forces: per slave, the candidate facet it penetrates (gap g < 0,
 * g > -depth_max, projection inside the triangle) with the largest g (ties:
 * smallest facet id); F = -penalty A_s g n on the slave (A_s = slave_area[s],
 * the node's tributary area, or 1 if slave_area is null), -F times the barycentric
 * weights on the facet nodes.  fc[n_dof] is zeroed then filled; master
 * reactions are gathered per master node through the static node->facet
 * adjacency (nf_ptr/nf_facet/nf_vert) and a bucket sort of slaves by chosen
 * facet (sb_start[F+1], sb_items[S]).  energy_partial (optional) receives
 * per-block sums of penalty A_s g^2 / 2 (fea_contact_blocks(S) partials). */
int fea_contact_forces(int n_slave, const int32_t *slave, const double *X, const double *U,
                       const int32_t *conn, int n_facets, const int32_t *cand, const int32_t *count,
                       int kmax, double penalty, const double *slave_area, double depth_max, int n_master,
                       const int32_t *master_nodes, const int32_t *nf_ptr, const int32_t *nf_facet,
                       const int32_t *nf_vert, int32_t *chosen, double *force, double *weight,
                       int32_t *sb_start, int32_t *sb_items, int32_t *work, int64_t n_dof, double *fc,
                       double *energy_partial, void *stream);

/* Node-to-surface contact for explicit steps with reused candidate lists (it stands
 * for the contact force the reference's explicit loop would take from its broken
 * force-only barrier branch, contact.py:658-659).  fea_contact_rebuild searches with
 * the radius radius + skin (grid build and search as above, cell >= radius + skin),
 * stores each slave's candidates (cand/count), facet -> slave lists of those pairs
 * (inv_start[F + 1], inv_items[S*kmax], slaves ascending), per master node the
 * pairs over its facets (m_ptr[n_master + 1], m_items[2 * 3*S*kmax]: slave, facet*4 +
 * local vertex; facet order, slaves ascending), the contact node positions x0 and
 * clears dmax.  fea_contact_eval, per step: dmax = the largest squared
 * displacement of a contact node since the rebuild; then, while dmax <= (skin/2)^2 and
 * no stored list overflowed kmax (sup_max), every exact candidate (the predicate of
 * fea_contact_search with `radius`) is in the stored lists, which the penalty and the
 * master reactions use; otherwise they scan every facet / slave.  Either way the
 * result is that of fea_contact_search + fea_contact_forces at U (choice ties go to
 * the smaller facet id, as in the sorted candidate list), with fc written only at the
 * contact nodes (fc[node] = master reaction + slave force; the caller zeroes fc once)
 * and max_count raised by exact candidate counts above kmax.  The host rebuilds when
 * dmax approaches (skin/2)^2; a late rebuild costs time, not results.
 *   cnode[n_cnode]: every slave and master node once; cn_slave[i] / cn_master[i]:
 *   its slave index / master index (into nf_ptr) or -1. */
/* dmax is kept in 64 slots a cache line apart (contact node i raises slot i % 64), so that
 * the per-step tracking's atomics do not all meet at one address. */
#define FEA_CONTACT_DMAX_SLOTS 64
#define FEA_CONTACT_DMAX_STRIDE 16
typedef struct fea_contact_set {
  int32_t n_slave, n_facets, n_cnode, kmax, table_bits, n_master;
  double cell, radius, skin, penalty, depth_max;
  double origin[3];
  const int32_t *slave, *conn, *cnode, *cn_slave, *cn_master, *nf_ptr, *nf_facet, *nf_vert;
  const double *slave_area;                  /* may be NULL (area 1) */
  double *centroid, *force, *weight, *x0;    /* [3F], [3S], [3S], [3*n_cnode] */
  uint64_t *dmax;  /* [FEA_CONTACT_DMAX_SLOTS * FEA_CONTACT_DMAX_STRIDE]: slot k at
                      dmax[k * FEA_CONTACT_DMAX_STRIDE]; the tracked value is the slots' max */
  int32_t *key, *bucket_start, *bucket_items, *work;   /* [F], [2^bits + 1], [F], [max(2^bits, F)] */
  int32_t *cand, *count, *max_count, *chosen;          /* [S*kmax], [S], [1] (may be NULL), [S] */
  int32_t *sup_max, *inv_start, *inv_items;            /* [1], [F + 1], [S*kmax] */
  int32_t *m_ptr, *m_items;                            /* [n_master + 1], [6*S*kmax] */
  int32_t *lists_ok;  /* [1]: validity of the stored lists at the last penalty launch (may be
                         NULL for fea_contact_eval; the explicit loop's fused launches need it) */
} fea_contact_set;
int fea_contact_set_size(void);
int fea_contact_rebuild(const fea_contact_set *set_host, const double *X, const double *U, void *stream);
int fea_contact_eval(const fea_contact_set *set_host, const double *X, const double *U, double *fc,
                     double *energy_partial, void *stream);
/* dmax = max(dmax, the largest squared displacement of a contact node from x0 at U): the
 * first launch of fea_contact_eval, alone (before the explicit loop's fused launches at a U
 * they did not produce). */
int fea_contact_track(const fea_contact_set *set_host, const double *X, const double *U, void *stream);

/* The explicit loop's step in two launches (no counterpart in the reference; they replace
 * fea_elem_residual_tiles + fea_contact_eval + fea_explicit_update_tiles of one step, with
 * the same results bit for bit).
 * fea_elem_residual_tiles_contact: fea_elem_residual_tiles (no node staging) and, in further
 * workgroups of the same launch, the penalty of fea_contact_eval at U on the stored lists
 * (or every facet when dmax > (skin/2)^2 or a list overflowed), which stores the lists'
 * validity in lists_ok.  dmax must hold the displacement at U (fea_contact_rebuild or
 * fea_contact_track at U, or the previous fea_explicit_update_tiles_contact).
 * fea_explicit_update_tiles_contact: fea_explicit_update_tiles with the contact force of
 * fea_contact_eval (the master reactions and slave force of each contact node, from the
 * penalty launch's choice and lists_ok), written to fc at the contact nodes; the nodes with
 * bit A % 32 of cmask[A / 32] set (exactly the set's contact nodes cnode) are updated by the
 * contact workgroups, which then raise dmax by each contact node's displacement at the new U.
 * minv_node[A] is 1/m of node A (its free DOFs share the lumped mass); a may be NULL (the
 * accelerations are not stored). */
int fea_elem_residual_tiles_contact(int elem_type, int dtype, int n_elem, const int32_t *conn,
                                    const double *X, const double *U, const double *params_host,
                                    const int32_t *tile_lbase, const int32_t *pslot_ptr, const uint16_t *pslot,
                                    double *partial, const fea_contact_set *set_host, void *stream);
int fea_explicit_update_tiles_contact(int n_nodes, const int32_t *node_pptr, const int32_t *node_plist,
                                      const int32_t *dof_map, const double *partial, const double *minv_node,
                                      const uint32_t *cmask, const fea_contact_set *set_host, const double *X,
                                      double *fc, double dt_v, double dt_u, double *a, double *vh, double *U,
                                      void *stream);

/* ---------------------------------------------------------------- adjoint --
 * Vector-Jacobian products for sensitivities.  The reference differentiates by
 * torch autograd through its assembly (C3base.py:387-452, pressure.py:80-132
 * under requires_grad); these replace that path.
 *
 * fea_elem_adjoint: for phi = lam . f_int(X, U; params) of one element group,
 * writes per (element, node) rows of d phi/dX (gx_stage) and d phi/dU
 * (gu_stage, i.e. K_int lam) through dst (slot per element node, -1 = skip;
 * rows of 3 doubles), and per-block partial sums of d phi/d{w1, w2, kv} into
 * gp_partial[k * fea_elem_adjoint_blocks(n) + block].  Any output may be null
 * (not all).  Float64 throughout; Gauss-point state is recomputed. */
int fea_elem_adjoint_blocks(int n_elem);
int fea_elem_adjoint(int elem_type, int n_elem, const int32_t *conn, const double *X, const double *U,
                     const double *params_host, const double *lam, const int32_t *dst, double *gx_stage,
                     double *gu_stage, double *gp_partial, void *stream);
/* Pressure tangent times lam without the matrix: rows of scale * Hess(V)(X+U) lam
 * per (facet, node) through dst.  V is the enclosed-volume functional of
 * fea_surface_pressure, so this is also d(scale dV/dx . lam)/dX. */
int fea_surface_pressure_hvp(int surf_type, int n_facets, const int32_t *conn, const double *X,
                             const double *U, double scale, const double *lam, const int32_t *dst,
                             double *stage, void *stream);
/* out[3A + j] = sum of stage rows seg[A] .. seg[A+1]-1, every node (no DOF map). */
int fea_gather_nodal(int n_nodes, const int32_t *seg, const double *stage, double *out, void *stream);

/* ---------------------------------------------------------------- couple --
 * Kinematic coupling of node sets to reference points (constraints/couple.py:7-292,
 * reference_points.py:5-24).  A reference point owns 6 generalized coordinates
 * (v, z): translation and rotation vector; a coupled node b follows
 * u_b = v + R(z) y_b - y_b, y_b = X_b - X_rp (Couple.modify_RGC, couple.py:36-43).
 *
 * The state vector is [3 n_state_nodes nodal displacements | 6 per reference point].
 * The element kernels assemble an *extended* system in which the coupled node DOFs
 * are kept as rows and columns (CSR "ext"); fea_couple_condense forms the reduced
 * system of the generalized coordinates (free node DOFs, then free reference-point
 * DOFs, the reference's GC order) by the chain rule:
 *   R_red = T^T R_ext - load_scale * (F, M),  K_red = T^T K_ext T + sum_b R_b d2u_b/dz2,
 * T = du/dGC (Couple.modify_R_K, couple.py:66-184, and the reduction of
 * _assemble_reduced_Matrix, assembly.py:357-387).  The (reference point, row)
 * entries are written as the transposes of the (row, reference point) entries,
 * as the reference does (K_ext symmetric).  Every output entry has one writer and
 * every sum a fixed order: results are deterministic.
 *
 * All arrays are device arrays built once per model by the host (fea_amd.couple). */
typedef struct fea_couple_map {
  int32_t n_couples;      /* couples, one reference point each */
  int32_t n_cnodes;       /* coupled nodes, couples contiguous: couple k owns [cnode_ptr[k], cnode_ptr[k+1]) */
  int32_t n_pairs;        /* (reduced node row, couple) pairs whose ext row has a coupled column */
  int32_t n_red;          /* reduced rows */
  int32_t n_red_nodes;    /* reduced rows that are node DOFs (the first n_red_nodes) */
  int32_t n_state_nodes;  /* nodes of the nodal part of the state vector */
  int64_t nnz_red;        /* entries of the reduced CSR */
  const int32_t *cnode_ptr; /* [n_couples + 1] */
  const int32_t *rp_state;  /* [n_couples] offset of the couple's (v, z) in the state vector */
  const int32_t *cnode;     /* [n_cnodes] node id */
  const double *y;          /* [n_cnodes * 3] X_b - X_rp */
  const int32_t *crow_ext;  /* [3 n_cnodes] ext row of coupled DOF 3b + c */
  const int32_t *cr_ptr;    /* [3 n_cnodes + 1] coupled-column entries of each coupled row */
  const int32_t *ce_slot;   /* ext CSR slot of each such entry */
  const int32_t *ce_dof;    /* its column as a coupled DOF index 3b' + c' */
  const int32_t *pair_ptr;  /* [n_pairs + 1] coupled-column entries of each pair */
  const int32_t *pe_slot, *pe_dof;
  const int32_t *pair_row_slot; /* [n_pairs * 6] reduced slot (row, point DOF d), -1 if d is fixed */
  const int32_t *pair_col_slot; /* [n_pairs * 6] reduced slot (point DOF d, row) */
  const int32_t *rp_slot;   /* [n_couples * 36] reduced slot of (point DOF d, point DOF e) or -1 */
  const int32_t *rp_row;    /* [n_couples * 6] reduced row of point DOF d or -1 */
  const double *rp_load;    /* [n_couples * 6] force (3), moment (3) on the point */
  const int32_t *ff_src;    /* [nnz_red] ext slot of a node-node entry, -1 for the others */
  const int32_t *rf_src;    /* [n_red_nodes] ext row of a reduced node row */
  double *ydot;             /* workspace [n_cnodes * 9]: d u_bc / d z_p at [(3b + c) * 3 + p] */
  double *ydot2;            /* workspace [n_cnodes * 27]: d2 u_bc / d z_p d z_q */
  double *partial;          /* workspace [3 n_cnodes * 42] */
} fea_couple_map;
int fea_couple_map_size(void);
/* Coupled node displacements from the reference points' (v, z) in `state`
 * (Couple.modify_RGC, couple.py:36-43 with _rotation3d, :272-292) and, with
 * with_derivs, their first and second derivatives in z into m->ydot / m->ydot2
 * (Couple._calculate_Ydotz, couple.py:186-270).  Rodrigues in the form
 * R(z) y - y = A(s) z x y + B(s) z x (z x y), s = |z|^2, A = sin t / t,
 * B = (1 - cos t) / t^2, with series of A, B and their s-derivatives below
 * s = 2.5, so every quantity is exact to rounding at z = 0 too. */
int fea_couple_kinematics(const fea_couple_map *m, double *state, int with_derivs, void *stream);
/* Reduced residual (and tangent when K_ext != NULL; val_red in val_prec) from the
 * extended ones.  Needs the derivatives of the last fea_couple_kinematics call. */
int fea_couple_condense(const fea_couple_map *m, int val_prec, const double *K_ext, const double *R_ext,
                        double load_scale, void *val_red, double *R_red, void *stream);

/* ---------------------------------------------------------------- barrier --
 * The reference's Gauss-point-pair contact barrier on T3 facets (ContactSelf /
 * Contact, loads/contact.py:86-122, :300-711, :771-1207).  For facet pairs whose
 * centroids lie within `radius` at the current configuration (FMA-free float64
 * centroid ((a + b) + c) / 3 and squared distance, as fea_contact_search), and each
 * of their 3 x 3 Gauss pairs (points (1/6,1/6), (2/3,1/6), (1/6,2/3), weight 1/6):
 *   phi = ratio w_p w_q c_g exp(k_g D) S((s_f - M)/(s_f - e_f)) S((h_t - L)/(r_h h_t)),
 * D = (n_p - n_q).(y_g - y_G)/2, M = n_p.n_q, L = |y_g - y_G|, S the clamped quintic
 * smoothstep, w = |X_xi x X_eta|/6 from the reference coordinates; ratio = 1 for two
 * surfaces, and for one surface (self_contact) the initial-geometry factor
 * S((-(n0_g - n0_G).dy0/|dy0| - ign_min)/(ign_max - ign_min)) of contact.py:312-331.
 * The barrier energy is + sum phi in the total potential energy, its gradient adds to
 * the residual and its Hessian to the tangent. */
typedef struct fea_barrier_params {
  double c_g, k_g;        /* penalty_distance_g (1e-5), penalty_factor_g (20) */
  double h_t, r_h;        /* penalty_threshold_h (1.5), penalty_ratio_h (0.5) */
  double s_f, e_f;        /* penalty_start_f (-0.0), penalty_end_f (-0.8) */
  double ign_min, ign_max;/* ignore_min_normal (0.8), ignore_max_normal (1.2), self contact */
  double radius;          /* candidate radius: search ratio (2.0) x h_t */
  int32_t self_contact;   /* 1: one surface (pairs j > i, ratio factor), 0: two surfaces */
  int32_t reserved;
} fea_barrier_params;
/* Candidate pairs (i, j): query facets qconn [n_query x 3] against the facets of a grid
 * built by fea_contact_grid_build (centroids, buckets; cell >= radius).  With
 * self_contact only j > i.  Count pass (offsets == NULL): count[i].  Fill pass:
 * pair_i/pair_j at offsets[i] (exclusive scan of the counts), deterministic order. */
int fea_barrier_search(int n_query, const int32_t *qconn, const double *X, const double *U, const double *centroid,
                       const int32_t *bucket_start, const int32_t *bucket_items, const double *origin_host,
                       double cell, int table_bits, double radius, int self_contact, const int32_t *offsets,
                       int32_t *count, int32_t *pair_i, int32_t *pair_j, void *stream);
/* out[0..n] = exclusive prefix sum of in[0..n) (out[n] = total), one workgroup. */
int fea_scan_i32(int n, const int32_t *in, int32_t *out, void *stream);
/* flag[p] = 1 if some Gauss pair of pair p has a nonzero ratio (the pairs ContactSelf keeps). */
int fea_barrier_active(const fea_barrier_params *P, int n_pairs, const int32_t *pair_i, const int32_t *pair_j,
                       const int32_t *conn1, const int32_t *conn2, const double *X, int32_t *flag, void *stream);
/* *missing += number of pairs with a nonzero ratio whose key i * stride + j is not in
 * keys[n_keys] (ascending): the pairs a staged pattern does not cover. */
int fea_barrier_missing(const fea_barrier_params *P, int n_pairs, const int32_t *pair_i, const int32_t *pair_j,
                        const int32_t *conn1, const int32_t *conn2, const double *X, const int64_t *keys,
                        int n_keys, int64_t stride, int32_t *missing, void *stream);
int fea_barrier_energy_blocks(int n_pairs);
/* per-block partial sums of phi over the listed pairs (all treated as candidates) */
int fea_barrier_energy(const fea_barrier_params *P, int n_pairs, const int32_t *pair_i, const int32_t *pair_j,
                       const int32_t *conn1, const int32_t *conn2, const double *X, const double *U,
                       double *partial, void *stream);
/* Gradient rows (dst_r: slot per pair node, 6 per pair, nodes = conn1[i] then conn2[j])
 * and Hessian 3x3 blocks (dst_k: 36 per pair, row-major node pairs; stage_prec) of every
 * listed pair that is a candidate at U; the slots of the others are zeroed.  One 64-lane
 * workgroup per pair.  Either output may be NULL. */
int fea_barrier_assemble(const fea_barrier_params *P, int stage_prec, int n_pairs, const int32_t *pair_i,
                         const int32_t *pair_j, const int32_t *conn1, const int32_t *conn2, const double *X,
                         const double *U, const int32_t *dst_k, void *kstage, const int32_t *dst_r, double *rstage,
                         void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FEA_AMD_H */
