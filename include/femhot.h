/*
 * femhot -- MI355X (gfx950) finite-element residual hot path, C ABI.
 *
 * Plain C interface: opaque context handle, raw device pointers, element counts,
 * hipStream_t passed as void*.  No torch / XLA types cross this boundary.  Every
 * call is stream-ordered on the caller's stream; the library allocates device
 * memory only in femhot_ctx_create / femhot_ctx_destroy and the lazily built
 * context caches (pattern, slot table, colouring, chunk plan, solver workspace),
 * which are created on the first call that needs them.
 *
 * Reference interfaces replaced (folax, paths relative to the repository root):
 *  - fol/loss_functions/fe_loss.py:57-122   FiniteElementLoss.Initialize  -> femhot_ctx_create
 *  - fol/loss_functions/fe_loss.py:166-176  ComputeElementsEnergies / ComputeTotalEnergy
 *                                           -> femhot_residual (energy output)
 *  - fol/loss_functions/fe_loss.py:250-262  ComputeBatchLoss (per-sample energies, B >= 1)
 *                                           -> femhot_residual / femhot_energy_grad with batch
 *  - fol/loss_functions/fe_loss.py:264-318  ComputeJacobianMatrixAndResidualVector
 *                                           -> femhot_pattern + femhot_tangent_csr
 *  - fol/loss_functions/ffi_functions/kr_small_displacement_element.cc:178-224,226-291
 *      (XLA FFI handlers compute_nodal_residuals / compute_elements)
 *                                           -> femhot_residual / femhot_tangent_csr
 *  - fol/responses/fe_response.py:434-526   ComputeLossElementControlGrad / adjoint nodal control
 *                                           derivatives (lambda^T dR/dk) -> femhot_residual_vjp
 *  - fol/responses/fe_response.py:312-332,365-406 adjoint nodal shape derivatives (lambda^T dR/dX)
 *                                           -> femhot_residual_shape_vjp
 *  - jax.grad of fe_loss.py:175 ComputeTotalEnergy (VJP of the energy) -> femhot_energy_grad
 *  - fol/solvers/fe_solver.py:60-78         LinearSolve (BiCGSTAB / direct on the assembled
 *                                           tangent) -> femhot_pcg (Jacobi-PCG, FP64 or FP32
 *                                           inner iterations with FP64 refinement)
 *  - fol/solvers/fe_nonlinear_residual_based_solver.py:76-170  load-stepped Newton Solve
 *                                           -> femhot_newton
 *  - fol/mesh_input_output/mesh.py:123-142  CheckAndOrientElements -> femhot_orient_elements
 */
#ifndef FEMHOT_H
#define FEMHOT_H

#ifdef __cplusplus
extern "C" {
#endif

/* element types (fol/geometries/__init__.py:18-21 names: "quad", "hexahedron", "tetra") */
#define FEMHOT_QUAD4 0
#define FEMHOT_HEX8 1
#define FEMHOT_TET4 2

/* physics (ThermalLoss thermal.py:15, MechanicalLoss mechanical.py:11,
 * NeoHookeMechanicalLoss mechanical_neohooke.py:16) */
#define FEMHOT_THERMAL 0
#define FEMHOT_MECHANICAL 1
#define FEMHOT_NEOHOOKE 2

/* arithmetic type of a call */
#define FEMHOT_F32 0
#define FEMHOT_F64 1

/* call flags */
#define FEMHOT_MASK_DIRICHLET 0x1u  /* zero Dirichlet dofs of residual / g_u (fe_loss.py:191-207) */
#define FEMHOT_ELEMENT_FIELD 0x2u   /* material/control field given per element, not per node */
#define FEMHOT_PICARD 0x4u          /* thermal VJP with k(T) frozen: thermal.py:31 stop_gradient */
#define FEMHOT_TRANSPOSE 0x8u       /* tangent: transpose element matrices before masking (:219-228) */
#define FEMHOT_ACCUMULATE 0x10u     /* add into outputs instead of overwriting them */
#define FEMHOT_SCATTER_COLOUR 0x20u /* colour-ordered, non-atomic, deterministic scatter */
#define FEMHOT_SCATTER_CHUNK 0x40u  /* LDS-staged chunk scatter (persistent launch) */
#define FEMHOT_SCATTER_ROWS 0x80u   /* tangent: row-owner assembly, each CSR value summed in a fixed
                                       order and written once (thermal, linear elastic) */

/* status codes */
#define FEMHOT_OK 0
#define FEMHOT_ERR_ARG 1          /* bad argument (null pointer, unsupported enum, ...) */
#define FEMHOT_ERR_SHAPE 2        /* sizes inconsistent */
#define FEMHOT_ERR_INDEX 3        /* connectivity / dof index out of range */
#define FEMHOT_ERR_NONFINITE 4    /* non-finite input where finite is required */
#define FEMHOT_ERR_HIP 5          /* HIP runtime error (message has the HIP error string) */
#define FEMHOT_ERR_NOT_CONVERGED 6
#define FEMHOT_ERR_UNSUPPORTED 7  /* combination not implemented */

typedef struct femhot_ctx femhot_ctx;

/* Mesh + loss settings.  All pointers are HOST pointers, read during create only. */
typedef struct {
  int elem_type;       /* FEMHOT_QUAD4 / FEMHOT_HEX8 / FEMHOT_TET4 */
  int physics;         /* FEMHOT_THERMAL / FEMHOT_MECHANICAL / FEMHOT_NEOHOOKE */
  int num_gp;          /* Gauss rule 1/2/3 (fe_loss.py:71-84); 0 = the reference default */
  double young_modulus;   /* material_dict E (mechanical: scales D; unused by neohooke) */
  double poisson_ratio;   /* material_dict nu */
  double beta, c;         /* thermal k(T) = k0 (1 + beta T^c) (thermal.py:22-26) */
  double body_force[3];   /* mechanical body_foce / thermal scalar source in [0] */
  long long n_nodes;
  long long n_elems;
  const double *coords;   /* (n_nodes, 3) row-major */
  const int *conn;        /* (n_elems, nodes_per_elem) row-major, int32 */
  long long n_dirichlet;
  const long long *dirichlet_idx; /* global dof ids ndpn*node+comp (fe_loss.py:34-55) */
  const double *dirichlet_val;
  int n_segments;                  /* >1: disjoint meshes, energies reported per segment */
  const long long *segment_offsets; /* (n_segments+1) element offsets, or NULL */
} femhot_mesh_desc;

const char *femhot_last_error(void);
const char *femhot_version(void);

int femhot_ctx_create(const femhot_mesh_desc *desc, femhot_ctx **out);
void femhot_ctx_destroy(femhot_ctx *ctx);

/* sizes: dofs, nodes per element, dofs per node, nnz of the tangent, colours, segments */
int femhot_ctx_info(const femhot_ctx *ctx, long long *n_dofs, int *nodes_per_elem, int *dofs_per_node,
                    long long *nnz, int *n_colours, int *n_segments);

/* CSR pattern of the tangent (scipy csr_array of the reference COO: sorted columns, summed
 * duplicates, explicit zeros kept).  row_ptr: (n_dofs+1) int64, col_idx: (nnz) int32, DEVICE. */
int femhot_pattern(femhot_ctx *ctx, long long *row_ptr, int *col_idx, void *stream);

/* greedy first-fit element colouring in element order (HOST output, n_elems int32). */
int femhot_colouring(femhot_ctx *ctx, int *colour);

/* Residual + element energies.  field: (batch, n_nodes) or (batch, n_elems) with
 * FEMHOT_ELEMENT_FIELD; u: (batch, n_dofs); R: (batch, n_dofs) or NULL; energy: double,
 * (batch * max(1,n_segments)) or NULL -- the reference element energies summed
 * (thermal T.r, mechanical u.r, neohooke Ee).  Strides in elements of the arithmetic type. */
int femhot_residual(femhot_ctx *ctx, int dtype, int batch, const void *field, long long field_stride,
                    const void *u, long long u_stride, void *R, long long R_stride, double *energy,
                    unsigned flags, void *stream);

/* Tangent values on the CSR pattern (+ optional residual R), BC-masked exactly like
 * fe_loss.py:191-207; FEMHOT_TRANSPOSE transposes each element matrix first.  Assembly: hardware
 * atomics into zero-filled values (no scatter flag), colour-ordered plain adds
 * (FEMHOT_SCATTER_COLOUR), or row-owner assembly (FEMHOT_SCATTER_ROWS: one workgroup per node sums
 * the node's rows from its incident elements in element order and writes each value once; no fill,
 * deterministic; thermal and linear elastic, Neo-Hookean falls back to colour-ordered adds);
 * FEMHOT_SCATTER_CHUNK takes the one femhot_autotune measured fastest (untuned: row-owner, for
 * Neo-Hookean colour-ordered). */
int femhot_tangent_csr(femhot_ctx *ctx, int dtype, const void *field, const void *u, void *vals, void *R,
                       unsigned flags, void *stream);

/* VJP of the residual: g_u = (dR/du)^T w, g_field = (dR/dfield)^T w (exact; FEMHOT_PICARD
 * freezes k(T) for thermal).  Outputs are overwritten unless FEMHOT_ACCUMULATE. */
int femhot_residual_vjp(femhot_ctx *ctx, int dtype, int batch, const void *field, long long field_stride,
                        const void *u, long long u_stride, const void *w, long long w_stride, void *g_u,
                        long long g_u_stride, void *g_field, long long g_field_stride, unsigned flags,
                        void *stream);

/* VJP of the residual with respect to the node coordinates (shape sensitivity): g_x = (dR/dX)^T w,
 * (n_nodes, 3) in the arithmetic type, overwritten; the z column is zero for 2-D elements.  Replaces
 * jax.jacrev(ComputeElement[1], argnums=0) contracted with the element adjoint and assembled per
 * node component (fol/responses/fe_response.py:312-332, 365-406).  One sample; FEMHOT_ELEMENT_FIELD
 * is honoured, other flags are ignored. */
int femhot_residual_shape_vjp(femhot_ctx *ctx, int dtype, const void *field, const void *u, const void *w,
                              void *g_x, unsigned flags, void *stream);

/* femhot_energy_grad that allocates nothing: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as femhot_energy_grad on the LDS-staged chunk path (the
 * configuration tuned or set for this call shape, else the default one; FEMHOT_SCATTER_* flags are
 * ignored), but its scratch -- the shared-node partials and per-chunk energies -- lives in a fixed
 * 64 MiB device array of the library, one per GPU, held until the process ends, instead of the
 * context's growable workspace.  Requirements and limits:
 *  - the chunk plan of the call shape must already be built (femhot_set_chunk_config, or one
 *    femhot_energy_grad call with FEMHOT_SCATTER_CHUNK), else FEMHOT_ERR_ARG;
 *  - a call whose scratch (batch x shared-node slots x (dofs per node [+1]) x element size + batch x
 *    chunks x 8 bytes) exceeds 64 MiB returns FEMHOT_ERR_SHAPE;
 *  - every call on one GPU shares that GPU's array: calls that overlap in time on one GPU (two
 *    streams, two threads) race and must be serialised by the caller.  Nothing in the array needs
 *    to start at zero: each call writes every slot it reads. */
int femhot_energy_grad_noalloc(femhot_ctx *ctx, int dtype, int batch, const void *field, long long field_stride,
                               const void *u, long long u_stride, double *energy, void *g_u, long long g_u_stride,
                               void *g_field, long long g_field_stride, unsigned flags, void *stream);

/* Integration data of every element at the context's Gauss rule (G points per element: QUAD4 g^2,
 * HEX8 g^3, TET4 1/4/8 for rule g = 1/2/3), DEVICE outputs in the arithmetic type, each optional:
 * N (G, nodes_per_elem) shape function values; wdetJ (n_elems, G) weight x signed detJ;
 * dNdx (n_elems, G, nodes_per_elem, dim) global shape gradients.  Replaces the per-Gauss-point
 * geometry that FiniteElementResponse.ComputeResponseElementValue evaluates before its response
 * formula (fol/responses/fe_response.py:92-113): the response integrand and its derivatives are
 * formed from these (d detJ / dX_a = detJ dN_a/dx gives the integrand's shape gradient). */
int femhot_gauss_point_data(femhot_ctx *ctx, int dtype, void *N, void *wdetJ, void *dNdx, void *stream);

/* Potential Pi and its gradient: g_u = dPi/du (= residual), g_field = dPi/dfield.
 * energy: double (batch * max(1,n_segments)) or NULL. */
int femhot_energy_grad(femhot_ctx *ctx, int dtype, int batch, const void *field, long long field_stride,
                       const void *u, long long u_stride, double *energy, void *g_u, long long g_u_stride,
                       void *g_field, long long g_field_stride, unsigned flags, void *stream);

/* Jacobi-preconditioned CG on the CSR tangent for K x = b (Dirichlet rows of b must be 0).
 * mode FEMHOT_F64: FP64 PCG.  mode FEMHOT_F32: FP32 matrix/vectors inside, FP64 dot products,
 * FP64 outer iterative refinement with the FP64 matrix.  x: in = initial guess, out = solution.
 * mode | FEMHOT_PCG_CONTEXT_PATTERN declares that row_ptr/col_idx are this context's pattern
 * (femhot_pattern): with 2 or 3 dofs per node the SpMV then reads one column index per node
 * block instead of one per nonzero.  Without it any CSR matrix of n_dofs rows is accepted.
 * Returns FEMHOT_ERR_NOT_CONVERGED (x still written) when maxiter is reached. */
#define FEMHOT_PCG_CONTEXT_PATTERN 0x100
int femhot_pcg(femhot_ctx *ctx, int mode, const long long *row_ptr, const int *col_idx, const double *vals,
               const double *b, double *x, double rel_tol, int maxiter, int *iters, double *rel_res,
               void *stream);

/* femhot_pcg in FP64 with its convergence test on the device: the host enqueues exactly `iterations`
 * iterations and never waits; once ||r|| <= rel_tol ||b|| (or the residual is not finite) the remaining
 * iterations' kernels return at once.  iters (int), rel_res (double) and status (int: 1 converged,
 * 2 non-finite residual, 3 iterations spent) are DEVICE pointers (any may be NULL), written at the end
 * of the stream's work.  Allocates nothing once the context's solver workspace is sized (its first
 * uncaptured call, or a femhot_pcg call), so a whole solve can be captured into a HIP graph.  mode:
 * FEMHOT_F64 (| FEMHOT_PCG_CONTEXT_PATTERN); the context's pattern must be built.  No counterpart in the
 * reference, whose iterative solve returns to the host (fe_solver.py:60-78); for callers that capture
 * solves or keep the host out of the loop. */
int femhot_pcg_async(femhot_ctx *ctx, int mode, const long long *row_ptr, const int *col_idx, const double *vals,
                     const double *b, double *x, double rel_tol, int iterations, int *iters, double *rel_res,
                     int *status, void *stream);

/* Load-stepped Newton-Raphson solve of the BC-masked residual R(u; field) = 0, replacing
 * FiniteElementNonLinearResidualBasedSolver.Solve (fol/solvers/fe_nonlinear_residual_based_solver.py:
 * 76-170).  For load step s = 1..load_steps the Dirichlet dofs are set to (s / load_steps) x their
 * value (fe_loss.py:91-92); each iteration assembles the FP64 CSR tangent and residual
 * (femhot_tangent_csr), takes ||R||_2 (a NaN norm stops the solve with FEMHOT_ERR_NONFINITE, the
 * reference's NaN guard, :118-132), solves K du = -R with femhot_pcg (mode FEMHOT_F64 or FEMHOT_F32;
 * a solve that reaches lin_maxiter continues with its last iterate, as the reference's iterative
 * solver does), and stops when ||R|| < abs_tol, ||du|| < rel_tol or after max_iter iterations,
 * else u += du (:135-159).  All vectors stay on the device; the host waits twice per iteration
 * (the two norms) besides the linear solver's own convergence checks.
 * field: FP64 DEVICE, (n_nodes) or (n_elems) with FEMHOT_ELEMENT_FIELD.  u: FP64 DEVICE (n_dofs),
 * initial guess in, solution out.  flags: FEMHOT_ELEMENT_FIELD | FEMHOT_SCATTER_COLOUR |
 * FEMHOT_SCATTER_ROWS | FEMHOT_SCATTER_CHUNK only: the tangent assembly, as for femhot_tangent_csr.
 * history: HOST (load_steps x max_iter x 2) doubles or NULL: (||R||, ||du||) per iteration, NaN
 * after a load step's last iteration.  steps: HOST (load_steps) ints or NULL: iterations taken per
 * load step.  Synchronises the stream. */
int femhot_newton(femhot_ctx *ctx, int mode, const double *field, double *u, unsigned flags, int load_steps,
                  int max_iter, double abs_tol, double rel_tol, double lin_tol, int lin_maxiter, double *history,
                  int *steps, void *stream);

/* femhot_newton with every decision taken on the device: the same load steps, assembly, norms and
 * stopping rules (fe_nonlinear_residual_based_solver.py:76-170), with the FP64 linear solve of
 * femhot_pcg_async (lin_iterations enqueued per Newton iteration; a solve that spends them continues
 * with its last iterate).  The host enqueues load_steps x max_iter iterations and never waits; once a
 * load step has ended its remaining iterations' updates and linear solves return at once (the tangent
 * and residual are still assembled, into the context's workspace), and once the solve has failed
 * every later kernel returns.  history: DEVICE (load_steps x max_iter x 2) doubles or NULL, (||R||,
 * ||du||) per iteration, NaN after a load step's last iteration.  steps: DEVICE (load_steps) ints or
 * NULL, iterations per load step (0 for steps after a failure).  status: DEVICE int or NULL,
 * FEMHOT_OK, or FEMHOT_ERR_NONFINITE when a residual norm was NaN or a linear solve non-finite (the
 * step's count is then the failing iteration).  The first call on a context must be uncaptured (it
 * builds the pattern and sizes the workspaces); later calls allocate nothing, so a whole nonlinear
 * solve can be captured into a HIP graph.  No counterpart in the reference, whose Newton loop
 * returns to the host each iteration; for callers that capture solves or keep the host out of the
 * loop. */
int femhot_newton_async(femhot_ctx *ctx, const double *field, double *u, unsigned flags, int load_steps, int max_iter,
                        double abs_tol, double rel_tol, double lin_tol, int lin_iterations, double *history,
                        int *steps, int *status, void *stream);

/* Autotuner of the scatter path that FEMHOT_SCATTER_CHUNK leaves to the library: times
 * elements-per-chunk x workgroup size x grid shape (one workgroup per chunk, or a persistent grid
 * of k workgroups per CU) x samples per work item of the LDS-staged chunk kernel on these inputs,
 * and beside them the element-parallel kernel with hardware atomics and the colour-ordered
 * element-parallel kernel, and keeps the fastest for (ctx, op, dtype, batch, field location);
 * FEMHOT_SCATTER_CHUNK calls of that shape then take it.  *epc = 0 reports that an
 * element-parallel path won, *block then FEMHOT_PATH_ATOMIC or FEMHOT_PATH_COLOUR.  Untuned,
 * FEMHOT_SCATTER_CHUNK calls take the chunk path except in FP64 Neo-Hookean, which takes the
 * atomic path.  Finally runs the operation once more so the outputs are valid.
 * op FEMHOT_OP_TANGENT times femhot_tangent_csr with atomics, colour-ordered and (thermal, linear
 * elastic) row-owner assembly (out_u = the CSR values, no residual) and keeps the fastest for
 * femhot_tangent_csr calls with FEMHOT_SCATTER_CHUNK.  Synchronises the stream (not for capture).
 * op: FEMHOT_OP_RESIDUAL / FEMHOT_OP_VJP / FEMHOT_OP_ENERGY_GRAD / FEMHOT_OP_TANGENT.  No
 * counterpart in the reference (XLA fuses the element vmap itself); it replaces fe_loss.py:105-114's
 * fixed batch size.  In FP32 the candidates include the software-pipelined chunk kernel, whose
 * staging is LDS DMA issued one step ahead of the arithmetic; *block then carries FEMHOT_CHUNK_DMA. */
#define FEMHOT_CHUNK_DMA 0x10000
#define FEMHOT_OP_RESIDUAL 0
#define FEMHOT_OP_VJP 1
#define FEMHOT_OP_ENERGY_GRAD 2
#define FEMHOT_OP_TANGENT 3   /* femhot_tangent_csr: atomic, colour-ordered and row-owner assembly; out_u = values */
/* *block when an element-parallel path won (*epc = 0): hardware atomics, or colour-ordered plain adds;
 * for the tangent also the row-owner assembly */
#define FEMHOT_PATH_ATOMIC 1
#define FEMHOT_PATH_COLOUR 2
#define FEMHOT_PATH_ROWS 3
int femhot_autotune(femhot_ctx *ctx, int op, int dtype, int batch, const void *field, long long field_stride,
                    const void *u, long long u_stride, const void *w, long long w_stride, void *out_u,
                    long long out_u_stride, void *out_field, long long out_field_stride, double *energy,
                    unsigned flags, int *epc, int *block, int *grid_per_cu, float *ms, void *stream);

/* Sets the chunk configuration femhot_autotune would record (e.g. one saved from an earlier
 * process) for (ctx, op, dtype, batch, field location) and builds its chunk plan now, so later
 * calls (and graph captures) allocate nothing.  block: 128 / 256 / 512, or 256 / 512 |
 * FEMHOT_CHUNK_DMA for the pipelined LDS-DMA kernel (FP32 only).  grid_per_cu: 0 = one workgroup
 * per work item, k > 0 = k persistent workgroups per CU, k < 0 = -k persistent workgroups in all.
 * epc = 0 (and always for op FEMHOT_OP_TANGENT) records an element-parallel path instead, block =
 * FEMHOT_PATH_ATOMIC or FEMHOT_PATH_COLOUR (for the tangent also FEMHOT_PATH_ROWS), as
 * femhot_autotune reports it.
 * No counterpart in the reference. */
int femhot_set_chunk_config(femhot_ctx *ctx, int op, int dtype, int batch, unsigned flags, int epc, int block,
                            int grid_per_cu);

/* Measurement hook (no counterpart in the reference): enable=1 starts recording HIP events on the
 * call's stream immediately before and after the dominant element kernel(s) of every residual /
 * VJP / potential-gradient call on this context; enable=0 waits for them, returns the summed
 * kernel time and the number of timed calls, and stops recording.  Not for use under capture. */
int femhot_kernel_timing(femhot_ctx *ctx, int enable, double *total_ms, long long *launches);

/* CSR SpMV y = A x (FP64), used by the solver and exposed for tests. */
int femhot_spmv(femhot_ctx *ctx, const long long *row_ptr, const int *col_idx, const double *vals,
                const double *x, double *y, void *stream);

/* Orientation fix (mesh.py:123-142) on a HOST connectivity array, in place: where detJ at the
 * one-point rule is negative, swap local nodes 0 and 1.  Returns the number swapped in *n_swapped. */
int femhot_orient_elements(int elem_type, long long n_nodes, const double *coords, long long n_elems, int *conn,
                           long long *n_swapped);

/* Test hook (no counterpart in the reference; for the library's own tests and host sanitizer runs):
 * builds the chunk plan of `epc` elements per chunk for the mesh `desc` on the HOST only (no device
 * memory, no GPU needed) and checks every invariant the chunk kernels index by: element
 * permutation, segment runs, chunk-local connectivity against the mesh, incidence runs and per-corner
 * ranks, coordinate records, zero / shared node lists and the partial-slot layout.  stats[0..7] =
 * n_chunks, max nodes per chunk, max elements per chunk, staged chunk nodes, shared nodes, partial
 * slots, nodes in no element, 1 if the chunks are runs of the caller's element order.  Returns
 * FEMHOT_ERR_INDEX, with the first violation in femhot_last_error(), when an invariant fails. */
int femhot_chunk_plan_check(const femhot_mesh_desc *desc, int epc, long long *stats);

#ifdef __cplusplus
}
#endif
#endif /* FEMHOT_H */
