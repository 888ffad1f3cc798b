/*
 * loadsg_amd.h -- C ABI of the MI355X (gfx950) implementation of LoAdSG's
 * matrix-free prewavelet operator application and CG solve on sparse grids
 * with zero Dirichlet boundaries.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (LSG_OK = 0, negative on error; lsg_last_error() describes the last error of
 * the calling thread).  Nothing throws across this boundary.
 *
 * Vectors cross the boundary in one of two layouts:
 *   LSG_HOST_SLOT   host double arrays in the reference's slot order, i.e. the
 *                   layout of VectorSparseG's data table
 *                   (library/source/extemp/vector.h:247-441, vector.cc:12-23):
 *                   entry i belongs to AdaptiveSparseGrid::getIndexOfTable(i); one
 *                   block of lsg_grid_size() entries per batch instance, instance-major
 *                   (batch x n doubles; a batch-1 operator reads and writes n).
 *   LSG_DEVICE      device arrays in this library's depth-blocked order
 *                   (lsg_grid_device_order gives the permutation), one block of
 *                   lsg_grid_size() entries per batch instance, value type of
 *                   the operator's precision (double for 64, float for 32).
 */
#ifndef LOADSG_AMD_H
#define LOADSG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSG_OK 0
#define LSG_ERR_ARG (-1)
#define LSG_ERR_UNSUPPORTED (-2)
#define LSG_ERR_CAPACITY (-3)
#define LSG_ERR_DEVICE (-4)
#define LSG_ERR_NAN (-5)
#define LSG_ERR_NOCONV (-6)

#define LSG_HOST_SLOT 0
#define LSG_DEVICE 1

/* coefficient families for -div(a grad u) + c u (evaluated at cell centres) */
#define LSG_COEF_CONST 0   /* a = a0, c = c0 (Poisson: a0=1,c0=0; HelmHoltz: a0=0,c0=1) */
#define LSG_COEF_CFG2 1    /* a = 1 + 0.5 prod cos(pi x_i), c = 1 + x_1 */
#define LSG_COEF_CFG3 2    /* a = 1 + |x|^2, c = 0 */
#define LSG_COEF_EXPSIN 3  /* a = exp(sum_k xi_k sin(k pi x_{1+(k mod d)}) / k), c = 0 */
/* General coefficients, sampled on the host at the centre of every cell of depth T+1 of
 * every depth T (the cells StencilTemplate::applyStencilOnCell visits, Stencil.h:187-246;
 * the value a plug-in's integration(cell, dirP, dirQ) would evaluate, Stencil.h:740):
 * element matrix = a(centre) * Poisson + c(centre) * HelmHoltz on that cell.  Operators
 * with these coefficients run on the cell-wise (sparse) path on every grid. */
#define LSG_COEF_CELL 4    /* cell_a[k], cell_c[k] for the cells in lsg_grid_cells() order */
#define LSG_COEF_FUNC 5    /* host functions a_fn(x, ctx), c_fn(x, ctx) called at every cell centre */
/* Stored element matrices: the counterpart of LocalStiffnessMatrices (a StencilTemplate's
 * integration(cell, dirP, dirQ) evaluated once per cell and kept,
 * localStiffnessMatrices/LocalStiffnessMatrices.h:76-321, LocalStiffnessMatrices.cpp:453-531)
 * and of MatrixVectorHomogen::multiplication with such a plug-in (StoreLocalStiffnessMatrix,
 * MatrixVectorHomogen.h:170-245,1619-1700).  Every cell of lsg_grid_cells() carries a
 * symmetric 2^d x 2^d matrix between its corners, lsg_cell_entries(d) = 2^(d-1) (2^d + 1)
 * values: corners p <= q (bit i of a corner = upper end of the cell in dimension i), entry
 * e(p, q) = p 2^d - p (p - 1) / 2 + (q - p), the order of the reference's nested
 * CellIndexIterator loops (cells/iterator_neu.cpp:34-49).  Entries of pairs with a corner on
 * the boundary are never read (the reference does not store them).  Operators with stored
 * matrices run on the cell-wise path and keep ncells x entries values per instance on the GPU. */
#define LSG_COEF_MATRIX 6      /* cell_mat[k * lsg_cell_entries(d) + e] for cell k, entry e */
#define LSG_COEF_INTEGRATOR 7  /* host integrator integ(d, level, pos, p, q, ctx) per cell and pair */

/* CG stopping rules */
#define LSG_STOP_REF_ABS 0   /* reference rule: iterate while delta = r^T D^-1 r > eps, stop early once
                                * delta < eps^2 (cg_method.h:845-871) */
#define LSG_STOP_REL 1       /* ||f - A x||_2 <= eps ||f||_2 */
/* or-ed into stop_rule of lsg_cg_solve: plain CG, no diagonal preconditioner (D = 1 in
 * LSG_STOP_REF_ABS's delta) */
#define LSG_CG_NO_PRECOND 0x100

typedef struct lsg_grid lsg_grid;
typedef struct lsg_op lsg_op;

/* Host coefficient function: value at the point x (d coordinates in (0,1)). */
typedef double (*lsg_coef_fn)(const double* x, void* ctx);
/* Host element integrator: the entry between corners p <= q of the cell
 * prod_i [pos_i 2^-level_i, (pos_i + 1) 2^-level_i] (a cell of depth level + 1), the value a
 * plug-in's StencilTemplate::integration(cell, dirP, dirQ) returns (stencils/Stencil.h:740);
 * called only for pairs whose corners are both interior. */
typedef double (*lsg_cell_integrator_fn)(int d, const int32_t* level, const int32_t* pos, uint32_t p, uint32_t q,
                                         void* ctx);

typedef struct lsg_coef {
    int kind;
    double a0, c0;        /* LSG_COEF_CONST */
    int nxi;              /* LSG_COEF_EXPSIN */
    const double* xi;
    const double* cell_a; /* LSG_COEF_CELL: one value per cell (NULL: no stiffness term) */
    const double* cell_c; /* LSG_COEF_CELL: one value per cell (NULL: no mass term) */
    lsg_coef_fn a_fn;     /* LSG_COEF_FUNC (NULL: no stiffness term); only called during the */
    lsg_coef_fn c_fn;     /*   lsg_op_create / lsg_op_set_coef call that receives it */
    void* ctx;            /* context of a_fn, c_fn and integ */
    const double* cell_mat;       /* LSG_COEF_MATRIX: ncells x lsg_cell_entries(d) values */
    lsg_cell_integrator_fn integ; /* LSG_COEF_INTEGRATOR: only called during the call that receives it */
} lsg_coef;

typedef struct lsg_solve_info {
    int iterations;        /* CG iterations (reference counting, cg_method.h:853-883) */
    int outer;             /* iterative-refinement sweeps (precision 32 only) */
    double final_delta;    /* last r^T D^-1 r */
    double rel_residual;   /* ||f - A x||_2 / ||f||_2 at exit (FP64) */
    int converged;
} lsg_solve_info;

/* ------------------------------------------------------------------ grids */
/* Grid from an index table in reference slot order: idx[i*d + k] is the 1D index
 * (0/1 boundary, 2 = 0.5, sons 2j, 2j+1; indices/index.h:35-81) of point i in
 * dimension k; active[i] = AdaptiveSparseGrid::getActiveTable()[i].
 * Replaces the grid handed to MatrixVectorHomogen(AdaptiveSparseGrid&, ...)
 * (MatrixVectorMultiplication/MatrixVectorHomogen.h:72-77). */
int lsg_grid_create(int d, int64_t n, const int32_t* idx, const uint8_t* active, lsg_grid** out);
/* Regular grid, identical slot order to AdaptiveSparseGrid::AddRecursiveSonsOfPoint(
 * center, level) (sgrid/sparseGrid.cc:557-566). */
int lsg_grid_regular(int d, int level, lsg_grid** out);
int64_t lsg_grid_size(const lsg_grid* g);
int lsg_grid_dim(const lsg_grid* g);
/* Index table and active flags in reference slot order
 * (AdaptiveSparseGrid_Base::getIndexOfTable / getActiveTable, sparseGrid.h:145-173). */
int lsg_grid_slot_order(const lsg_grid* g, int32_t* idx_out, uint8_t* active_out);
/* dev_of_slot[i] = position of slot i in the device (depth-blocked) order. */
int lsg_grid_device_order(const lsg_grid* g, int64_t* dev_of_slot);
/* ------------------------------------------------------- grid generation */
/* A growing set of interior points with the reference's AdaptiveSparseGrid generation
 * calls (host code, no GPU).  Points are 1D index tuples as above (no boundary points:
 * LSG_ERR_ARG); slots are handed out in insertion order, as the reference's data table
 * does, so the slot order of the result equals the reference's. */
typedef struct lsg_gridgen lsg_gridgen;
lsg_gridgen* lsg_gridgen_create(int d);
void lsg_gridgen_destroy(lsg_gridgen* g);
/* AdaptiveSparseGrid::AddPoint(I) (sgrid/sparseGrid.cc:482-517): added active, or made
 * active when already stored. */
int lsg_gridgen_add_point(lsg_gridgen* g, const int32_t* idx);
/* AdaptiveSparseGrid::AddRecursiveSonsOfPoint(I, level) (sgrid/sparseGrid.cc:557-566). */
int lsg_gridgen_add_recursive_sons(lsg_gridgen* g, const int32_t* idx, int level);
/* Refinement towards the origin (corners = 1) or towards the origin and (1, ..., 1)
 * (corners = 2): every active point whose support box [x - 2^-t, x + 2^-t] is closer than
 * `radius` to the corner gets both sons in every dimension whose sum of (depth_i - 1) is
 * at most max_level, repeated until nothing changes; the locally adaptive benchmark grids
 * are this refinement of a regular grid followed by lsg_gridgen_complete_dirichlet.  No
 * counterpart in the reference library (its drivers refine with their own loops). */
int lsg_gridgen_refine_towards(lsg_gridgen* g, double radius, int max_level, int corners);
/* AdaptiveSparseGrid::completeDirichletGrid (sgrid/sparseGrid.cc:771-787):
 * completeGridWithoutBoundary, CompleteToLocalTensorProductGrid
 * (sgrid/localTensorProduct.cc:189-213), completeGridWithoutBoundary, addHangingNodes
 * (sparseGrid.cc:587-621), addPoints2 (sgrid/GridGeneration.cpp:333-378). */
int lsg_gridgen_complete_dirichlet(lsg_gridgen* g);
int64_t lsg_gridgen_size(const lsg_gridgen* g);
/* Index table (size x d) and active flags in slot order, ready for lsg_grid_create. */
int lsg_gridgen_points(const lsg_gridgen* g, int32_t* idx_out, uint8_t* active_out);
/* Structural statistics: out[0]=#depths, [1]=multilevel size N_ML, [2]=#apply items,
 * [3]=#wavefront levels, [4]=algorithmic values moved per matvec (streaming model),
 * [5]=largest depth box, [6]=case-buffer size, [7]=W-buffer size, [8]=part-buffer size,
 * [9]=values this implementation moves per matvec, [10]=of which the apply stage,
 * [11]=streaming-model values of the stages the apply stage performs (prolongation along
 * the last prolongation dimension, stencil, restrictions with the gM exchange, conversion). */
int lsg_grid_stats(const lsg_grid* g, int64_t* out, int nout);
/* The same statistics for the plan an operator runs on (dense grids: every prolongation
 * dimension materialised and single-depth apply items by default; [10] and [11] then
 * exclude the prolongation, which the prolongation stage performs). */
int lsg_op_stats(const lsg_op* op, int64_t* out, int nout);
/* Cells of depth T+1 of every depth T in the grid (CellData / CellStructure,
 * cells/CellStructure.h:23-61: level-T cells whose interior corners are all stored), in
 * the order LSG_COEF_CELL values are read: depth by depth in DepthList order
 * (iterator/depthIterator.h:14-50).  *ncells receives the count; centres (ncells x d,
 * may be NULL) the cell centres; depth_index (ncells, may be NULL) the position of the
 * cell's depth in the depth list. */
int lsg_grid_cells(lsg_grid* g, int64_t* ncells, double* centres, int32_t* depth_index);
/* The same cells as level-T boxes: level[k*d + i] = T_i of the cell's depth, pos[k*d + i] =
 * lower position (the cell is prod_i [pos_i 2^-T_i, (pos_i + 1) 2^-T_i]).  Either output
 * may be NULL. */
int lsg_grid_cell_boxes(lsg_grid* g, int32_t* level, int32_t* pos);
/* Values per cell of an LSG_COEF_MATRIX coefficient: 2^(d-1) (2^d + 1). */
int64_t lsg_cell_entries(int d);
int lsg_grid_destroy(lsg_grid* g);

/* -------------------------------------------------------------- operators */
/* Operator a(u,v) = int a grad u . grad v + c u v on the grid, matrix-free in the
 * prewavelet basis: replaces MatrixVectorHomogen + a StencilTemplate plug-in
 * (Poisson / HelmHoltz / a variable-coefficient integration(), stencils/Stencil.h:138-988,
 * 1605-1793).  precision: 64 (FP64) or 32 (FP32 storage and arithmetic, FP64 reductions).
 * batch: number of instances sharing the grid (coefficients may differ per instance). */
int lsg_op_create(lsg_grid* g, const lsg_coef* coef, int precision, int batch, lsg_op** out);
int lsg_op_set_coef(lsg_op* op, int instance, const lsg_coef* coef);
/* lsg_op_set_coef for instances first .. first + count - 1 (coefs[0 .. count)) in one call: the
 * tables of the separable families are built on several host threads and uploaded with one
 * copy.  No counterpart in the reference (one operator per coefficient there). */
int lsg_op_set_coefs(lsg_op* op, int first, int count, const lsg_coef* coefs);
int lsg_op_batch(const lsg_op* op);
/* Preconditioner of lsg_cg_solve and of the inner FP32 solves of lsg_ir_solve on this
 * operator.  LSG_PRECOND_DIAG (default): the reference's diagonal precond3
 * (applications/cg_method.h:322-451), so that lsg_cg_solve reproduces CG::solveHomogen.
 * LSG_PRECOND_DEPTH_BLOCK: no counterpart in the reference -- block Jacobi over the depth
 * blocks, each block the Galerkin block of the constant-coefficient Laplacian on the
 * prewavelets of that depth (a sum of Kronecker products of 1D prewavelet stiffness and mass
 * matrices, inverted once per grid), scaled symmetrically by sqrt(D0 / D) with D this
 * operator's diagonal and D0 that of the constant-coefficient blocks.  Blocks of up to 128
 * points on FP32 operators (64 on FP64 ones) keep their dense inverse; larger blocks whose
 * points form a product set of 1D positions (up to 256 per dimension, 2048 points) are
 * applied by fast diagonalisation of the Kronecker sum; others keep the diagonal.  On
 * adaptive grids a block is taken over the active points of its depth (hanging points: the
 * identity).  Solutions meet the same stopping rules; iteration counts differ from the
 * reference's. */
#define LSG_PRECOND_DIAG 0
#define LSG_PRECOND_DEPTH_BLOCK 1
int lsg_op_set_precond(lsg_op* op, int kind);
/* y = A x: MatrixVectorHomogen::multiplication(prew, Ax, stencil)
 * (MatrixVectorHomogen.h:660-759), for every batch instance (x and y hold batch x n
 * values in either layout). */
int lsg_matvec(lsg_op* op, const void* x, void* y, int layout);
/* Diagonal preconditioner z_p = a(psi_p, psi_p): Preconditioning<S>::precond3
 * (applications/cg_method.h:322-451). */
int lsg_precond_diag(lsg_op* op, void* z, int layout);
/* CG solve A x = f: CG::solveHomogen(eps, x, f, iterations, usol, matrix, stencil, time)
 * (applications/cg_method.h:759-886).  x holds the start vector on entry.
 * info: array of lsg_op_batch() entries (may be NULL).  Returns LSG_ERR_NOCONV (x and
 * info still filled) when an instance stops at maxit before the stopping rule holds
 * (the reference stops at CG::maxIteration = 500, cg_method.h:35,852), LSG_ERR_NAN when
 * a CG scalar turns NaN. */
int lsg_cg_solve(lsg_op* op, const void* f, void* x, double eps, int stop_rule, int maxit,
                 lsg_solve_info* info, int layout);
/* lsg_cg_solve (layout LSG_DEVICE) without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The CG work vectors, the
 * diagonal, the partial sums and the per-instance state live in one fixed 64 MiB device
 * scratch per GPU, held until the process ends: 5 * n * batch values of the operator's
 * precision plus a few KB must fit, otherwise LSG_ERR_CAPACITY.  The operator itself is
 * created beforehand (lsg_op_create allocates).  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU (two host threads, or two operators' streams)
 * race.  No CG chunk graphs, no compaction of finished instances; results equal
 * lsg_cg_solve's within rounding.  Same statuses as lsg_cg_solve. */
int lsg_cg_solve_noalloc(lsg_op* op, const void* f, void* x, double eps, int stop_rule, int maxit,
                         lsg_solve_info* info);
/* Mixed precision: FP32 operator inner PCG inside FP64 iterative refinement; the
 * outer residual f - A x, all dot products/norms and the stop test are FP64 on the
 * device.  op64 and op32 must share the grid and batch.  stop rule LSG_STOP_REL.
 * Returns LSG_ERR_NOCONV (x and info filled) when an instance misses rel_tol after
 * maxit refinement sweeps, LSG_ERR_NAN on a NaN residual. */
int lsg_ir_solve(lsg_op* op64, lsg_op* op32, const void* f, void* x, double rel_tol,
                 double inner_rel_tol, int maxit, lsg_solve_info* info, int layout);
/* Load vector b_p = int f psi_p of f = sum_j eta_j prod_i sin(m_j pi x_i) (exact 1D
 * quadrature of each factor), for every instance: eta[inst*nterms + j], m[j]. */
int lsg_rhs_sines(lsg_op* op, int nterms, const double* eta, const int32_t* m, void* b, int layout);
/* lsg_ir_solve for several operator pairs at once (instances on different grids: an
 * uneven batch), device layout: group q solves op64[q] / op32[q] (one grid and batch per
 * group) with f[q], x[q] and fills info[q] (lsg_op_batch entries, may be NULL).  One host
 * thread advances every group in lock-step; each phase (FP64 residuals, inner FP32 CG
 * chunks, corrections) is issued for all groups on their operators' streams, so the GPU
 * runs the groups side by side, and the host joins them once per phase.  No counterpart
 * in the reference. */
int lsg_ir_solve_many(int ngroups, lsg_op* const* op64, lsg_op* const* op32, const void* const* f, void* const* x,
                      double rel_tol, double inner_rel_tol, int maxit, lsg_solve_info* const* info);
/* Load vector of a sum of separable functions f = sum_j w_j prod_i g_ji(x_i):
 * b_p = int f psi_p = sum_j w[inst*nterms + j] prod_i int g_ji psi^(i)_p, the 1D moments
 * integrated with a 4-point Gauss rule on every cell of the grid's finest 1D level (exact
 * for polynomial factors of degree <= 6).  vals[(j*d + i)*count + q] is g_ji at node q of
 * lsg_rhs_nodes.  Replaces the load-vector assembly of a driver (the reference forms
 * f_h = M prew(f_I) instead, examples_doxygen/poisson_regular.cpp:57-113); no counterpart
 * in the reference library. */
int lsg_rhs_nodes(const lsg_op* op, int64_t* count, double* nodes);
int lsg_rhs_separable(lsg_op* op, int nterms, const double* vals, const double* w, void* b, int layout);
/* ------------------------------------------------- basis transformations */
/* FP64 operators (LSG_ERR_UNSUPPORTED on FP32 ones), regular and adaptive grids.  On
 * adaptive grids only active points are written (hanging points keep 0 in the outputs of
 * lsg_prew_by_nodal and lsg_nodal_by_prew, and their input values in lsg_hierarchize /
 * lsg_dehierarchize), as the reference's workonindex tests do.  Vectors as in
 * lsg_matvec; in and out must be different arrays.
 * Hierarchical surpluses from nodal values: CalcHierarchicalBasis(u)
 * (BasisTransformations/BasisTransformations.cpp:390-409). */
int lsg_hierarchize(lsg_op* op, const void* u, void* hier, int layout);
/* Nodal values from hierarchical surpluses: CalcNodalBasis(u) (BasisTransformations.cpp:174-198). */
int lsg_dehierarchize(lsg_op* op, const void* hier, void* u, int layout);
/* Prewavelet coefficients of the sparse-grid interpolant of nodal values u:
 * calcPrewByNodal(prew, u) (BasisTransformations.cpp:486-545). */
int lsg_prew_by_nodal(lsg_op* op, const void* u, void* prew, int layout);
/* Nodal values of the prewavelet expansion prew: calcNodalByPrew(prew, u)
 * (BasisTransformations.cpp:670-707). */
int lsg_nodal_by_prew(lsg_op* op, const void* prew, void* u, int layout);
/* out[b] = sqrt(v^T A v), v = prewavelets of the nodal values u of instance b, A the
 * operator: H1::getValue with a Poisson operator, L2::getValue with a mass operator
 * (applications/H1_Norm.cpp:8-78).  out holds lsg_op_batch() values. */
int lsg_energy_norm(lsg_op* op, const void* u, double* out, int layout);

/* Vector norms per instance: out[2b] = max |x_i| over all stored points (L_infty,
 * applications/norms.h:12-52), out[2b+1] = sqrt(sum x_i^2 / dofs) over the active points
 * (L_2, norms.h:54-100).  Any grid and precision; out holds 2 lsg_op_batch() values. */
int lsg_vector_norms(lsg_op* op, const void* x, double* out, int layout);

/* Device timing of the last lsg_matvec: out[0]=total ms, out[1]=apply-kernel ms,
 * out[2]=number of apply launches. */
int lsg_last_timing(const lsg_op* op, double* out, int nout);
/* Tuning: threads per workgroup for the apply stage (a multiple of 64 up to 512); 0 = autotune. */
int lsg_op_set_block(lsg_op* op, int stage, int threads);
int lsg_op_destroy(lsg_op* op);

/* ------------------------------------------------------ warm-start network */
/* MLP n_in -> width (x hidden layers, GELU) -> n_out that maps the parameters of a
 * problem instance to an initial guess of its solution (device order).  No
 * counterpart in the reference.  BF16 operands on the matrix cores, FP32
 * accumulation, FP32 master weights and Adam moments.  All x / y / target arrays
 * are device pointers to FP32, row-major batch x n_in / batch x n_out.  Parameters
 * are flattened layer by layer as W (in x out, row-major) followed by b (out). */
typedef struct lsg_mlp lsg_mlp;
int lsg_mlp_create(int n_in, int width, int hidden_layers, int64_t n_out, uint64_t seed, lsg_mlp** out);
int lsg_mlp_forward(lsg_mlp* m, const float* x, float* y, int batch);
/* mean squared error over batch x n_out and its gradients (kept inside; lsg_mlp_get_grads) */
int lsg_mlp_gradients(lsg_mlp* m, const float* x, const float* target, int batch, double* loss);
/* gradients + one Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr */
int lsg_mlp_train_step(lsg_mlp* m, const float* x, const float* target, int batch, float lr, double* loss);
int lsg_mlp_num_params(const lsg_mlp* m, int64_t* n);
int lsg_mlp_get_params(const lsg_mlp* m, float* host);
int lsg_mlp_set_params(lsg_mlp* m, const float* host);
int lsg_mlp_get_grads(const lsg_mlp* m, float* host);
/* Training targets of a family whose solution is linear in part of the input:
 * out[b][i] = sum_j w[b*J + j] U[(rows[b]*J + j)*n + i] (device pointers; U FP32
 * S x J x n, rows int64, w FP32 batch x J, out FP32 batch x n).  No counterpart in
 * the reference. */
int lsg_mlp_mix_rows(lsg_mlp* m, const float* U, int64_t n, int J, const int64_t* rows, const float* w, int batch,
                     float* out);
int lsg_mlp_destroy(lsg_mlp* m);

const char* lsg_last_error(void);
int lsg_device_available(void);

#ifdef __cplusplus
}
#endif
#endif /* LOADSG_AMD_H */
