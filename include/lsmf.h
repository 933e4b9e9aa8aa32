/*
 * lsmf.h -- C ABI of the MI355X matrix-free Neo-Hookean tangent operator and
 * the GMG-preconditioned CG / Newton solve it drives.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository davydden/large-strain-matrix-free).  The
 * reference's operator is the C++ template
 *   NeoHookOperator<dim, fe_degree, n_q_points_1d, number>
 * (include/mf_nh_operator.h:147-296); this ABI exposes the same member set
 * through an opaque handle, with plain pointers and sizes only.
 *
 * Conventions
 *   - Status codes: 0 = OK, negative = error; lsmf_last_error() returns a
 *     thread-local message for the last failing call on this thread.
 *   - Vectors passed to vmult / diagonal / residual are DEVICE pointers,
 *     borrowed for the duration of the call.  Their element type is the
 *     operator's precision (double for LSMF_FP64, float for LSMF_FP32).
 *   - Global DoF numbering: dof = dim * node + component.  Local node order
 *     inside a cell is lexicographic (x fastest), as in deal.II's
 *     FEEvaluation (include/mf_nh_operator.h:827-832 uses
 *     j + c * dofs_per_component for the cell-local layout).
 *   - All device work is stream-ordered on the stream argument (a
 *     hipStream_t; NULL = default stream); vmult never synchronises the host.
 *     One handle must not be used from two streams at once (the reference's
 *     vmult is not thread-safe either: include/mf_elasticity.h:1755).
 */
#ifndef LSMF_H
#define LSMF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int lsmf_status;
#define LSMF_OK 0
#define LSMF_ERR_INVALID_ARGUMENT (-1)
#define LSMF_ERR_UNKNOWN_CACHING (-2) /* ExcMessage("Unknown caching"), include/mf_nh_operator.h:438-442 */
#define LSMF_ERR_OUT_OF_MEMORY (-3)
#define LSMF_ERR_NONPOSITIVE_DET_F (-4) /* Assert(det_F > 0), include/mf_nh_operator.h:497-503 */
#define LSMF_ERR_NOT_CACHED (-5)        /* vmult before cache() */
#define LSMF_ERR_HIP (-6)
#define LSMF_ERR_NOT_CONVERGED (-7) /* "No convergence in nonlinear solver!", include/mf_elasticity.h:2557 */
#define LSMF_ERR_UNSUPPORTED (-8)   /* (dim, degree, n_q_1d) not instantiated, src/main.cc:102-110 */
#define LSMF_ERR_NONPOSITIVE_DIAGONAL (-9) /* Assert(diag > 0), include/mf_nh_operator.h:1543-1547 */

typedef void *lsmf_stream; /* hipStream_t */

/* MF caching variants, include/mf_nh_operator.h:284-293 ("MF caching" prm key,
 * include/mf_elasticity.h:445-450). */
enum {
  LSMF_CACHING_SCALAR = 0,
  LSMF_CACHING_TENSOR2 = 1,
  LSMF_CACHING_TENSOR4 = 2,
  LSMF_CACHING_TENSOR4_NS = 3,
  LSMF_CACHING_SCALAR_REFERENTIAL = 4
};

/* operator number type: the reference instantiates NeoHookOperator with
 * number = double on the fine level and float on the MG levels
 * (include/mf_elasticity.h:673-674, 900-902). */
enum { LSMF_FP64 = 0, LSMF_FP32 = 1 };

/* vmult phase mask, the reference's MFMask (include/mf_nh_operator.h:73-86).
 * ZERO and MPI have no effect on one GPU (dst is overwritten unless add != 0;
 * there are no ghosts).  As in the reference, RW gates the gather, the
 * scatter and the identity on constrained rows (:703-705): a mask without RW
 * does no cell work (dst = 0 in set mode, unchanged with add); RW without SF
 * times gather and scatter only, RW | SF adds the sum factorization.  PLAIN
 * is new: apply the whole operator without constraints (constrained DoFs are
 * read and written like free ones; the other bits are then ignored), which
 * the residual needs for the Dirichlet lifting -K g
 * (include/mf_elasticity.h:2894-2898). */
enum {
  LSMF_MASK_ZERO = 1u << 0,
  LSMF_MASK_MPI = 1u << 1,
  LSMF_MASK_RW = 1u << 2,
  LSMF_MASK_SF = 1u << 3,
  LSMF_MASK_QD = 1u << 4,
  LSMF_MASK_DEFAULT = 0x1fu,
  LSMF_MASK_CELLLOOP = (1u << 2) | (1u << 3) | (1u << 4),
  LSMF_MASK_RWSF = (1u << 2) | (1u << 3),
  LSMF_MASK_PLAIN = 1u << 5
};

/* Mesh of one level.  Replaces the DoFHandler + MatrixFree index data the
 * reference builds in Solid::system_setup / setup_matrix_free
 * (include/mf_elasticity.h:1610-1708, 1752-1831). */
typedef struct {
  int dim;                    /* 2 or 3 */
  int degree;                 /* FE_Q degree p */
  int n_cells;
  int n_nodes;                /* (p+1)^dim support points per cell are shared */
  const int32_t *cell_nodes;  /* host, [n_cells][(p+1)^dim], lexicographic */
  const double *coords;       /* host, [n_nodes][dim] reference coordinates of the
                                 support points; cell geometry is the Q1 map of the
                                 2^dim corner nodes (StaticMappingQ1,
                                 include/mf_elasticity.h:1828-1829) */
  const uint8_t *constrained; /* host, [n_nodes*dim], 1 = homogeneous Dirichlet
                                 DoF of this operator, or NULL */
  const int32_t *cell_colour; /* host, [n_cells] colouring with no node shared
                                 inside a colour, or NULL (greedy colouring) */
  /* Hanging-node (general affine, homogeneous) constraints, the constraint
   * pools deal.II's read_dof_values / distribute_local_to_global resolve per
   * cell (include/mf_nh_operator.h:745-749, 769-780): node nc_nodes[k] has,
   * in every component, the value sum_j nc_weights[j] * value(nc_masters[j])
   * over j in [nc_ptr[k], nc_ptr[k+1]).  Masters must be unconstrained by
   * this list (closed constraints); a master's Dirichlet components read as
   * zero.  Constrained nodes' rows act as identity, like Dirichlet rows.
   * 0 / NULL: none.  With constraints the cells layout is used; the
   * constrained nodes' results reach their masters in one pass after the
   * colours (in a fixed order, so products are reproducible), and
   * compute_diagonal evaluates, on each cell with constrained nodes, the
   * closed form for the constrained basis function of every node the cell
   * couples (its own unconstrained nodes and the masters). */
  int n_node_constraints;
  const int32_t *nc_nodes;   /* [n_node_constraints] */
  const int32_t *nc_ptr;     /* [n_node_constraints + 1] */
  const int32_t *nc_masters; /* [nc_ptr[n]] */
  const double *nc_weights;  /* [nc_ptr[n]] */
} lsmf_mesh_desc;

typedef struct {
  int n_q_1d;    /* QGauss points per direction (n_q_points_1d) */
  int precision; /* LSMF_FP64 or LSMF_FP32 */
  int caching;   /* LSMF_CACHING_* */
  int layout;    /* LSMF_LAYOUT_AUTO (0) or LSMF_LAYOUT_CELLS */
} lsmf_op_options;

/* Cell loop layout.  AUTO: on a structured mesh (the lexicographic numbering
 * lsmf_mesh_create produces, detected from cell_nodes) cells are processed in
 * tiles (rows of 9x1x1 Q2 cells in 3D, ...) whose shared nodes are gathered once and
 * accumulated in LDS, with the tiles coloured by parity; otherwise, and with
 * CELLS, one cell per lane group with the cells coloured (given colouring or
 * greedy).  Both give the same operator to round-off. */
enum { LSMF_LAYOUT_AUTO = 0, LSMF_LAYOUT_CELLS = 1 };

typedef struct lsmf_op lsmf_op;

/* NeoHookOperator::initialize (include/mf_nh_operator.h:394-443): binds the
 * mesh and the caching mode and allocates the per-qp caches.  Uploads index
 * tables, colouring and geometry to the current HIP device. */
lsmf_status lsmf_op_create(const lsmf_mesh_desc *mesh, const lsmf_op_options *opt, lsmf_op **out);

/* NeoHookOperator::clear (include/mf_nh_operator.h:379-390) + destructor. */
lsmf_status lsmf_op_destroy(lsmf_op *op);

/* NeoHookOperator::set_material (include/mf_nh_operator.h:596-608): the
 * reference passes a matrix and an inclusion Material selected by
 * material_id (:930-932); here mu and lambda are per-cell host arrays
 * [n_cells], in the mesh's cell order, so the two-material case is a special
 * case.  lambda is Lame's first parameter, lambda = 2 mu nu / (1 - 2 nu)
 * (include/material.h:82-94). */
lsmf_status lsmf_set_material_per_cell(lsmf_op *op, const double *mu, const double *lambda);

/* The reference's two-material form of set_material(matrix, inclusion)
 * (include/mf_nh_operator.h:596-608, selection at :930-932): cell c takes
 * the inclusion's (mu1, lambda1) where material_id[c] != 0 and the matrix's
 * (mu0, lambda0) elsewhere; material_id is a host array [n_cells] in the
 * mesh's cell order. */
lsmf_status lsmf_set_material_two(lsmf_op *op, const uint8_t *material_id, double mu0, double lambda0, double mu1,
                                  double lambda1);

/* The formulation argument of Material_Compressible_Neo_Hook_One_Field
 * (include/material.h:82-94), which the reference's set_material passes in:
 * 1 (default): psi = mu/2 (tr C - d - 2 ln J) + lambda ln^2 J;
 * 0: the volumetric/isochoric split, psi = kappa/4 (J^2 - 1 - 2 ln J)
 * + mu/2 (tr b_bar - d), kappa = lambda + 2 mu / 3, whose operator caches
 * J^{-1/d} only (include/mf_nh_operator.h:467-487, 1001-1181).  Formulation 0
 * runs with scalar or tensor4 caching (LSMF_ERR_UNSUPPORTED otherwise) and
 * has no residual VJP.  Call before cache(). */
lsmf_status lsmf_set_formulation(lsmf_op *op, int formulation);

/* The displacement argument of NeoHookOperator::initialize
 * (include/mf_nh_operator.h:397, 404): the linearisation state.  u_dev is a
 * device FP64 vector [n_nodes*dim]; it is copied (and converted for FP32
 * operators), so it need not outlive the call. */
lsmf_status lsmf_set_linearization(lsmf_op *op, const double *u_dev, lsmf_stream stream);

/* NeoHookOperator::cache (include/mf_nh_operator.h:447-592).  Kinematics
 * (F, det F, ln J) are evaluated in FP64 for both precisions.  Returns
 * LSMF_ERR_NONPOSITIVE_DET_F if any quadrature point has det F <= 0 (this
 * call synchronises the stream to read that count). */
lsmf_status lsmf_cache(lsmf_op *op, lsmf_stream stream);

/* lsmf_cache without any allocation: the same cache kernel and result,
 * with the det F <= 0 counter in a file-scope device array of the library
 * (64 ints, of which this call uses one) instead of the operator's own
 * counter.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (HIP graph set-up code, serving loops, harnesses
 * that forbid hipMalloc).  It calls no hipMalloc / hipFree: the array's
 * address is looked up once per device, on the first call with that device
 * current.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (several host threads or streams) race; calls in
 * sequence are safe (each clears the counter first).  Its scratch does not
 * grow with the operator, so no size can exceed it.  Like lsmf_cache it
 * synchronises the stream to read the count. */
lsmf_status lsmf_cache_noalloc(lsmf_op *op, lsmf_stream stream);

/* NeoHookOperator::compute_diagonal (include/mf_nh_operator.h:1503-1557):
 * exact diagonal, constrained entries 1, and its inverse.  DoFs that no cell
 * references have diagonal 0 and inverse 0.  Any other entry <= 0 returns
 * LSMF_ERR_NONPOSITIVE_DIAGONAL (the reference's Assert(diag > 0),
 * :1543-1547); this call synchronises the stream to read that count. */
lsmf_status lsmf_compute_diagonal(lsmf_op *op, lsmf_stream stream);

/* NeoHookOperator::vmult / vmult_add / Tvmult / Tvmult_add
 * (include/mf_nh_operator.h:612-706): dst = A src (add = 0) or dst += A src
 * (add != 0).  Constrained rows act as identity (:703-705) unless
 * LSMF_MASK_PLAIN.  The operator is symmetric, so Tvmult = vmult.  dst and
 * src must not alias. */
lsmf_status lsmf_vmult(lsmf_op *op, void *dst, const void *src, int add, unsigned mask, lsmf_stream stream);

/* n_vectors products dst_b = A src_b of lsmf_vmult (set mode, constrained
 * rows as identity), src_b = src + b * n and dst_b = dst + b * n with n =
 * lsmf_op_size(op), in one launch per colour for all vectors; each dst_b is
 * bit-identical to lsmf_vmult's.  No counterpart in the reference, whose
 * coarse level is solved iteratively: it serves the exact coarse solve,
 * which probes the coarse matrix with the n unit vectors at once. */
lsmf_status lsmf_vmult_batched(lsmf_op *op, void *dst, const void *src, int n_vectors, lsmf_stream stream);

/* get_matrix_diagonal_inverse() / el(i,i) (include/mf_nh_operator.h:206-211,
 * 1561-1571): device pointers to the diagonal and its inverse (operator
 * precision), valid until the next compute_diagonal or destroy. */
lsmf_status lsmf_get_diagonal(lsmf_op *op, const void **diag_dev, const void **inv_diag_dev);

/* precondition_Jacobi (include/mf_nh_operator.h:348-358): dst = omega D^-1 src */
lsmf_status lsmf_precondition_jacobi(lsmf_op *op, void *dst, const void *src, double omega, lsmf_stream stream);

/* m() / n() (include/mf_nh_operator.h:362-375) */
size_t lsmf_op_size(const lsmf_op *op);

/* memory_consumption (include/mf_nh_operator.h:300-328): bytes of device
 * memory vmult reads (caches, index tables, geometry), diagonals excluded. */
size_t lsmf_memory_consumption(const lsmf_op *op);

/* FP64 internal-force residual, the cell part of Solid::assemble_system
 * (include/mf_elasticity.h:2758-2804): rhs = - sum_cells int P : Grad N dV,
 * P = tau F^-T, at the linearisation state set by lsmf_set_linearization.
 * No constraints applied; rhs_dev is FP64 [n_nodes*dim].  The operator must
 * be FP64. */
lsmf_status lsmf_residual(lsmf_op *op, double *rhs_dev, lsmf_stream stream);

/* Vector-Jacobian product of lsmf_residual (new; the reference has no
 * adjoint -- its nearest relative differentiates P(F) per quadrature point
 * with Sacado in include/mf_ad_nh_operator.h:415-485).  For the cotangent w
 * (device FP64, n_nodes*dim):
 *   du_dev      = (dR/du)^T w = -K w            (K: unconstrained tangent)
 *   dmu[c]      = (dR/dmu_c)^T w,  dlambda[c] = (dR/dlambda_c)^T w
 * with dmu, dlambda device FP64 arrays [n_cells] in mesh cell order.
 * Kinematics (F, J, ln J) are recomputed from the linearisation state, no
 * cache is read or stored.  FP64 operators only. */
lsmf_status lsmf_residual_vjp(lsmf_op *op, const double *w_dev, double *du_dev, double *dmu, double *dlambda,
                              lsmf_stream stream);

/* thread-local description of the last error on this thread */
const char *lsmf_last_error(void);

/* Per-launch timing of the operator's fused cell kernels with HIP events on
 * the launch stream, the counterpart of the reference's TimerOutput and LIKWID
 * regions around vmult (include/mf_nh_operator.h:108-119,
 * include/mf_elasticity.h:2354-2448).  enable != 0 starts recording (and
 * discards earlier records); lsmf_profile_read synchronises the device,
 * returns the summed kernel time in milliseconds and the number of launches
 * since enabling, and resets the counters. */
lsmf_status lsmf_profile_enable(lsmf_op *op, int enable);
lsmf_status lsmf_profile_read(lsmf_op *op, double *total_ms, int *n_launches);

/* ------------------------------------------------------------------------
 * Structured meshes of the unit square / cube (and the Cook membrane used by
 * the reference's tests).  Replaces GridGenerator::subdivided_hyper_rectangle
 * + GridTools::transform + refine_global (include/mf_elasticity.h:1339-1398)
 * for the geometries this library runs.  Host-only; no GPU needed.
 * ---------------------------------------------------------------------- */
enum { LSMF_GEOMETRY_BOX = 0, LSMF_GEOMETRY_COOK = 1 };

typedef struct {
  int dim, degree;
  int n_cells_dir[3];
  int n_cells, n_nodes;
  int32_t *cell_nodes; /* [n_cells][(p+1)^dim] */
  double *coords;      /* [n_nodes][dim] */
  int32_t *cell_colour; /* [n_cells] parity colouring, 2^dim colours */
} lsmf_mesh;

/* Box [lo, hi] with n_cells_dir cells per direction; geometry COOK applies the
 * reference's grid_y_transform (include/mf_elasticity.h:1252-1270) to the box
 * and then scales all coordinates by `scale`. */
lsmf_status lsmf_mesh_create(int dim, int degree, const int *n_cells_dir, const double *lo, const double *hi,
                             int geometry, double scale, lsmf_mesh **out);
lsmf_status lsmf_mesh_destroy(lsmf_mesh *mesh);
/* side s in [0, 2*dim): face 2d (+0) is the low-x_d side, 2d+1 the high side,
 * as deal.II's face numbering.  Writes (cell, local face) pairs; returns the
 * count (call with NULL outputs to query). */
int lsmf_mesh_boundary_faces(const lsmf_mesh *mesh, int side, int32_t *cells, int32_t *faces);
/* nodes on side s; returns the count (NULL output to query) */
int lsmf_mesh_boundary_nodes(const lsmf_mesh *mesh, int side, int32_t *nodes);

/* ------------------------------------------------------------------------
 * Newton solve with GMG-preconditioned CG on a structured hierarchy.
 * Replaces Solid::run / solve_nonlinear_timestep / setup_matrix_free /
 * solve_linear_system (include/mf_elasticity.h:1187-1222, 1712-2241,
 * 2245-2559, 3012-3214): FP64 outer CG on the fine operator, multigrid
 * V-cycle with Chebyshev smoothing (Jacobi-preconditioned, eigenvalues by a
 * CG/Lanczos estimate), Chebyshev used as the coarse solver, tensor-product
 * transfer, level operators in FP32 or FP64 linearised at the displacement
 * interpolated to each level.  Dirichlet values are increments applied at
 * Newton iteration 0 of every load step (include/mf_elasticity.h:2963,
 * 3204-3211); the lifting -K g enters the first residual.
 * ---------------------------------------------------------------------- */
typedef struct {
  int dim, degree, n_q_1d;
  int n_cells_coarse[3];         /* cells per direction of the coarsest level */
  int n_levels;                  /* finest level has n_cells_coarse * 2^(n_levels-1) cells */
  int geometry;                  /* LSMF_GEOMETRY_BOX / _COOK */
  double lo[3], hi[3], scale;
  int caching;                   /* fine (FP64) operator */
  int mg_precision;              /* LSMF_FP32 or LSMF_FP64 level operators */
  int mg_caching;
  int smoother_degree;           /* Chebyshev degree, levels > 0 (5: include/mf_elasticity.h:2104) */
  double smoothing_range;        /* 20 (:2102) */
  int eig_cg_iterations;         /* 30 ("MF Chebyshev number CG iterations") */
  double coarse_reduction;       /* 1e-3: coarse Chebyshev as a solver (:2087-2096) */
  int coarse_eig_iterations;     /* 0: as many as coarse DoFs ("accurate eigenvalues") */
  double tol_lin;                /* CG: |r| <= tol_lin |rhs| ("Residual", :3027) */
  int max_lin_iterations;
  int max_newton;                /* "Max iterations Newton-Raphson" */
  double tol_f, tol_u;           /* "Tolerance force", "Tolerance displacement" */
  double dt, t_end;              /* "Time step size", "End time" */
  unsigned dirichlet_components[6]; /* per side s: bit c = component c prescribed */
  double dirichlet_increment[6][3]; /* value per side and component (increment per load step) */
  unsigned neumann_sides;            /* bit s: traction on side s */
  double neumann_traction[6][3];     /* traction = t * value, reference configuration */
  int verbose;
  int coarse_solver;             /* LSMF_COARSE_CHEBYSHEV (reference) or LSMF_COARSE_DIRECT */
  int use_graph;                 /* replay the V-cycle as a captured HIP graph */
  int preconditioner;            /* LSMF_PRECONDITIONER_GMG / _JACOBI / _NONE ("Preconditioner type") */
  double preconditioner_relaxation; /* Jacobi damping omega ("Preconditioner relaxation") */
  double tol_f_abs, tol_u_abs;   /* "Absolute tolerance force" / "... displacement" (:2262-2266) */
  int formulation;               /* material formulation, "Formulation" (1 or 0) */
  int cond_number_cg_iterations; /* "Condition number CG iterations": an unpreconditioned CG of
                                    this many steps before each linear solve estimates the
                                    operator's condition number from its Lanczos coefficients
                                    (include/mf_elasticity.h:3029-3046); 0 skips it */
} lsmf_solver_options;

enum { LSMF_PRECONDITIONER_GMG = 0, LSMF_PRECONDITIONER_JACOBI = 1, LSMF_PRECONDITIONER_NONE = 2 };

/* coarse-level solver: the reference's Chebyshev iteration used as a solver
 * (include/mf_elasticity.h:2087-2096, 2113, "MF Chebyshev coarse = true"), an
 * exact FP64 solve with the inverse of the assembled coarse matrix (coarsest
 * level up to 4096 DoFs), or the reference's iterative coarse solver
 * ("MF Chebyshev coarse = false", :2115-2127): CG from zero preconditioned by
 * the coarse level's Chebyshev smoother, to a residual reduction of 1e-3 or
 * 1e-10 absolute, at most max(n, 100) steps.  Its iteration count depends
 * on the data, so the V-cycle then runs without graph capture. */
enum { LSMF_COARSE_CHEBYSHEV = 0, LSMF_COARSE_DIRECT = 1, LSMF_COARSE_CG = 2 };

typedef struct {
  int newton_iterations;  /* as printed by "Converged in N Newton iterations" */
  int linear_iterations;  /* CG iterations summed over the step */
  int max_linear_iterations;
  double residual_0, residual_final;
  double time_total_s, time_setup_s, time_linear_s, time_residual_s;
  double cond_number;     /* last estimate of the step (0 when not estimated) */
  double time_setup_cache_s; /* of time_setup_s: caches and diagonals of the fine operator and
                                all levels (the rest: Chebyshev eigenvalue estimates, coarse solve) */
} lsmf_step_stats;

typedef struct lsmf_solver lsmf_solver;

lsmf_status lsmf_solver_options_default(lsmf_solver_options *opt);
/* mu, lambda: host arrays over the finest cells (mesh order); levels use the
 * volume average of their children. */
lsmf_status lsmf_solver_create(const lsmf_solver_options *opt, const double *mu, const double *lambda,
                               lsmf_solver **out);
lsmf_status lsmf_solver_destroy(lsmf_solver *s);
/* load stepping t = dt, 2 dt, ... <= t_end; stats[k] per step */
lsmf_status lsmf_solver_run(lsmf_solver *s, int max_steps, lsmf_step_stats *stats, int *n_steps);
/* one load step (the next t = t_prev + dt); *done = 1 and no work when the
 * next t exceeds t_end; *t_out receives the step's time */
lsmf_status lsmf_solver_step(lsmf_solver *s, lsmf_step_stats *stats, double *t_out, int *done);
/* Dirichlet increments of the next load steps (the values interpolate_boundary_values
 * puts into the constraints at Newton iteration 0, include/mf_elasticity.h:2963-2990):
 * host array over all finest DoFs, entries of unconstrained DoFs ignored. Replaces
 * the per-side constants of the options. */
lsmf_status lsmf_solver_set_dirichlet_increment(lsmf_solver *s, const double *inc_host);
/* surface traction per side and component (reference configuration) for the next
 * load steps, applied as given (the options' traction is scaled by t instead);
 * the Neumann term of assemble_system (include/mf_elasticity.h:2859-2892) */
lsmf_status lsmf_solver_set_traction(lsmf_solver *s, const double traction[6][3]);
lsmf_status lsmf_solver_get_solution(const lsmf_solver *s, double *u_host);
size_t lsmf_solver_n_dofs(const lsmf_solver *s, int level); /* level -1 = finest */
/* linearise the fine operator and every level at u (host FP64, finest numbering) */
lsmf_status lsmf_solver_set_state(lsmf_solver *s, const double *u_host);
/* one V-cycle z = M r (device FP64 vectors of the finest level) */
lsmf_status lsmf_solver_precondition(lsmf_solver *s, double *z_dev, const double *r_dev);
/* FP64 PCG with the V-cycle: x = A^-1 b to |r| <= tol; x_dev starts at zero */
lsmf_status lsmf_solver_cg(lsmf_solver *s, double *x_dev, const double *b_dev, double tol, int max_it, int *its,
                           double *res);
/* Chebyshev data of a level after set_state */
lsmf_status lsmf_solver_level_info(const lsmf_solver *s, int level, double *lambda_max_est, double *lambda_min_est,
                                   int *degree);
/* FP64 transfer between level-1 and level on device vectors: mode 0
 * prolongate (dst fine), 1 restrict = P^T (dst coarse), 2 interpolate the
 * fine function at the coarse support points (dst coarse).  No constraints. */
lsmf_status lsmf_solver_transfer(lsmf_solver *s, int level, int mode, double *dst, const double *src);
/* the FP64 fine operator (borrowed) */
lsmf_op *lsmf_solver_fine_operator(lsmf_solver *s);

#ifdef __cplusplus
}
#endif

#endif /* LSMF_H */
