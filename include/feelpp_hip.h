/*
 * feelpp_hip.h -- C-ABI of the MI355X (gfx950) device layer for Feel++'s
 * form assembly and Jacobi-CG solve, continuous Lagrange Pk on simplices.
 *
 * This is the boundary a Feel++ "hip" backend kind binds: every entry point
 * below replaces one step of the reference path, cited as file:line relative
 * to the Feel++ source tree (feelpp/feel/...).  Plain pointers and sizes only;
 * host arrays are borrowed for the duration of a call; device memory is owned
 * by the handles.  Every function returns 0 on success or a negative status;
 * feel_hip_last_error() describes the last failure of the calling thread.
 * Solver non-convergence is not an error: it is reported in
 * feel_hip_solve_data, as SolveData is (feelalg/solverlinear.hpp:81).
 *
 * Threading: a context owns one HIP stream; calls on one context must not
 * overlap.  Use one context per host thread.
 */
#ifndef FEELPP_HIP_H
#define FEELPP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct feel_hip_ctx feel_hip_ctx;
typedef struct feel_hip_space feel_hip_space;
typedef struct feel_hip_graph feel_hip_graph;
typedef struct feel_hip_mat feel_hip_mat;
typedef struct feel_hip_vec feel_hip_vec;
typedef struct feel_hip_rows feel_hip_rows;

enum {
    FEEL_HIP_OK = 0,
    FEEL_HIP_EINVAL = -1,      /* bad argument (shape, order, null handle) */
    FEEL_HIP_EUNSUPPORTED = -2,/* expression / space outside the supported subset */
    FEEL_HIP_EDEVICE = -3,     /* HIP runtime error */
    FEEL_HIP_ENOMEM = -4,      /* device allocation failed */
    FEEL_HIP_EPATTERN = -5     /* an element entry falls outside the sparsity graph */
};

/* precision of stored matrix values / of the solve */
enum { FEEL_HIP_FP64 = 1, FEEL_HIP_FP32 = 2, FEEL_HIP_FP64_FP32 = 3 };

/* Expression subset of integrate(_range=elements(mesh), _expr=...)
 * (feelvf/integrate.hpp:64-98; bilinearformcontext.hpp:455-628):
 *   MASS        mass_scale * idv(m) * idt(u) * id(v)
 *   DIFFUSION   diff_scale * inner(idv(k) * gradt(u), grad(v))   (feelpde/cg_laplacian.hpp:124-125)
 *   ELASTICITY  inner(lambda tr(eps(u)) I + 2 mu eps(u), grad(v)), lambda/mu from P0 E, nu
 *               (quickstart/qs_elasticity.cpp:41-63)
 * MASS|DIFFUSION may be combined.  P0 coefficients are per-element device
 * vectors (nullptr = 1). */
enum { FEEL_HIP_EXPR_MASS = 1, FEEL_HIP_EXPR_DIFFUSION = 2, FEEL_HIP_EXPR_ELASTICITY = 4 };

typedef struct {
    int kind;                       /* FEEL_HIP_EXPR_* bitmask */
    int quad_order;                 /* "im(d,order,simplex)", pointsetquadrature.hpp:85-159 */
    int precision;                  /* which value arrays to write: FEEL_HIP_FP64 / FP32 / both */
    double diff_scale;
    const feel_hip_vec* diff_coef;  /* length nelements or null */
    double mass_scale;
    const feel_hip_vec* mass_coef;  /* length nelements or null */
    const feel_hip_vec* young;      /* elasticity: E per element */
    const feel_hip_vec* poisson;    /* elasticity: nu per element */
    int matrices;                   /* element-matrix path, FEEL_HIP_MATRICES_* (0 = autotuned) */
} feel_hip_bilinear_desc;

/* How element matrices reach the CSR values.  FUSED: each entry is contracted
 * inside the scatter from the element factors and reference tensors.  MFMA
 * (P3 and P4 on tetrahedra, scalar forms and elasticity): all element matrices
 * are first materialised with FP64 matrix-core GEMMs into a workspace of about
 * nelements * (ncomp nloc)^2 doubles held by the context (one per context, grown
 * to the largest assembly, freed with the context), then summed by the same
 * scatter.  AUTO: other spaces use FUSED; P3/P4 tetrahedra time both on the
 * graph's first assembly with add = 0 and keep the faster (that one assembly
 * waits for the context stream; later ones are asynchronous). */
enum { FEEL_HIP_MATRICES_AUTO = 0, FEEL_HIP_MATRICES_FUSED = 1, FEEL_HIP_MATRICES_MFMA = 2 };

/* Source of integrate(_range=elements(mesh), _expr=f*id(v)) (linearformcontext.hpp:294-305) */
enum { FEEL_HIP_SRC_CONST = 0, FEEL_HIP_SRC_SINSIN = 1, FEEL_HIP_SRC_VECTOR = 2 };
typedef struct {
    int kind;            /* CONST: f = p[0]; SINSIN: f = p[0] sin(pi x) sin(pi y); VECTOR: f = (p0,p1,p2) */
    int quad_order;
    double params[4];
} feel_hip_linear_desc;

/* Dirichlet elimination modes, on.type (feelpp/feel/options.cpp:338, feelalg/enums.hpp:59-62) */
enum { FEEL_HIP_ON_ELIMINATION = 0, FEEL_HIP_ON_ELIMINATION_SYMMETRIC = 1 };

typedef struct {
    double rtol;        /* ksp-rtol: ||B r_k|| <= max(rtol ||B r_0||, atol) (UIRNorm, solverlinearpetsc.cpp:592,610) */
    double atol;        /* ksp-atol */
    double dtol;        /* ksp-dtol */
    int maxit;          /* ksp-maxit */
    int precision;      /* FEEL_HIP_FP64: plain FP64 Jacobi-CG; FEEL_HIP_FP32: FP32 Jacobi-CG with
                           FP64 reliable updates (FP64 solution, residual and stopping test) */
    double inner_rtol;  /* FP32: reduction of the residual estimate between reliable updates (1e-2) */
    int inner_maxit;    /* unused by the single solve */
    int max_outer;      /* FP32: cap on reliable updates */
    int check_every;    /* host polls the device convergence flag every this many iterations (0 = auto) */
    int inner_operator; /* FP32: operator of the inner iterations, FEEL_HIP_INNER_* */
} feel_hip_solve_opts;

/* Operator of the FP32 inner iterations.  ASSEMBLED: the matrix's FP32 values
 * (rounded once from the FP64 assembly).  ELEMENT: element by element from the
 * per-element factors k J K^{-1} K^{-T} in FP32, available while the matrix
 * holds a P2 tetrahedral diffusion assembly (operator=, quadrature order >= 2)
 * followed only by symmetric unit-diagonal eliminations; FP64 residuals still
 * use the assembled FP64 values; the FP32 matrix image is then not needed.
 * AUTO: ASSEMBLED (faster on MI355X: the element operator writes its element
 * results and gathers them back by row, about as many bytes per product as the
 * assembled FP32 values and columns).
 * No counterpart in the reference (PETSc applies the assembled matrix). */
enum { FEEL_HIP_INNER_AUTO = 0, FEEL_HIP_INNER_ASSEMBLED = 1, FEEL_HIP_INNER_ELEMENT = 2 };

typedef struct {
    int converged;      /* SolveData::isConverged */
    int its;            /* total Jacobi-CG iterations (A*p products) */
    double residual;    /* final ||B r|| (preconditioned norm), FP64 */
    int reason;         /* KSPConvergedReason: 2 rtol, 3 atol, -3 maxit, -4 dtol, -8 indefinite preconditioner,
                           -9 nan, -10 indefinite matrix */
    int outer_its;      /* FP64 residual evaluations after the first (reliable updates; 1 for FP64) */
    double rnorm0;      /* initial ||B r_0|| */
    int fp64_its;       /* iterations run in FP64: all of them for FP64; for FP32, those after the solve
                           found the matrix too ill-conditioned for FP32 (Lanczos estimate of kappa) */
} feel_hip_solve_data;

/* ---- post-solve evaluation ------------------------------------------------ */
/* Reference functions for error norms: LINEAR c0 + c1 x + c2 y + c3 z,
 * SINSIN c0 sin(pi x) sin(pi y). */
enum { FEEL_HIP_FN_ZERO = 0, FEEL_HIP_FN_LINEAR = 1, FEEL_HIP_FN_SINSIN = 2 };
typedef struct {
    int kind;
    double c[4];
} feel_hip_function;

/* ---- lifetime ------------------------------------------------------------ */
/* Handles are reference counted the way MatrixSparse/GraphCSR shared_ptrs are
   (feelalg/matrixsparse.hpp:113): each object keeps what it was built on
   (space -> context, graph -> space, matrix -> graph, vector/rows/batch ->
   context, batch family -> graph) alive, so the *_destroy calls may come in
   any order.  One host thread per context. */
const char* feel_hip_last_error(void);
int feel_hip_ctx_create(int device, feel_hip_ctx** out);
int feel_hip_ctx_destroy(feel_hip_ctx* ctx);
int feel_hip_ctx_synchronize(feel_hip_ctx* ctx);
/* HIP stream (hipStream_t) the context launches on */
void* feel_hip_ctx_stream(feel_hip_ctx* ctx);
/* Device memory held by the context's handles, in bytes */
int64_t feel_hip_ctx_device_bytes(feel_hip_ctx* ctx);

/* Kernel timing with HIP events on the context stream.  period 0: off; N > 0:
 * every N-th launch of each timed kernel is bracketed by events (1 = every
 * launch; a larger period keeps the events' own cost out of a timed loop).
 * name: "spmv_fp64", "spmv_fp32", "assemble_gather", "assemble_geo". */
int feel_hip_ctx_set_timing(feel_hip_ctx* ctx, int period);
int feel_hip_ctx_timing(feel_hip_ctx* ctx, const char* name, int64_t* launches, double* total_ms);
int feel_hip_ctx_reset_timing(feel_hip_ctx* ctx);

/* ---- mesh + function space ----------------------------------------------
 * Replaces the upload of a FunctionSpace: mesh geometry (Mesh<Simplex<d>>,
 * feeldiscr/meshimpl.hpp:163-240) and the DofTable local-to-global map
 * (feeldiscr/doftable.hpp:732-750).  elem2dof is ne x (nloc*ncomp), local index
 * nloc*c + l (doftable.hpp:1213-1300). */
int feel_hip_space_create(feel_hip_ctx* ctx, int dim, int order, int ncomp, int64_t nverts,
                          const double* coords, int64_t nelems, const int32_t* elems,
                          const int32_t* elem2dof, int64_t ndof, feel_hip_space** out);
/* The same space with the DofTable built on the device from the mesh alone:
 * Feel++'s first-touch numbering (DofTable::build/buildDofMap,
 * feeldiscr/doftable.hpp:1802-1890,2364-2455; DofFromElement::add,
 * feeldiscr/doffromelement.hpp:492-530), identical to the host restatement. */
int feel_hip_space_create_from_mesh(feel_hip_ctx* ctx, int dim, int order, int ncomp, int64_t nverts,
                                    const double* coords, int64_t nelems, const int32_t* elems,
                                    feel_hip_space** out);
/* copies the space's elem2dof (nelements x nloc*ncomp) to host memory
 * (DofTable::localToGlobalIndices, doftable.hpp:732-750) */
int feel_hip_space_elem2dof(const feel_hip_space* s, int32_t* elem2dof);
int feel_hip_space_destroy(feel_hip_space* s);
int64_t feel_hip_space_ndof(const feel_hip_space* s);
int64_t feel_hip_space_nelements(const feel_hip_space* s);

/* ---- sparsity graph ------------------------------------------------------
 * Replaces MatrixPetsc::init(m, n, ..., graph) (feelalg/matrixpetsc.cpp:355-415):
 * uploads Feel++'s own CSR graph (GraphCSR::close, feelalg/graphcsr.cpp:663-1021)
 * and derives the owner-computes scatter map of the space's element entries. */
int feel_hip_graph_create(feel_hip_space* space, int64_t nrows, const int32_t* ia, const int32_t* ja,
                          feel_hip_graph** out);
/* The space's COUPLED pattern built on the device (Stencil::computeGraph,
 * feeldiscr/stencil.hpp:1502-1692; GraphCSR::close, feelalg/graphcsr.cpp:663-1021),
 * identical to the graph feel_hip_graph_create takes from the host. */
int feel_hip_graph_create_from_space(feel_hip_space* space, feel_hip_graph** out);
/* copies ia (nrows+1) and ja (nnz) to host memory (GraphCSR::ia/ja, graphcsr.hpp:61-81) */
int feel_hip_graph_pattern(const feel_hip_graph* g, int32_t* ia, int32_t* ja);
int feel_hip_graph_destroy(feel_hip_graph* g);
int64_t feel_hip_graph_nnz(const feel_hip_graph* g);

/* ---- matrices (MatrixSparse<double>, feelalg/matrixsparse.hpp:97) -------- */
int feel_hip_mat_create(feel_hip_graph* g, int precision, feel_hip_mat** out);   /* Backend::newMatrix, backend.hpp:349-440 */
/* FEEL_HIP_FP64_FP32: the FP32 values are always the FP64 values rounded (no
 * FP32 CSR copy is kept; get_values(FP32) and the FP32 solver image round). */
int feel_hip_mat_destroy(feel_hip_mat* A);
int feel_hip_mat_zero(feel_hip_mat* A);                                          /* zero(), matrixsparse.hpp:420 */
/* addMatrix(rows, nrows, cols, ncols, data) ADD semantics, data row-major (matrixsparse.hpp:496-498) */
int feel_hip_mat_add_matrix(feel_hip_mat* A, int nrows, const int32_t* rows, int ncols, const int32_t* cols,
                            const double* data);
int feel_hip_mat_get_values(feel_hip_mat* A, int precision, void* host_values);  /* nnz values in CSR order */
int feel_hip_mat_set_values(feel_hip_mat* A, const double* host_values);        /* FP64 values (and FP32 copy) */
int feel_hip_mat_mult(feel_hip_mat* A, const feel_hip_vec* x, feel_hip_vec* y);  /* multVector, matrixsparse.hpp:530 */

/* The rest of MatrixSparse's pure virtual interface on the device.
 * Structure of addMatrix(a, X, structure), Feel::MatrixStructure (feelalg/enums.hpp:194-200):
 * SAME requires X on the same pattern as Y; SUBSET and DIFFERENT accept any X whose
 * entries all lie in Y's pattern (the graph of Y cannot grow: other X fail with
 * FEEL_HIP_EPATTERN and leave Y unchanged). */
enum { FEEL_HIP_SAME_NONZERO_PATTERN = 0, FEEL_HIP_DIFFERENT_NONZERO_PATTERN = 1,
       FEEL_HIP_SUBSET_NONZERO_PATTERN = 3 };
/* MatNorm types: l1Norm (max column sum), linftyNorm (max row sum), Frobenius */
enum { FEEL_HIP_NORM_1 = 0, FEEL_HIP_NORM_FROBENIUS = 2, FEEL_HIP_NORM_INF = 3 };
int feel_hip_mat_scale(feel_hip_mat* A, double a);                                 /* scale(a), matrixsparse.hpp:524 */
int feel_hip_mat_axpy(feel_hip_mat* Y, double a, const feel_hip_mat* X, int structure); /* addMatrix(a, X), :512 */
int feel_hip_mat_copy(feel_hip_mat* dst, const feel_hip_mat* src);                 /* operator=(M), :593 */
int feel_hip_mat_diagonal(const feel_hip_mat* A, feel_hip_vec* d);                  /* diagonal(d), :602 */
/* transpose(Mt), :626: At(i,j) = A(j,i) into a matrix on the same, structurally
 * symmetric pattern (every graph built from one space is); EUNSUPPORTED otherwise */
int feel_hip_mat_transpose(const feel_hip_mat* A, feel_hip_mat* At);
/* l1Norm :697 (needs a structurally symmetric pattern), linftyNorm :709, Frobenius */
int feel_hip_mat_norm(const feel_hip_mat* A, int type, double* out);
/* energy(v, u, transpose), :673: v^T A u, or v^T A^T u */
int feel_hip_mat_energy(const feel_hip_mat* A, const feel_hip_vec* v, const feel_hip_vec* u, int transpose,
                        double* out);
/* set(i, j, value) over a block (INSERT_VALUES), :466; repeated entries: the last wins */
int feel_hip_mat_set_matrix(feel_hip_mat* A, int nrows, const int32_t* rows, int ncols, const int32_t* cols,
                            const double* data);
/* operator()(i, j), :591: m entries (rows[k], cols[k]); entries outside the pattern read 0 */
int feel_hip_mat_get_entries(const feel_hip_mat* A, int64_t m, const int32_t* rows, const int32_t* cols, double* out);
/* feel_hip_mat_add_matrix without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same ordered ADD and the same errors;
 * rows, columns, values and CSR positions are staged in one static device array
 * per GPU (1.5 MiB, held until the process ends), so nrows * ncols must not
 * exceed 65536 (FEEL_HIP_EINVAL otherwise).  Calls that overlap on one GPU share
 * that array and race: order them. */
int feel_hip_mat_add_matrix_noalloc(feel_hip_mat* A, int nrows, const int32_t* rows, int ncols, const int32_t* cols,
                                    const double* data);

/* ---- vectors (Vector<double>, feelalg/vector.hpp:143-576) ------------------ */
int feel_hip_vec_create(feel_hip_ctx* ctx, int64_t n, feel_hip_vec** out);      /* Backend::newVector, backend.hpp:571-592 */
int feel_hip_vec_destroy(feel_hip_vec* v);
int feel_hip_vec_set(feel_hip_vec* v, const double* host);                      /* host -> device */
int feel_hip_vec_get(const feel_hip_vec* v, double* host);                      /* device -> host */
int feel_hip_vec_zero(feel_hip_vec* v);
int feel_hip_vec_dot(const feel_hip_vec* a, const feel_hip_vec* b, double* out); /* dot(), vector.hpp:544 */
int feel_hip_vec_axpy(feel_hip_vec* y, double a, const feel_hip_vec* x);          /* add(a, v), vector.hpp:474 */
int feel_hip_vec_scale(feel_hip_vec* v, double a);                                /* scale(a), vector.hpp:576 */
int feel_hip_vec_copy(feel_hip_vec* dst, const feel_hip_vec* src);                /* operator=, vector.hpp:271 */
/* v[r] = values[r] (0 when values is null) for the rows of `rows`: the rhs part of
 * on() re-applied to a new right-hand side (feelvf/on.hpp:457-701) */
int feel_hip_vec_set_rows(feel_hip_vec* v, const feel_hip_rows* rows, const feel_hip_vec* values);
/* Reductions of Vector (vector.hpp:291-339): sum, l1Norm, l2Norm, min, max, linftyNorm */
enum { FEEL_HIP_VEC_SUM = 0, FEEL_HIP_VEC_L1 = 1, FEEL_HIP_VEC_L2 = 2, FEEL_HIP_VEC_MIN = 3, FEEL_HIP_VEC_MAX = 4,
       FEEL_HIP_VEC_LINF = 5 };
int feel_hip_vec_reduce(const feel_hip_vec* v, int kind, double* out);
int feel_hip_vec_set_constant(feel_hip_vec* v, double s);                         /* setConstant(s), vector.hpp:186 */
int feel_hip_vec_shift(feel_hip_vec* v, double s);                                /* add(s), vector.hpp:460 */
/* setVector(i, n, v), vector.hpp:443 (set(i, v) :438 for n = 1): v[idx[k]] = vals[k];
 * repeated indices: the last value wins */
int feel_hip_vec_set_indexed(feel_hip_vec* v, int64_t n, const int32_t* idx, const double* vals);
/* addVector(i, n, v, K, K2), vector.hpp:453 (add(i, v) :448 for n = 1): v[idx[k]] += vals[k];
 * repeated indices add up in order */
int feel_hip_vec_add_indexed(feel_hip_vec* v, int64_t n, const int32_t* idx, const double* vals);
/* addVector(V, A), vector.hpp:507: y += A x (x and y distinct) */
int feel_hip_vec_mult_add(feel_hip_vec* y, const feel_hip_mat* A, const feel_hip_vec* x);
int64_t feel_hip_vec_size(const feel_hip_vec* v);
/* raw device pointer (double*) of the vector's storage */
void* feel_hip_vec_device_ptr(feel_hip_vec* v);

/* ---- assembly -------------------------------------------------------------
 * form2 = integrate(elements(mesh), expr): BilinearForm::operator= / +=
 * (feelvf/bilinearform.hpp:1437-1468) -> Integrator::assemble
 * (feelvf/integrator.hpp:975-1211) -> element matrices + ADD scatter
 * (bilinearformcontext.hpp:772-889, matrixpetsc.cpp:919-937).
 * add = 0 replaces the values (operator=), add = 1 accumulates (operator+=). */
int feel_hip_assemble_bilinear(feel_hip_mat* A, const feel_hip_bilinear_desc* desc, int add);
/* form1 = integrate(elements(mesh), f*id(v)) (feelvf/linearform.hpp:1214) */
int feel_hip_assemble_linear(feel_hip_space* space, const feel_hip_linear_desc* desc, feel_hip_vec* b,
                             int add);

/* ---- Dirichlet ------------------------------------------------------------
 * on(_range=..., _rhs=l, _element=u, _expr=g) -> MatrixSparse::zeroRows
 * (feelvf/on.hpp:457-701, feelalg/matrixpetsc.cpp:1764-1850).  The row set is
 * uploaded once and stays resident. */
int feel_hip_rows_create(feel_hip_ctx* ctx, int64_t n, int64_t nrows, const int32_t* rows, feel_hip_rows** out);
int feel_hip_rows_destroy(feel_hip_rows* r);
/* values: g at every dof (length n) or null for g = 0 */
int feel_hip_mat_zero_rows(feel_hip_mat* A, const feel_hip_rows* rows, const feel_hip_vec* values,
                           feel_hip_vec* rhs, int mode, double diag);

/* ---- solve ----------------------------------------------------------------
 * BilinearFormBase::solve (feelvf/bilinearformbase.hpp:446-466) ->
 * Backend::solve -> SolverLinearPetsc::solve (feelalg/solverlinearpetsc.cpp:575-640)
 * with ksp-type=cg, pc-type=jacobi.  x holds the initial guess on entry. */
void feel_hip_solve_opts_default(feel_hip_solve_opts* o);
/* SpMV launch shape (block size, grid) of a graph, per precision.  The shape
 * decides how rows map to wavefronts and how many p.Ap block partials the CG
 * dots sum, so solutions are bitwise reproducible for a given shape.  By default
 * (-1) the first solve times every candidate once and keeps the fastest, which
 * can differ between processes; a caller that needs bitwise reproducibility
 * across processes pins a shape (0 .. feel_hip_graph_spmv_shapes()-1, 0 = the
 * default (1024, 2048)), or reads the tuned one and pins it in the next run.
 * No counterpart in the reference (PETSc's MatMult has no launch shape). */
int feel_hip_graph_spmv_shapes(void);
int feel_hip_graph_set_spmv_shape(feel_hip_graph* g, int precision, int shape);
int feel_hip_graph_spmv_shape(const feel_hip_graph* g, int precision, int* shape, int* block, int* grid);
int feel_hip_solve_pcg(feel_hip_mat* A, const feel_hip_vec* b, feel_hip_vec* x, const feel_hip_solve_opts* opts,
                       feel_hip_solve_data* out);

/* normL2(_range=elements(mesh), _expr=idv(u) - f) = sqrt(|integrate(inner(u - f))|)
 * (feelvf/norml2.hpp:37-57) and the H1 seminorm of u - f (feelvf/normsemih1.hpp):
 * f holds one function per component (null: zero); quad_order < 0 picks
 * 2k (+2 for SINSIN), the order of the squared integrand. */
int feel_hip_norm_l2(feel_hip_space* s, const feel_hip_vec* u, const feel_hip_function* f, int quad_order, double* out);
int feel_hip_norm_semi_h1(feel_hip_space* s, const feel_hip_vec* u, const feel_hip_function* f, int quad_order,
                          double* out);

/* ---- batched and continuously batched solves ------------------------------
 * Affine-parameter families A(mu) = sum_t mu_t A_t (the affine decomposition of
 * Feel++'s reduced-basis layer, mor/mor/feel/feelmor/modelcrbbase.hpp:1244, used
 * here for many full-order solves): the term matrices, rhs and homogeneous
 * Dirichlet rows stay resident; each submitted instance is assembled (A(mu), D^-1)
 * and solved with Jacobi-CG (FP32 inner solves inside FP64 refinement, or FP64),
 * all in-flight instances sharing kernel launches.  Instances are admitted in
 * submission order while the HBM budget allows (0 = unlimited).  No counterpart
 * in the reference, which solves one system at a time (SolverLinearPetsc::solve). */
typedef struct feel_hip_batch feel_hip_batch;
typedef struct {
    double rtol, atol;        /* outer FP64 test ||D^-1 r|| <= max(rtol ||D^-1 b||, atol) */
    double inner_rtol;        /* FP32: residual reduction between FP64 reliable updates */
    int maxit, inner_maxit, max_outer;
    int precision;            /* FEEL_HIP_FP32 (refined) or FEEL_HIP_FP64 */
    int64_t hbm_budget;       /* bytes of instance buffers allowed at once, 0 = unlimited */
    int use_graph;            /* capture each iteration chunk in a hipGraph */
    int group_spmv;           /* one SpMV pass per family group, A(mu) formed on the fly (FP32 and FP64) */
    int backfill;             /* 0: strict submission order; N > 0: requests that fit the budget may pass
                                 one that does not, until it has waited N steps */
} feel_hip_batch_opts;
void feel_hip_batch_opts_default(feel_hip_batch_opts* o);
int feel_hip_batch_create(feel_hip_ctx* ctx, const feel_hip_batch_opts* opts, feel_hip_batch** out);
int feel_hip_batch_destroy(feel_hip_batch* b);
/* terms: nterms assembled FP64 matrices on one graph; rhs: the family's load vector;
 * dirichlet: rows eliminated symmetrically with a unit diagonal (g = 0), or null */
int feel_hip_batch_add_family(feel_hip_batch* b, int nterms, feel_hip_mat* const* terms, const feel_hip_vec* rhs,
                              const feel_hip_rows* dirichlet, int* family);
int feel_hip_batch_submit(feel_hip_batch* b, int family, const double* mu, int64_t* request);
/* admits pending instances, then runs `iterations` CG iterations of every in-flight
 * instance and advances their refinement; n_active counts pending + in-flight */
int feel_hip_batch_step(feel_hip_batch* b, int iterations, int* n_active, int* n_done);
/* pops one finished request (returns 1), or returns 0 when none is waiting */
int feel_hip_batch_poll(feel_hip_batch* b, int64_t* request, feel_hip_solve_data* sd);
/* copies the solution (Feel++ dof numbering) and releases the request's buffers */
int feel_hip_batch_solution(feel_hip_batch* b, int64_t request, double* host_x);
int64_t feel_hip_batch_device_bytes(const feel_hip_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* FEELPP_HIP_H */
