/*
 * feelpp_host.h -- C-ABI of the host side: seeded meshes, Feel++ DofTable
 * numbering, Stencil/GraphCSR sparsity pattern and boundary-dof sets, in the
 * exact form Feel++ produces them (so that the device layer scatters into
 * Feel++'s own graph).  Reference interfaces cited file:line relative to the
 * Feel++ source tree.  All functions return >= 0 on success, < 0 on error
 * (feel_host_last_error() explains).
 */
#ifndef FEELPP_HOST_H
#define FEELPP_HOST_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* feel_host_last_error(void);

/* Seeded synthetic meshes (SplitMix64); coords nv x dim, elems ne x (dim+1).
 * unit square: (n+1)^2 vertices, 2n^2 triangles; Kuhn box: 6 tets per cube. */
int feel_mesh_unit_square(int n, double jitter, uint64_t seed, double* coords, int32_t* elems);
int feel_mesh_kuhn_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter, uint64_t seed,
                       double* coords, int32_t* elems);
/* Gmsh MSH 2.2 writer/reader with Feel++ import semantics (feelfilters/importergmsh.hpp:2040-2066,2851-2883) */
int feel_mesh_write_msh22(const char* path, int dim, int64_t nv, const double* coords, int64_t ne, const int32_t* elems);
/* mesh followed by an $ElementNodeData block of a Pk field (values at each element's
 * vertices, ncomp per vertex), the layout of Feel++'s Gmsh exporter
 * (feelfilters/exportergmsh_impl.hpp:1210-1324) */
int feel_mesh_write_msh22_solution(const char* path, int dim, int64_t nv, const double* coords, int64_t ne,
                                   const int32_t* elems, int order, int ncomp, const int32_t* elem2dof,
                                   const double* values, const char* name, double time, int step);
/* returns nv and ne through pointers; call with null arrays first to size them */
int feel_mesh_read_msh22(const char* path, int dim, int64_t* nv, double* coords, int64_t* ne, int32_t* elems);

int feel_field_lognormal(int64_t n, double mu, double sigma, uint64_t seed, double* out);
int feel_field_uniform(int64_t n, double lo, double hi, int logscale, uint64_t seed, double* out);

/* DofTable::build for Pch<order> (ncomp = 1) / Pchv<order> (ncomp = dim)
 * (feeldiscr/doftable.hpp:1802-1890, 2364-2455).  elem2dof: ne x nloc*ncomp.
 * Returns ndof. */
int64_t feel_doftable_build(int dim, int order, int ncomp, int64_t nv, int64_t ne, const int32_t* elems,
                            int32_t* elem2dof);

/* Stencil::computeGraph + GraphCSR::close (feeldiscr/stencil.hpp:1502-1692,
 * feelalg/graphcsr.cpp:663-1021).  Builds into a handle, then copy out. */
typedef struct feel_host_graph feel_host_graph;
int feel_graph_build(int nldof, int64_t ne, int64_t ndof, const int32_t* elem2dof, feel_host_graph** out);
int64_t feel_graph_nnz(const feel_host_graph* g);
int feel_graph_copy(const feel_host_graph* g, int32_t* ia, int32_t* ja);
void feel_graph_free(feel_host_graph* g);

/* dofs of on(_range=boundaryfaces(mesh)) (sel 0) or of faces on x == 0 (sel 1)
 * (feelvf/on.hpp:457-701).  flags: one byte per dof.  Returns the count. */
int64_t feel_boundary_dofs(int dim, int order, int ncomp, int64_t nv, const double* coords, int64_t ne,
                           const int32_t* elems, const int32_t* elem2dof, int64_t ndof, int sel, uint8_t* flags);

/* Reference element data: nodes (nloc x dim), quadrature (npts x dim, npts),
 * and reference tensors R (dim*dim x nloc x nloc), M (nloc x nloc), m (nloc). */
int feel_lagrange_nodes(int dim, int order, double* nodes);
int feel_quadrature(int dim, int order, double* pts, double* w);
int feel_reference_tensors(int dim, int order, int qorder, double* R, double* M, double* m);

/* Condition number of the Jacobi-preconditioned operator estimated from k CG
 * iterations, given as k + 1 pairs (gamma_j = (r_j, z_j), pAp_j = (p_j, A p_j)):
 * the extreme eigenvalues of the Lanczos tridiagonal of the run.  No Feel++
 * counterpart in this path (PETSc offers it as KSPComputeExtremeSingularValues);
 * the FP32 solve uses it to decide when the matrix is too ill-conditioned for
 * FP32 and the rest of the solve must run in FP64. */
int feel_cg_condition_estimate(const double* gamma_pap, int k, double* kappa);
/* Whether that estimate exceeds limit (the test the FP32 solve makes): the
 * largest eigenvalue to 1 % and one Sturm count, so it agrees with
 * estimate > limit except within about 1 % of the limit.  No Feel++
 * counterpart. */
int feel_cg_condition_exceeds(const double* gamma_pap, int k, double limit, int* exceeds);

#ifdef __cplusplus
}
#endif
#endif
