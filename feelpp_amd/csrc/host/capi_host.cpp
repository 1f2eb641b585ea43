// C-ABI of the host side (include/feelpp_host.h).
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "feelpp_host.h"
#include "fe.hpp"
#include "space.hpp"
#include "krylov.hpp"

using namespace feelpp_amd;

namespace {
thread_local std::string gHostError;
int64_t fail(const std::string& m) {
    gHostError = m;
    return -1;
}
SimplexMesh view(int dim, int64_t nv, const double* coords, int64_t ne, const int32_t* elems) {
    SimplexMesh m;
    m.dim = dim; m.nv = nv; m.ne = ne;
    m.coords.assign(coords, coords + nv * dim);
    m.elems.assign(elems, elems + ne * (dim + 1));
    return m;
}
}  // namespace

#define HOST_TRY try {
#define HOST_CATCH \
    }              \
    catch (const std::exception& e) { return fail(e.what()); }

struct feel_host_graph {
    GraphCSR g;
};

extern "C" {

const char* feel_host_last_error(void) { return gHostError.c_str(); }

int feel_mesh_unit_square(int n, double jitter, uint64_t seed, double* coords, int32_t* elems) {
    HOST_TRY
    if (n < 1 || !coords || !elems) return (int)fail("bad argument");
    SimplexMesh m = unitSquareMesh(n, jitter, seed);
    std::memcpy(coords, m.coords.data(), sizeof(double) * m.coords.size());
    std::memcpy(elems, m.elems.data(), sizeof(int32_t) * m.elems.size());
    return 0;
    HOST_CATCH
}

int feel_mesh_kuhn_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter, uint64_t seed,
                       double* coords, int32_t* elems) {
    HOST_TRY
    if (nx < 1 || ny < 1 || nz < 1 || !coords || !elems) return (int)fail("bad argument");
    SimplexMesh m = kuhnBoxMesh(nx, ny, nz, lx, ly, lz, jitter, seed);
    std::memcpy(coords, m.coords.data(), sizeof(double) * m.coords.size());
    std::memcpy(elems, m.elems.data(), sizeof(int32_t) * m.elems.size());
    return 0;
    HOST_CATCH
}

int feel_mesh_write_msh22(const char* path, int dim, int64_t nv, const double* coords, int64_t ne,
                          const int32_t* elems) {
    HOST_TRY
    std::ofstream os(path);
    if (!os) return (int)fail(std::string("cannot open ") + path);
    writeMsh22(view(dim, nv, coords, ne, elems), os);
    return os.good() ? 0 : (int)fail("write failed");
    HOST_CATCH
}

int feel_mesh_write_msh22_solution(const char* path, int dim, int64_t nv, const double* coords, int64_t ne,
                                   const int32_t* elems, int order, int ncomp, const int32_t* elem2dof,
                                   const double* values, const char* name, double time, int step) {
    HOST_TRY
    if (!path || !coords || !elems || !elem2dof || !values || ncomp < 1 || ncomp > 3)
        return (int)fail("bad argument");
    if (order < 1 || order > maxSupportedOrder(dim)) return (int)fail("unsupported order");
    std::ofstream os(path);
    if (!os) return (int)fail(std::string("cannot open ") + path);
    SimplexMesh m = view(dim, nv, coords, ne, elems);
    writeMsh22(m, os);
    writeMsh22ElementNodeData(m, nLocalDof(dim, order) * ncomp, ncomp, elem2dof, values, name ? name : "u", time, step,
                              os);
    return os.good() ? 0 : (int)fail("write failed");
    HOST_CATCH
}

int feel_mesh_read_msh22(const char* path, int dim, int64_t* nv, double* coords, int64_t* ne, int32_t* elems) {
    HOST_TRY
    std::ifstream is(path);
    if (!is) return (int)fail(std::string("cannot open ") + path);
    SimplexMesh m = readMsh22(is, dim);
    if (nv) *nv = m.nv;
    if (ne) *ne = m.ne;
    if (coords) std::memcpy(coords, m.coords.data(), sizeof(double) * m.coords.size());
    if (elems) std::memcpy(elems, m.elems.data(), sizeof(int32_t) * m.elems.size());
    return 0;
    HOST_CATCH
}

int feel_field_lognormal(int64_t n, double mu, double sigma, uint64_t seed, double* out) {
    HOST_TRY
    auto v = lognormalField(n, mu, sigma, seed);
    std::memcpy(out, v.data(), sizeof(double) * n);
    return 0;
    HOST_CATCH
}

int feel_field_uniform(int64_t n, double lo, double hi, int logscale, uint64_t seed, double* out) {
    HOST_TRY
    auto v = uniformField(n, lo, hi, logscale != 0, seed);
    std::memcpy(out, v.data(), sizeof(double) * n);
    return 0;
    HOST_CATCH
}

int64_t feel_doftable_build(int dim, int order, int ncomp, int64_t nv, int64_t ne, const int32_t* elems,
                            int32_t* elem2dof) {
    HOST_TRY
    if (!elems || !elem2dof || nv <= 0 || ne <= 0) return fail("bad argument");
    SimplexMesh m;
    m.dim = dim; m.nv = nv; m.ne = ne;
    m.elems.assign(elems, elems + ne * (dim + 1));
    for (int32_t v : m.elems)
        if (v < 0 || v >= nv) return fail("vertex index out of range");
    DofTable dt = buildDofTable(m, order, ncomp);
    std::memcpy(elem2dof, dt.elem2dof.data(), sizeof(int32_t) * dt.elem2dof.size());
    return dt.ndof;
    HOST_CATCH
}

int feel_graph_build(int nldof, int64_t ne, int64_t ndof, const int32_t* elem2dof, feel_host_graph** out) {
    HOST_TRY
    if (!elem2dof || !out || nldof <= 0 || ne <= 0 || ndof <= 0) return (int)fail("bad argument");
    DofTable dt;
    dt.nldof = nldof; dt.ne = ne; dt.ndof = ndof;
    dt.elem2dof.assign(elem2dof, elem2dof + ne * nldof);
    for (int32_t d : dt.elem2dof)
        if (d < 0 || d >= ndof) return (int)fail("dof index out of range");
    auto* g = new feel_host_graph();
    g->g = buildGraph(dt);
    *out = g;
    return 0;
    HOST_CATCH
}

int64_t feel_graph_nnz(const feel_host_graph* g) { return g ? g->g.nnz : -1; }

int feel_graph_copy(const feel_host_graph* g, int32_t* ia, int32_t* ja) {
    if (!g) return (int)fail("null graph");
    if (ia) std::memcpy(ia, g->g.ia.data(), sizeof(int32_t) * g->g.ia.size());
    if (ja) std::memcpy(ja, g->g.ja.data(), sizeof(int32_t) * g->g.ja.size());
    return 0;
}

void feel_graph_free(feel_host_graph* g) { delete g; }

int64_t feel_boundary_dofs(int dim, int order, int ncomp, int64_t nv, const double* coords, int64_t ne,
                           const int32_t* elems, const int32_t* elem2dof, int64_t ndof, int sel, uint8_t* flags) {
    HOST_TRY
    SimplexMesh m = view(dim, nv, coords, ne, elems);
    DofTable dt;
    dt.dim = dim; dt.order = order; dt.ncomp = ncomp;
    dt.nloc = nLocalDof(dim, order);
    dt.nldof = dt.nloc * ncomp;
    dt.ne = ne; dt.ndof = ndof;
    dt.elem2dof.assign(elem2dof, elem2dof + ne * dt.nldof);
    auto f = boundaryDofs(m, dt, sel);
    int64_t cnt = 0;
    for (int64_t i = 0; i < ndof; ++i) { flags[i] = f[i]; cnt += f[i]; }
    return cnt;
    HOST_CATCH
}

int feel_lagrange_nodes(int dim, int order, double* nodes) {
    HOST_TRY
    auto v = lagrangeNodes(dim, order);
    if (nodes) std::memcpy(nodes, v.data(), sizeof(double) * v.size());
    return (int)(v.size() / dim);
    HOST_CATCH
}

int feel_quadrature(int dim, int order, double* pts, double* w) {
    HOST_TRY
    auto q = quadrature(dim, order);
    if (pts) std::memcpy(pts, q.pts.data(), sizeof(double) * q.pts.size());
    if (w) std::memcpy(w, q.w.data(), sizeof(double) * q.w.size());
    return q.npts;
    HOST_CATCH
}

int feel_reference_tensors(int dim, int order, int qorder, double* R, double* M, double* m) {
    HOST_TRY
    auto T = referenceTensors(dim, order, qorder);
    if (R) std::memcpy(R, T.R.data(), sizeof(double) * T.R.size());
    if (M) std::memcpy(M, T.M.data(), sizeof(double) * T.M.size());
    if (m) std::memcpy(m, T.m.data(), sizeof(double) * T.m.size());
    return T.n;
    HOST_CATCH
}

int feel_cg_condition_estimate(const double* gamma_pap, int k, double* kappa) {
    HOST_TRY
    if (!gamma_pap || !kappa || k < 0) return (int)fail("bad argument");
    for (int j = 0; j < k; ++j)
        if (!(gamma_pap[2 * j] > 0) || !(gamma_pap[2 * j + 1] > 0))
            return (int)fail("coefficients must be positive");
    *kappa = cgConditionEstimate(gamma_pap, k);
    return 0;
    HOST_CATCH
}

int feel_cg_condition_exceeds(const double* gamma_pap, int k, double limit, int* exceeds) {
    HOST_TRY
    if (!gamma_pap || !exceeds || k < 0 || !(limit > 0)) return (int)fail("bad argument");
    for (int j = 0; j < k; ++j)
        if (!(gamma_pap[2 * j] > 0) || !(gamma_pap[2 * j + 1] > 0))
            return (int)fail("coefficients must be positive");
    *exceeds = cgConditionExceeds(gamma_pap, k, limit) ? 1 : 0;
    return 0;
    HOST_CATCH
}

}  // extern "C"
