// Host-side mesh, DofTable and sparsity graph for continuous Lagrange Pk on
// simplices.  These reproduce Feel++'s numbering and pattern bit for bit
// (first-touch dof numbering, COUPLED stencil, sorted CSR) with array-based
// algorithms instead of std::map / std::set.
#pragma once
#include <cstdint>
#include <iosfwd>
#include <string>
#include <vector>

namespace feelpp_amd {

struct SimplexMesh {
    int dim = 0;
    int64_t nv = 0, ne = 0;
    std::vector<double> coords;    // nv x dim
    std::vector<int32_t> elems;    // ne x (dim+1), vertex indices in point-id order
};

// Seeded synthetic meshes (SplitMix64).  Interior vertices move by
// h*U(-jitter,jitter) per coordinate, drawn in vertex-id order (x fastest).
SimplexMesh unitSquareMesh(int n, double jitter, uint64_t seed);
SimplexMesh kuhnBoxMesh(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                        uint64_t seed);

// Gmsh MSH 2.2 (ASCII) I/O with Feel++ import semantics (feelfilters/importergmsh.hpp):
// element id = insertion order of elements of the mesh dimension (:2631,:2851),
// point id = file node id (:2050; kept as an order-preserving relabelling).
void writeMsh22(const SimplexMesh& m, std::ostream& os);
// $ElementNodeData block of a Pk field (ncomp components, local dof nloc*c + l)
// with the values at each element's vertices, as Feel++'s Gmsh exporter writes
// nodal fields (feelfilters/exportergmsh_impl.hpp:1210-1324); elements are
// numbered 1..ne as writeMsh22 writes them.
void writeMsh22ElementNodeData(const SimplexMesh& m, int nldof, int ncomp, const int32_t* elem2dof,
                               const double* values, const std::string& name, double time, int step,
                               std::ostream& os);
SimplexMesh readMsh22(std::istream& is, int dim);

// Per-element random fields
std::vector<double> lognormalField(int64_t n, double mu, double sigma, uint64_t seed);
std::vector<double> uniformField(int64_t n, double lo, double hi, bool logscale, uint64_t seed);

struct DofTable {
    int dim = 0, order = 0, ncomp = 1;
    int nloc = 0;     // scalar local dofs
    int nldof = 0;    // nloc * ncomp
    int64_t ne = 0, ndof = 0;
    std::vector<int32_t> elem2dof;   // ne x nldof, local index nloc*c + l
};
DofTable buildDofTable(const SimplexMesh& m, int order, int ncomp);

struct GraphCSR {
    int64_t nrows = 0, nnz = 0;
    std::vector<int32_t> ia, ja;
};
GraphCSR buildGraph(const DofTable& dt, int nthreads = 0);

// Boundary dofs of on(boundaryfaces(mesh)) (sel = 0) or of the faces whose
// vertices all lie on x == 0 (sel = 1).  Returns a 0/1 flag per dof.
std::vector<uint8_t> boundaryDofs(const SimplexMesh& m, const DofTable& dt, int sel);

// dof -> (element, local dof) incidence in element order: entry = e*nldof + l
struct Incidence {
    std::vector<int64_t> ptr;     // ndof + 1
    std::vector<uint32_t> entry;
};
Incidence buildIncidence(const DofTable& dt);

// Sliced ELLPACK (SELL-C-sigma) image of a CSR matrix for the device SpMV:
// rows are sorted by decreasing length inside windows of `window` rows
// (perm: new -> old row, iperm: old -> new), grouped in slices of C = 64 rows
// (one wavefront, one lane per row) and stored column-major inside a slice:
// slot = slicePtr[s] + k*64 + lane.  Columns are renumbered with iperm so the
// solver works entirely in the permuted numbering; padding slots point at the
// row itself with a zero value (slot2csr = -1).
// Row order of the solver's SELL-C-sigma image (new -> old): reverse
// Cuthill-McKee when reorder is set (bandwidth reduction, so slices span few
// columns and x gathers stay local), then a stable sort by decreasing row
// length inside windows of `window` rows.  The image itself is built on the
// device.
std::vector<int32_t> sellPermutation(const GraphCSR& g, int window, bool reorder = true);

// Reverse Cuthill-McKee ordering of a symmetric CSR graph (new -> old)
std::vector<int32_t> reverseCuthillMcKee(const GraphCSR& g);

int defaultThreads();

}  // namespace feelpp_amd
