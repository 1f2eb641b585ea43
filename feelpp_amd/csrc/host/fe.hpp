// Reference-element data for continuous Lagrange Pk on simplices: quadrature
// rules, node sets, nodal basis and the reference tensors the device kernels
// contract with per-element geometric factors.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace feelpp_amd {

// Local numbering of the reference simplex (feelmesh/simplexordering.hpp:107-112,
// :256-264, :297-303, :345-350).
extern const int kTriEdge[3][2];
extern const int kTetEdge[6][2];
extern const int kTetFace[4][3];
extern const int kTetFaceEdges[4][3];

inline int numEdges(int dim) { return dim == 2 ? 3 : 6; }
inline int numFaces(int dim) { return dim == 3 ? 4 : 0; }
inline int dofPerEdge(int k) { return k - 1; }
inline int dofPerFace(int k) { return (k - 1) * (k - 2) / 2; }
inline int dofPerVolume(int k) { return (k - 1) * (k - 2) * (k - 3) / 6; }
inline int nLocalDof(int dim, int k) {
    return dim == 2 ? (k + 1) * (k + 2) / 2 : (k + 1) * (k + 2) * (k + 3) / 6;
}
int maxSupportedOrder(int dim);

struct QuadratureRule {
    int dim = 0, order = 0, npts = 0;
    std::vector<double> pts;   // npts x dim, on the [-1,1] reference simplex
    std::vector<double> w;     // npts
};
// "im(dim,order,simplex)"; throws std::invalid_argument for unknown orders
QuadratureRule quadrature(int dim, int order);

// Lagrange nodes (npts x dim); local order vertices, edges, faces, interior.
std::vector<double> lagrangeNodes(int dim, int k);

// Tabulation of the nodal basis: phi[i*npts+q], dphi[(i*npts+q)*dim+d]
struct Tabulation {
    int dim = 0, k = 0, n = 0, npts = 0;
    std::vector<double> phi, dphi;
};
Tabulation tabulate(int dim, int k, const std::vector<double>& pts);

// Reference tensors on the reference element for quadrature order qo:
//   R[a*dim+b][i*n+j] = sum_q w_q dphi_i/dx_a dphi_j/dx_b   (a,b < dim)
//   M[i*n+j]          = sum_q w_q phi_i phi_j
//   m[i]              = sum_q w_q phi_i
// accumulated in extended precision.
struct ReferenceTensors {
    int dim = 0, k = 0, n = 0, qorder = 0;
    std::vector<double> R, M, m;
};
ReferenceTensors referenceTensors(int dim, int k, int qorder);

}  // namespace feelpp_amd
