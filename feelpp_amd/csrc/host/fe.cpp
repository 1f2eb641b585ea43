// Lagrange Pk on simplices: node sets, nodal basis, quadrature and reference tensors.
//
// Node sets follow Feel++'s default Lagrange point set (PointSetFekete, which
// is the warp-blend family on tetrahedra: feelpoly/fekete.hpp:1251-1265,
// feelpoly/warpblend.hpp:50-138).  On edges the warp-blend family reduces to
// Gauss-Lobatto-Legendre points; for k <= 2 every family gives vertices and
// midpoints; for k = 3 the single face node is the face centroid.
//
// The nodal basis is obtained by inverting a generalised Vandermonde matrix
// (feelpoly/fe.hpp:132-161 does this on the Dubiner basis).  Here the
// spanning set is the homogeneous barycentric monomials lambda^alpha,
// |alpha| = k, and the inversion runs in long double.
#include "fe.hpp"

#include <cmath>
#include <stdexcept>

#include "im_simplex_tables.inc"

namespace feelpp_amd {

const int kTriEdge[3][2] = {{1, 2}, {2, 0}, {0, 1}};
const int kTetEdge[6][2] = {{1, 2}, {2, 0}, {0, 1}, {0, 3}, {1, 3}, {2, 3}};
const int kTetFace[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
const int kTetFaceEdges[4][3] = {{5, 4, 0}, {5, 3, 1}, {4, 3, 2}, {0, 1, 2}};

int maxSupportedOrder(int dim) { return dim == 2 ? 2 : 4; }

QuadratureRule quadrature(int dim, int order) {
    for (const auto& t : im_simplex_tables) {
        if (t.dim != dim || t.order != order) continue;
        QuadratureRule r;
        r.dim = dim; r.order = order; r.npts = t.npts;
        r.pts.resize((size_t)t.npts * dim);
        r.w.resize(t.npts);
        for (int q = 0; q < t.npts; ++q) {
            for (int c = 0; c < dim; ++c) r.pts[q * dim + c] = t.q[q * (dim + 1) + c];
            r.w[q] = t.q[q * (dim + 1) + dim];
        }
        return r;
    }
    throw std::invalid_argument("no simplex quadrature of order " + std::to_string(order) +
                                " in dimension " + std::to_string(dim));
}

static std::vector<double> gllInterior(int k) {
    switch (k) {
        case 2: return {0.0};
        case 3: return {-1.0 / std::sqrt(5.0), 1.0 / std::sqrt(5.0)};
        case 4: return {-std::sqrt(3.0 / 7.0), 0.0, std::sqrt(3.0 / 7.0)};
        default: return {};
    }
}

// Warp & blend (Warburton) face node of P4 tetrahedra as Feel++ builds it
// (feelpoly/warpblend.hpp:366-500 with alpha_2d[4] = 0.1001): the equispaced
// face-interior point with barycentrics (1/2,1/4,1/4) is displaced along its
// median by the blended 1D warp; returns the weight at the near vertex.
static double p4FaceNearWeight() {
    const int N = 4;
    double req[N + 1], rgl[N + 1];
    const auto gl = gllInterior(N);
    for (int i = 0; i <= N; ++i) req[i] = -1.0 + 2.0 * i / N;
    rgl[0] = -1; rgl[N] = 1;
    for (int i = 1; i < N; ++i) rgl[i] = gl[i - 1];
    // warp(r) = sum_i (rgl_i - req_i) l_i(r) / (1 - r^2), l_i the equispaced Lagrange basis
    auto warp = [&](double r) {
        double w = 0;
        for (int i = 0; i <= N; ++i) {
            double li = 1;
            for (int j = 0; j <= N; ++j)
                if (j != i) li *= (r - req[j]) / (req[i] - req[j]);
            w += (rgl[i] - req[i]) * li;
        }
        return w / (1 - r * r);
    };
    const double alpha = 0.1001, L1 = 0.5, L2 = 0.25, L3 = 0.25;
    // only the two warps with a component along the median through vertex 1 survive
    const double w2 = 4 * L3 * L1 * warp(L1 - L3) * (1 + alpha * alpha * L2 * L2);
    const double w3 = 4 * L1 * L2 * warp(L2 - L1) * (1 + alpha * alpha * L3 * L3);
    const double dy = std::sqrt(3.0) / 2 * (w2 - w3);
    const double y = (2 * L1 - L2 - L3) / std::sqrt(3.0) + dy;   // side-2 equilateral triangle
    return (std::sqrt(3.0) * y + 1) / 3;
}

std::vector<double> lagrangeNodes(int dim, int k) {
    if (dim < 2 || dim > 3 || k < 1 || k > maxSupportedOrder(dim))
        throw std::invalid_argument("unsupported Lagrange space P" + std::to_string(k) + " in " +
                                    std::to_string(dim) + "D");
    const double V3[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    std::vector<double> out;
    for (int v = 0; v <= dim; ++v)
        for (int c = 0; c < dim; ++c) out.push_back(V3[v][c]);
    const auto t = gllInterior(k);
    for (int e = 0; e < numEdges(dim); ++e) {
        const int* ab = dim == 2 ? kTriEdge[e] : kTetEdge[e];
        for (double s : t)
            for (int c = 0; c < dim; ++c)
                out.push_back(0.5 * (1.0 - s) * V3[ab[0]][c] + 0.5 * (1.0 + s) * V3[ab[1]][c]);
    }
    if (dim == 3 && k == 3)
        for (int f = 0; f < 4; ++f)
            for (int c = 0; c < 3; ++c)
                out.push_back((V3[kTetFace[f][0]][c] + V3[kTetFace[f][1]][c] + V3[kTetFace[f][2]][c]) / 3.0);
    if (dim == 3 && k == 4) {
        // face l-th point near the face's l-th vertex (equispaced interior lattice order,
        // feelpoly/equispaced.hpp:520-545), then the interior centroid
        const double tb = p4FaceNearWeight(), ts = 0.5 * (1.0 - tb);
        for (int f = 0; f < 4; ++f)
            for (int l = 0; l < 3; ++l)
                for (int c = 0; c < 3; ++c)
                    out.push_back(tb * V3[kTetFace[f][l]][c] + ts * V3[kTetFace[f][(l + 1) % 3]][c] +
                                  ts * V3[kTetFace[f][(l + 2) % 3]][c]);
        for (int c = 0; c < 3; ++c) out.push_back(-0.5);
    }
    return out;
}

namespace {

// barycentric coordinates on the [-1,1] simplex and their constant gradients
void bary(int dim, const double* x, long double* lam) {
    long double s = 0;
    for (int c = 0; c < dim; ++c) { lam[c + 1] = 0.5L * (1.0L + x[c]); s += lam[c + 1]; }
    lam[0] = 1.0L - s;
}
long double dbary(int v, int c) { return v == 0 ? -0.5L : (v == c + 1 ? 0.5L : 0.0L); }

std::vector<std::vector<int>> multiIndices(int dim, int k) {
    std::vector<std::vector<int>> out;
    std::vector<int> a(dim + 1, 0);
    // all alpha with |alpha| = k over dim+1 barycentric coordinates
    std::vector<int> cur(dim + 1, 0);
    auto rec = [&](auto&& self, int pos, int left) -> void {
        if (pos == dim) { cur[pos] = left; out.push_back(cur); return; }
        for (int v = left; v >= 0; --v) { cur[pos] = v; self(self, pos + 1, left - v); }
    };
    rec(rec, 0, k);
    return out;
}

long double ipow(long double x, int e) {
    long double r = 1;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

struct NodalBasis {
    int dim, k, n;
    std::vector<std::vector<int>> alpha;
    std::vector<long double> C;   // C[m*n+i]: coefficient of monomial m in phi_i
};

NodalBasis buildBasis(int dim, int k) {
    NodalBasis B;
    B.dim = dim; B.k = k;
    B.alpha = multiIndices(dim, k);
    B.n = (int)B.alpha.size();
    const int n = B.n;
    const auto nodes = lagrangeNodes(dim, k);
    if ((int)nodes.size() != n * dim) throw std::logic_error("node count mismatch");
    // V[j][m] = lambda(x_j)^alpha_m ; phi_i = sum_m C[m][i] mono_m with V C = I
    std::vector<long double> V((size_t)n * n), C((size_t)n * n, 0.0L);
    long double lam[4];
    for (int j = 0; j < n; ++j) {
        bary(dim, &nodes[j * dim], lam);
        for (int m = 0; m < n; ++m) {
            long double v = 1;
            for (int t = 0; t <= dim; ++t) v *= ipow(lam[t], B.alpha[m][t]);
            V[(size_t)j * n + m] = v;
        }
        C[(size_t)j * n + j] = 1.0L;
    }
    // Gauss-Jordan elimination with partial pivoting on [V | I]
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (fabsl(V[(size_t)r * n + col]) > fabsl(V[(size_t)piv * n + col])) piv = r;
        if (piv != col)
            for (int c = 0; c < n; ++c) {
                std::swap(V[(size_t)piv * n + c], V[(size_t)col * n + c]);
                std::swap(C[(size_t)piv * n + c], C[(size_t)col * n + c]);
            }
        const long double d = V[(size_t)col * n + col];
        for (int c = 0; c < n; ++c) { V[(size_t)col * n + c] /= d; C[(size_t)col * n + c] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const long double f = V[(size_t)r * n + col];
            if (f == 0) continue;
            for (int c = 0; c < n; ++c) {
                V[(size_t)r * n + c] -= f * V[(size_t)col * n + c];
                C[(size_t)r * n + c] -= f * C[(size_t)col * n + c];
            }
        }
    }
    B.C = std::move(C);
    return B;
}

// values and reference gradients of all basis functions at x (long double)
void evalBasis(const NodalBasis& B, const double* x, long double* phi, long double* dphi) {
    const int n = B.n, dim = B.dim;
    long double lam[4];
    bary(dim, x, lam);
    std::vector<long double> mv(n), md((size_t)n * dim);
    for (int m = 0; m < n; ++m) {
        const auto& a = B.alpha[m];
        long double v = 1;
        for (int t = 0; t <= dim; ++t) v *= ipow(lam[t], a[t]);
        mv[m] = v;
        for (int c = 0; c < dim; ++c) {
            long double g = 0;
            for (int t = 0; t <= dim; ++t) {
                if (a[t] == 0 || dbary(t, c) == 0) continue;
                long double term = a[t] * ipow(lam[t], a[t] - 1) * dbary(t, c);
                for (int s = 0; s <= dim; ++s)
                    if (s != t) term *= ipow(lam[s], a[s]);
                g += term;
            }
            md[(size_t)m * dim + c] = g;
        }
    }
    for (int i = 0; i < n; ++i) {
        long double s = 0;
        for (int m = 0; m < n; ++m) s += B.C[(size_t)m * n + i] * mv[m];
        phi[i] = s;
        for (int c = 0; c < dim; ++c) {
            long double g = 0;
            for (int m = 0; m < n; ++m) g += B.C[(size_t)m * n + i] * md[(size_t)m * dim + c];
            dphi[(size_t)i * dim + c] = g;
        }
    }
}

}  // namespace

Tabulation tabulate(int dim, int k, const std::vector<double>& pts) {
    NodalBasis B = buildBasis(dim, k);
    Tabulation T;
    T.dim = dim; T.k = k; T.n = B.n; T.npts = (int)(pts.size() / dim);
    T.phi.resize((size_t)T.n * T.npts);
    T.dphi.resize((size_t)T.n * T.npts * dim);
    std::vector<long double> p(B.n), dp((size_t)B.n * dim);
    for (int q = 0; q < T.npts; ++q) {
        evalBasis(B, &pts[(size_t)q * dim], p.data(), dp.data());
        for (int i = 0; i < B.n; ++i) {
            T.phi[(size_t)i * T.npts + q] = (double)p[i];
            for (int c = 0; c < dim; ++c) T.dphi[((size_t)i * T.npts + q) * dim + c] = (double)dp[(size_t)i * dim + c];
        }
    }
    return T;
}

ReferenceTensors referenceTensors(int dim, int k, int qorder) {
    NodalBasis B = buildBasis(dim, k);
    QuadratureRule Q = quadrature(dim, qorder);
    const int n = B.n;
    ReferenceTensors T;
    T.dim = dim; T.k = k; T.n = n; T.qorder = qorder;
    std::vector<long double> R((size_t)dim * dim * n * n, 0.0L), M((size_t)n * n, 0.0L), m(n, 0.0L);
    std::vector<long double> p(n), dp((size_t)n * dim);
    for (int q = 0; q < Q.npts; ++q) {
        evalBasis(B, &Q.pts[(size_t)q * dim], p.data(), dp.data());
        const long double w = Q.w[q];
        for (int i = 0; i < n; ++i) {
            m[i] += w * p[i];
            for (int j = 0; j < n; ++j) {
                M[(size_t)i * n + j] += w * p[i] * p[j];
                for (int a = 0; a < dim; ++a)
                    for (int b = 0; b < dim; ++b)
                        R[((size_t)(a * dim + b) * n + i) * n + j] += w * dp[(size_t)i * dim + a] * dp[(size_t)j * dim + b];
            }
        }
    }
    T.R.assign(R.begin(), R.end());
    T.M.assign(M.begin(), M.end());
    T.m.assign(m.begin(), m.end());
    return T;
}

}  // namespace feelpp_amd
