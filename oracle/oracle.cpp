/*
 * oracle.cpp -- FP64 CPU restatement of Feel++'s form assembly and Jacobi-CG
 * solve for Lagrange Pk on simplices.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product library (feelpp_amd) never links or calls it.
 *
 * Every function follows the Feel++ algorithm it restates (file:line cited,
 * paths relative to the Feel++ source tree).  The restatement is deliberately
 * naive (std::map / std::set / per-entry quadrature loops / binary-search
 * scatter) so that it reads like the reference and shares no code with the
 * GPU path.
 *
 * Parity pins (what checks each part independently of this file):
 *   quadrature      exactness on monomials vs closed forms, weight sums
 *   Lagrange nodes  data/operatorlagrangep1/simplex-3d-lagrange-order{2,3-warpblend}.msh
 *   basis           nodal property, partition of unity, P1 closed form
 *   dof numbering   hand-derived first-touch numbering on tiny meshes
 *   CSR pattern     brute-force neighbour sets, row-count stencils 65/27/19
 *   matrices        P1 closed forms, energy(1,1)=k*measure, zero row sums,
 *                   rigid-body modes for elasticity
 *   solve           patch tests, scipy.sparse direct solve on small meshes
 * The Krylov trajectory (iteration counts) has no reference fixture: parity
 * unpinned for iteration counts; the solution itself is pinned.
 * P4 tetrahedra: the face-interior node position (warp & blend) and the
 * numbering of the three dofs of a face have no Feel++ fixture (the reference
 * ships 3D node sets up to order 3 only): parity unpinned for those two; they
 * are checked only against what the mathematics fixes (nodal property,
 * quartic reproduction, one dof per physical point across faces).
 * orc_assemble_scalar_mt / orc_pcg_mt are the threaded CPU baseline of the
 * benchmark, not checkers.
 */
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "im_simplex_tables.inc"

namespace {

// ---------------------------------------------------------------------------
// SplitMix64 (published by S. Vigna); the seeded synthetic inputs use it.
// ---------------------------------------------------------------------------
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }   // [0,1)
};

// ---------------------------------------------------------------------------
// Reference simplex numbering, feelmesh/simplexordering.hpp:107-112 (triangle
// edges) and :256-264 (tetra edges), :297-303 (tetra faces), :345-350 (f2e).
// ---------------------------------------------------------------------------
const int TRI_E2P[3][2] = {{1, 2}, {2, 0}, {0, 1}};
const int TET_E2P[6][2] = {{1, 2}, {2, 0}, {0, 1}, {0, 3}, {1, 3}, {2, 3}};
const int TET_F2P[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
const int TET_F2E[4][3] = {{5, 4, 0}, {5, 3, 1}, {4, 3, 2}, {0, 1, 2}};

int nvert(int dim) { return dim + 1; }
int nedges(int dim) { return dim == 2 ? 3 : 6; }
int nfaces3(int dim) { return dim == 3 ? 4 : 0; }

// Dofs per entity for continuous Lagrange Pk on simplices.
int dof_per_edge(int k) { return k - 1; }
int dof_per_face(int k) { return (k - 1) * (k - 2) / 2; }          // triangle interior
int dof_per_vol(int k) { return (k - 1) * (k - 2) * (k - 3) / 6; }  // tetra interior
int nloc_scalar(int dim, int k) {
    return dim == 2 ? (k + 1) * (k + 2) / 2 : (k + 1) * (k + 2) * (k + 3) / 6;
}

// ---------------------------------------------------------------------------
// Lagrange node sets on the [-1,1] reference simplex.  Local order: vertices,
// edges (e2p order, points from the edge's first vertex to its second),
// faces, interior (feelpoly/lagrange.hpp, point sets warpblend.hpp:50-138 via
// fekete.hpp:1251-1265).  Edge points are Gauss-Lobatto points for k>=3
// (the warp-blend family reduces to GLL on edges); for k<=2 they coincide
// with equispaced points.
// ---------------------------------------------------------------------------
std::vector<double> gll_interior(int k) {
    // interior Gauss-Lobatto-Legendre points on [-1,1] for polynomial degree k
    if (k == 2) return {0.0};
    if (k == 3) return {-1.0 / std::sqrt(5.0), 1.0 / std::sqrt(5.0)};
    if (k == 4) return {-std::sqrt(3.0 / 7.0), 0.0, std::sqrt(3.0 / 7.0)};
    return {};
}

// Warp & blend face warp of PointSetWarpBlend (feelpoly/warpblend.hpp:366-500)
// evaluated for the face-interior points of P4 (equispaced face barycentrics
// (1/2,1/4,1/4)): warpFactor (:383-407) interpolates the GLL - equispaced
// displacement divided by 1-r^2; the 2D deformation (:409-442) adds
// blend_i * warp_i * (1 + (alpha_2d[4] L_i)^2) along the three edge directions,
// alpha_2d[4] = 0.1001 (:334).  By symmetry the point stays on the face median;
// returns its barycentric weight at the near vertex.
double p4_face_weight() {
    const int N = 4;
    const double xeq[5] = {-1, -0.5, 0, 0.5, 1};
    const double gll[5] = {-1, -std::sqrt(3.0 / 7.0), 0, std::sqrt(3.0 / 7.0), 1};
    auto warp = [&](double r) {
        double w = 0;
        for (int i = 0; i <= N; ++i) {
            double d = gll[i] - xeq[i];
            for (int j = 1; j < N; ++j)
                if (j != i) d *= (r - xeq[j]) / (xeq[i] - xeq[j]);
            if (i != 0) d *= -1.0 / (xeq[i] - xeq[0]);
            if (i != N) d *= 1.0 / (xeq[i] - xeq[N]);
            w += d;
        }
        return w;
    };
    const double alpha = 0.1001;
    const double L[3] = {0.5, 0.25, 0.25};
    double dy = 0;
    for (int i = 0; i < 3; ++i) {
        const double blend = 4 * L[(i + 1) % 3] * L[(i + 2) % 3];
        const double wi = blend * warp(L[(i + 2) % 3] - L[(i + 1) % 3]) * (1 + alpha * alpha * L[i] * L[i]);
        dy += std::sin(2.0 * i * M_PI / 3.0) * wi;
    }
    // equilateral triangle of side 2, L[0] at (0, 2/sqrt3): y = (3 L0 - 1)/sqrt3
    const double y = (3 * L[0] - 1) / std::sqrt(3.0) + dy;
    return (std::sqrt(3.0) * y + 1) / 3;
}

// Test hook: a positive value replaces the P4 face-interior weight above (0.5
// is the equispaced lattice point), so that a test can solve one problem in
// two nodal bases of the same P4 space; 0 restores the warp & blend weight.
double g_p4_face_weight = 0.0;

bool lagrange_nodes(int dim, int k, std::vector<double>& nodes) {
    nodes.clear();
    if (k < 1 || k > 4) return false;
    if (dim == 2 && k > 2) return false;
    std::vector<std::vector<double>> V;
    if (dim == 2) V = {{-1, -1}, {1, -1}, {-1, 1}};
    else V = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    for (auto& v : V) nodes.insert(nodes.end(), v.begin(), v.end());
    auto t = gll_interior(k);
    int ne = nedges(dim);
    for (int e = 0; e < ne; ++e) {
        int a = dim == 2 ? TRI_E2P[e][0] : TET_E2P[e][0];
        int b = dim == 2 ? TRI_E2P[e][1] : TET_E2P[e][1];
        for (double s : t)
            for (int c = 0; c < dim; ++c)
                nodes.push_back(V[a][c] * (1 - s) / 2 + V[b][c] * (1 + s) / 2);
    }
    if (dim == 3 && k == 3) {
        for (int f = 0; f < 4; ++f)
            for (int c = 0; c < 3; ++c)
                nodes.push_back((V[TET_F2P[f][0]][c] + V[TET_F2P[f][1]][c] + V[TET_F2P[f][2]][c]) / 3.0);
    }
    if (dim == 3 && k == 4) {
        // three points per face; the l-th sits near the face's l-th vertex (equispaced
        // interior triangle lattice order, equispaced.hpp:520-545, mapped by
        // pt_to_face_tetrahedron :671-690)
        const double tb = g_p4_face_weight > 0 ? g_p4_face_weight : p4_face_weight(), ts = (1 - tb) / 2;
        for (int f = 0; f < 4; ++f)
            for (int l = 0; l < 3; ++l)
                for (int c = 0; c < 3; ++c) {
                    double x = 0;
                    for (int m = 0; m < 3; ++m) x += (m == l ? tb : ts) * V[TET_F2P[f][m]][c];
                    nodes.push_back(x);
                }
        for (int c = 0; c < 3; ++c) nodes.push_back(-0.5);   // interior: centroid
    }
    return true;
}

// ---------------------------------------------------------------------------
// Lagrange basis: the nodal basis of P_k, phi_i(x_j) = delta_ij.  Feel++ builds
// it by inverting the Vandermonde matrix of the Dubiner basis at the nodes
// (feelpoly/fe.hpp:132-161).  The oracle inverts the monomial Vandermonde
// matrix in extended precision instead; both span P_k, so the nodal basis is
// the same function (round-off differences are ~1e-16).
// ---------------------------------------------------------------------------
struct Basis {
    int dim, k, n;
    std::vector<std::array<int, 3>> mono;   // exponents
    std::vector<long double> coef;          // n x n, column i = coefficients of phi_i
};

Basis make_basis(int dim, int k, const std::vector<double>& nodes) {
    Basis B;
    B.dim = dim; B.k = k;
    for (int t = 0; t <= k; ++t)
        for (int a = t; a >= 0; --a)
            for (int b = t - a; b >= 0; --b) {
                int c = t - a - b;
                if (dim == 2 && c != 0) continue;
                B.mono.push_back({a, b, c});
            }
    int n = (int)B.mono.size();
    B.n = n;
    // V[j][m] = mono_m(x_j); solve V C = I
    std::vector<long double> A(n * n), X(n * n, 0.0L);
    for (int j = 0; j < n; ++j)
        for (int m = 0; m < n; ++m) {
            long double v = 1.0L;
            for (int c = 0; c < dim; ++c) v *= std::pow((long double)nodes[j * dim + c], B.mono[m][c]);
            A[j * n + m] = v;
        }
    for (int i = 0; i < n; ++i) X[i * n + i] = 1.0L;
    // Gaussian elimination with partial pivoting: A * C = I  (C = X after solve)
    std::vector<int> perm(n);
    for (int col = 0; col < n; ++col) {
        int p = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[r * n + col]) > std::fabs(A[p * n + col])) p = r;
        if (p != col) {
            for (int c = 0; c < n; ++c) std::swap(A[p * n + c], A[col * n + c]);
            for (int c = 0; c < n; ++c) std::swap(X[p * n + c], X[col * n + c]);
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            long double f = A[r * n + col] / A[col * n + col];
            if (f == 0) continue;
            for (int c = 0; c < n; ++c) A[r * n + c] -= f * A[col * n + c];
            for (int c = 0; c < n; ++c) X[r * n + c] -= f * X[col * n + c];
        }
    }
    B.coef.resize(n * n);
    for (int m = 0; m < n; ++m)
        for (int i = 0; i < n; ++i) B.coef[m * n + i] = X[m * n + i] / A[m * n + m];
    return B;
}

// phi[i], dphi[i*dim+c] at point x
void eval_basis(const Basis& B, const double* x, double* phi, double* dphi) {
    int n = B.n, dim = B.dim;
    std::vector<long double> mv(n), md(n * dim);
    for (int m = 0; m < n; ++m) {
        long double v = 1.0L;
        for (int c = 0; c < dim; ++c) v *= std::pow((long double)x[c], B.mono[m][c]);
        mv[m] = v;
        for (int d = 0; d < dim; ++d) {
            int e = B.mono[m][d];
            if (e == 0) { md[m * dim + d] = 0; continue; }
            long double w = e;
            for (int c = 0; c < dim; ++c)
                w *= std::pow((long double)x[c], c == d ? e - 1 : B.mono[m][c]);
            md[m * dim + d] = w;
        }
    }
    for (int i = 0; i < n; ++i) {
        long double s = 0;
        for (int m = 0; m < n; ++m) s += B.coef[m * n + i] * mv[m];
        phi[i] = (double)s;
        for (int d = 0; d < dim; ++d) {
            long double g = 0;
            for (int m = 0; m < n; ++m) g += B.coef[m * n + i] * md[m * dim + d];
            dphi[i * dim + d] = (double)g;
        }
    }
}

// ---------------------------------------------------------------------------
// Quadrature: "im(d,order,simplex)" from IMFactory (pointsetquadrature.hpp:85-159)
// ---------------------------------------------------------------------------
bool quadrature(int dim, int order, std::vector<double>& pts, std::vector<double>& w) {
    for (const auto& t : im_simplex_tables)
        if (t.dim == dim && t.order == order) {
            pts.assign(t.npts * dim, 0.0);
            w.assign(t.npts, 0.0);
            for (int i = 0; i < t.npts; ++i) {
                for (int j = 0; j < dim; ++j) pts[i * dim + j] = t.q[(dim + 1) * i + j];
                w[i] = t.q[(dim + 1) * i + dim];
            }
            return true;
        }
    return false;
}

// ---------------------------------------------------------------------------
// Affine geometric map (feelpoly/geomap.hpp:2258-2355): K = G * grad(phi_geo),
// J = |det K| (:1972-1976), B = K^{-T} (:1986-1991).  With P1 geometry on the
// [-1,1] simplex, column a of K is (x_{a+1} - x_0) / 2.
// ---------------------------------------------------------------------------
struct Geo {
    double J;
    double B[3][3];   // B = K^{-T}
};

Geo geomap(int dim, const double* X /* (dim+1) x dim vertex coords */) {
    double K[3][3] = {{0}};
    for (int a = 0; a < dim; ++a)
        for (int r = 0; r < dim; ++r) K[r][a] = (X[(a + 1) * dim + r] - X[r]) / 2.0;
    Geo g;
    if (dim == 2) {
        double det = K[0][0] * K[1][1] - K[0][1] * K[1][0];
        g.J = std::fabs(det);
        // K^{-1} = [k11 -k01; -k10 k00]/det ; B = K^{-T}
        double inv[2][2] = {{K[1][1] / det, -K[0][1] / det}, {-K[1][0] / det, K[0][0] / det}};
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) g.B[r][c] = inv[c][r];
    } else {
        double c00 = K[1][1] * K[2][2] - K[1][2] * K[2][1];
        double c01 = K[1][2] * K[2][0] - K[1][0] * K[2][2];
        double c02 = K[1][0] * K[2][1] - K[1][1] * K[2][0];
        double det = K[0][0] * c00 + K[0][1] * c01 + K[0][2] * c02;
        g.J = std::fabs(det);
        double inv[3][3];
        inv[0][0] = c00 / det;
        inv[1][0] = c01 / det;
        inv[2][0] = c02 / det;
        inv[0][1] = (K[0][2] * K[2][1] - K[0][1] * K[2][2]) / det;
        inv[1][1] = (K[0][0] * K[2][2] - K[0][2] * K[2][0]) / det;
        inv[2][1] = (K[0][1] * K[2][0] - K[0][0] * K[2][1]) / det;
        inv[0][2] = (K[0][1] * K[1][2] - K[0][2] * K[1][1]) / det;
        inv[1][2] = (K[0][2] * K[1][0] - K[0][0] * K[1][2]) / det;
        inv[2][2] = (K[0][0] * K[1][1] - K[0][1] * K[1][0]) / det;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) g.B[r][c] = inv[c][r];
    }
    return g;
}

// Tabulated reference basis at quadrature points (bilinearformcontext.hpp:56-59 PreCompute)
struct Tab {
    int n, Q, dim;
    std::vector<double> w, pts, phi, dphi;   // phi[i*Q+q], dphi[(i*Q+q)*dim+d]
};

bool tabulate(int dim, int k, int qorder, Tab& T) {
    std::vector<double> nodes;
    if (!lagrange_nodes(dim, k, nodes)) return false;
    Basis B = make_basis(dim, k, nodes);
    if (!quadrature(dim, qorder, T.pts, T.w)) return false;
    T.dim = dim; T.n = B.n; T.Q = (int)T.w.size();
    T.phi.assign(T.n * T.Q, 0); T.dphi.assign(T.n * T.Q * dim, 0);
    std::vector<double> p(T.n), dp(T.n * dim);
    for (int q = 0; q < T.Q; ++q) {
        eval_basis(B, &T.pts[q * dim], p.data(), dp.data());
        for (int i = 0; i < T.n; ++i) {
            T.phi[i * T.Q + q] = p[i];
            for (int d = 0; d < dim; ++d) T.dphi[(i * T.Q + q) * dim + d] = dp[i * dim + d];
        }
    }
    return true;
}

void elem_coords(int dim, const double* coords, const int* el, double* X) {
    for (int v = 0; v <= dim; ++v)
        for (int c = 0; c < dim; ++c) X[v * dim + c] = coords[(long)el[v] * dim + c];
}

// MatSetValues(ADD_VALUES) on a SeqAIJ matrix: binary search of the column in
// the row (matrixpetsc.cpp:919-937).  Returns false if (r,c) is not in the
// preallocated pattern (Feel++ keeps MAT_NEW_NONZERO_ALLOCATION_ERR semantics).
bool csr_add(const int* ia, const int* ja, double* val, int r, int c, double v) {
    const int* b = ja + ia[r];
    const int* e = ja + ia[r + 1];
    const int* p = std::lower_bound(b, e, c);
    if (p == e || *p != c) return false;
    val[p - ja] += v;
    return true;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Seeded synthetic inputs
// ---------------------------------------------------------------------------

uint64_t orc_splitmix64(uint64_t* state) {
    SplitMix64 g{*state};
    uint64_t v = g.next();
    *state = g.s;
    return v;
}

// Unit square, n x n cells, each split by the (0,0)-(1,1) diagonal into
// (v00,v10,v11) and (v00,v11,v01); vertex ids x-fastest, element id = 2*cell+t.
// Interior vertices move by h*U(-a,a) per coordinate, drawn in vertex-id order.
int orc_mesh_square(int n, double jitter, uint64_t seed, double* coords, int* elems) {
    SplitMix64 g{seed};
    double h = 1.0 / n;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            long v = i + (long)(n + 1) * j;
            double x = i * h, y = j * h;
            if (i > 0 && i < n && j > 0 && j < n) {
                x += h * jitter * (2.0 * g.uniform() - 1.0);
                y += h * jitter * (2.0 * g.uniform() - 1.0);
            }
            coords[2 * v] = x;
            coords[2 * v + 1] = y;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            long cell = i + (long)n * j;
            int v00 = i + (n + 1) * j, v10 = v00 + 1, v01 = v00 + (n + 1), v11 = v01 + 1;
            int* t0 = elems + 6 * cell;
            t0[0] = v00; t0[1] = v10; t0[2] = v11;
            t0[3] = v00; t0[4] = v11; t0[5] = v01;
        }
    return 0;
}

// Box [0,lx]x[0,ly]x[0,lz], nx*ny*nz cubes, Kuhn 6-tetrahedron subdivision:
// cube (i,j,k) -> tets (v, v+e_p0, v+e_p0+e_p1, v+e_p0+e_p1+e_p2) for the six
// axis permutations p in lexicographic order; element id = 6*cube + perm.
int orc_mesh_kuhn(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                  uint64_t seed, double* coords, int* elems) {
    SplitMix64 g{seed};
    double hx = lx / nx, hy = ly / ny, hz = lz / nz;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                long v = i + (long)(nx + 1) * (j + (long)(ny + 1) * k);
                double x = i * hx, y = j * hy, z = k * hz;
                if (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
                    x += hx * jitter * (2.0 * g.uniform() - 1.0);
                    y += hy * jitter * (2.0 * g.uniform() - 1.0);
                    z += hz * jitter * (2.0 * g.uniform() - 1.0);
                }
                coords[3 * v] = x; coords[3 * v + 1] = y; coords[3 * v + 2] = z;
            }
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    long stride[3] = {1, nx + 1, (long)(nx + 1) * (ny + 1)};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long cube = i + (long)nx * (j + (long)ny * k);
                long v0 = i + stride[1] * j + stride[2] * k;
                for (int p = 0; p < 6; ++p) {
                    int* t = elems + 4 * (6 * cube + p);
                    long v = v0;
                    t[0] = (int)v;
                    for (int s = 0; s < 3; ++s) { v += stride[perms[p][s]]; t[s + 1] = (int)v; }
                }
            }
    return 0;
}

// exp(mu + sigma*Z), Z standard normal by Box-Muller on consecutive uniforms
// (both outputs of each pair used, in order).
int orc_lognormal(long n, double mu, double sigma, uint64_t seed, double* out) {
    SplitMix64 g{seed};
    for (long i = 0; i < n; i += 2) {
        double u1 = 1.0 - g.uniform();   // (0,1]
        double u2 = g.uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        double z0 = r * std::cos(2.0 * M_PI * u2), z1 = r * std::sin(2.0 * M_PI * u2);
        out[i] = std::exp(mu + sigma * z0);
        if (i + 1 < n) out[i + 1] = std::exp(mu + sigma * z1);
    }
    return 0;
}

// lo + (hi-lo)*U, or log-uniform when logscale != 0
int orc_uniform(long n, double lo, double hi, int logscale, uint64_t seed, double* out) {
    SplitMix64 g{seed};
    for (long i = 0; i < n; ++i) {
        double u = g.uniform();
        out[i] = logscale ? std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * u) : lo + (hi - lo) * u;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Reference element data
// ---------------------------------------------------------------------------
int orc_quadrature(int dim, int order, double* pts, double* w) {
    std::vector<double> p, ww;
    if (!quadrature(dim, order, p, ww)) return -1;
    if (pts) std::copy(p.begin(), p.end(), pts);
    if (w) std::copy(ww.begin(), ww.end(), w);
    return (int)ww.size();
}

void orc_set_p4_face_weight(double w) { g_p4_face_weight = w; }

int orc_lagrange_nodes(int dim, int order, double* nodes) {
    std::vector<double> nd;
    if (!lagrange_nodes(dim, order, nd)) return -1;
    if (nodes) std::copy(nd.begin(), nd.end(), nodes);
    return (int)nd.size() / dim;
}

int orc_lagrange_eval(int dim, int order, int npts, const double* pts, double* phi, double* dphi) {
    std::vector<double> nd;
    if (!lagrange_nodes(dim, order, nd)) return -1;
    Basis B = make_basis(dim, order, nd);
    std::vector<double> p(B.n), dp(B.n * dim);
    for (int q = 0; q < npts; ++q) {
        eval_basis(B, pts + q * dim, p.data(), dp.data());
        for (int i = 0; i < B.n; ++i) {
            phi[i * npts + q] = p[i];
            for (int d = 0; d < dim; ++d) dphi[(i * npts + q) * dim + d] = dp[i * dim + d];
        }
    }
    return B.n;
}

// ---------------------------------------------------------------------------
// DofTable::buildDofMap (feeldiscr/doftable.hpp:2364-2455): loop over the
// elements in ascending id (elements.hpp:630-645); DofFromElement::add
// (doffromelement.hpp:492-530) inserts vertex dofs (key = point id, :163-194),
// edge dofs (key = edge id, reversed when the element's local edge runs from
// the larger to the smaller point id, :279-330 with mesh3d.hpp:755-782), face
// dofs (:337-463) and volume dofs (:464) through insertDof
// (doftable.hpp:1213-1300): map.try_emplace(key, next_free_dof); a new key
// advances next_free_dof by nComponents.  Local index = nLocalDof*c + l,
// global = base + c.
// Edge ids are only map keys here; their numerical values do not influence the
// numbering (first touch decides), so the oracle keys edges by their sorted
// vertex pair.  Faces with one dof (P3) need no permutation.
// ---------------------------------------------------------------------------
long orc_doftable(int dim, int order, int ncomp, int ne, const int* elems, int* elem2dof) {
    if (order < 1 || order > 4 || (dim == 2 && order > 2)) return -1;
    const int nv = nvert(dim), nle = nedges(dim), nlf = nfaces3(dim);
    const int nde = dof_per_edge(order), ndf = dim == 3 ? dof_per_face(order) : 0;
    const int nl = nloc_scalar(dim, order);
    std::map<std::tuple<int, long, long, long>, long> gdof;   // (entity type, key...) -> base
    long next_free = 0;
    for (long e = 0; e < ne; ++e) {
        const int* el = elems + (long)nv * e;
        int* l2g = elem2dof + (long)e * nl * ncomp;
        int lc = 0;
        auto insert = [&](std::tuple<int, long, long, long> key) {
            auto [it, inserted] = gdof.try_emplace(key, next_free);
            if (inserted) next_free += ncomp;
            for (int c = 0; c < ncomp; ++c) l2g[nl * c + lc] = (int)(it->second + c);
            ++lc;
        };
        for (int i = 0; i < nv; ++i) insert({0, el[i], 0, 0});
        for (int i = 0; i < nle; ++i) {
            int a = el[dim == 2 ? TRI_E2P[i][0] : TET_E2P[i][0]];
            int b = el[dim == 2 ? TRI_E2P[i][1] : TET_E2P[i][1]];
            bool reverse = a > b;
            long lo = std::min(a, b), hi = std::max(a, b);
            for (int l = 0; l < nde; ++l) {
                long sub = reverse ? nde - 1 - l : l;
                insert({1, lo, hi, sub});
            }
        }
        // Face dofs: with one dof per face (P3) the key is the face.  With three (P4)
        // the element that creates the face numbers them in its local order and the
        // neighbour maps its local dofs through vector_permutation
        // (doffromelement.hpp:372-437, doftable.hpp:1506-1516); a conforming
        // permutation sends each dof to the one at the same physical point, i.e. near
        // the same mesh vertex, so the key here is (face, nearest vertex).
        for (int f = 0; f < nlf; ++f) {
            int v[3] = {el[TET_F2P[f][0]], el[TET_F2P[f][1]], el[TET_F2P[f][2]]};
            int s3[3] = {v[0], v[1], v[2]};
            std::sort(s3, s3 + 3);
            for (int l = 0; l < ndf; ++l) {
                const long near = ndf == 1 ? 0 : v[l];
                insert({2, s3[0], (long)s3[1] * (1L << 31) + s3[2], near});
            }
        }
        if (dim == 2 && order >= 3) return -1;
        for (int l = 0; l < (dim == 3 ? dof_per_vol(order) : 0); ++l) insert({3, e, l, 0});
    }
    return next_free;
}

// ---------------------------------------------------------------------------
// Stencil::computeGraph (feeldiscr/stencil.hpp:1502-1692): for every element
// and every local test dof, insert the element's trial dofs into the row's
// std::set; COUPLED pattern for vector spaces (all components).  GraphCSR::close
// (feelalg/graphcsr.cpp:663-1021) then emits rows ascending, columns ascending;
// addMissingZeroEntriesDiagonal (:634) adds the diagonal.
// If ja == nullptr only ia is filled; returns nnz.
// ---------------------------------------------------------------------------
long orc_graph(int nldof, int ne, long ndof, const int* elem2dof, int* ia, int* ja) {
    std::vector<std::set<int>> rows(ndof);
    for (long e = 0; e < ne; ++e) {
        const int* l2g = elem2dof + (long)e * nldof;
        for (int i = 0; i < nldof; ++i)
            for (int j = 0; j < nldof; ++j) rows[l2g[i]].insert(l2g[j]);
    }
    for (long r = 0; r < ndof; ++r) rows[r].insert((int)r);
    long nnz = 0;
    ia[0] = 0;
    for (long r = 0; r < ndof; ++r) {
        if (ja)
            for (int c : rows[r]) ja[nnz++] = c;
        else
            nnz += rows[r].size();
        ia[r + 1] = (int)nnz;
    }
    return nnz;
}

// ---------------------------------------------------------------------------
// Bilinear form assembly, scalar Lagrange space:
//   a = integrate(elements(mesh), diff*idv(k)*gradt(u)*trans(grad(v))
//                                 + mass*idv(m)*idt(u)*id(v))
// Integrator::assemble element loop (feelvf/integrator.hpp:1093-1207):
// gmc->update (geomap), M_prod[q] = w_q*J (pointsetquadrature.hpp:481-495),
// M_rep(i,j) = sum_q M_prod[q]*expr(i,j,q) (bilinearformcontext.hpp:455-628
// with pointsetquadrature.hpp:403-416), grad phi = B * grad_ref phi
// (polynomialset_impl.hpp:274-302), then MatSetValues ADD into the CSR
// (bilinearformcontext.hpp:772-889, matrixpetsc.cpp:919-937).
// kdiff / kmass are per-element P0 coefficients (nullptr -> 1).
// Returns 0, or -1 on unsupported input, -2 if an entry falls outside the pattern.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
// Element matrix of kd*grad(u).grad(v) + km*u*v on one simplex with vertex
// coordinates X; Ae row-major (test index i, bilinearform.hpp:462-492).
void scalar_element(const Tab& T, int dim, const double* X, double kd, double km, std::vector<double>& gphi,
                    std::vector<double>& prod, std::vector<double>& Ae) {
    const int n = T.n, Q = T.Q;
    Geo g = geomap(dim, X);
    for (int q = 0; q < Q; ++q) prod[q] = T.w[q] * g.J;
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < Q; ++q)
            for (int r = 0; r < dim; ++r) {
                double s = 0;
                for (int c = 0; c < dim; ++c) s += g.B[r][c] * T.dphi[(i * Q + q) * dim + c];
                gphi[(i * Q + q) * dim + r] = s;
            }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int q = 0; q < Q; ++q) {
                double ex = 0;
                if (kd != 0) {
                    double gg = 0;
                    for (int r = 0; r < dim; ++r) gg += gphi[(j * Q + q) * dim + r] * gphi[(i * Q + q) * dim + r];
                    ex += kd * gg;
                }
                if (km != 0) ex += km * T.phi[j * Q + q] * T.phi[i * Q + q];
                s += prod[q] * ex;
            }
            Ae[i * n + j] = s;
        }
}
}  // namespace

extern "C" {

int orc_assemble_scalar(int dim, int order, int qorder, int ne, const double* coords, const int* elems,
                        const int* elem2dof, double diff, const double* kdiff, double mass,
                        const double* kmass, const int* ia, const int* ja, double* val) {
    Tab T;
    if (!tabulate(dim, order, qorder, T)) return -1;
    const int n = T.n, Q = T.Q;
    std::vector<double> gphi(n * Q * dim), prod(Q), Ae(n * n);
    double X[12];
    for (long e = 0; e < ne; ++e) {
        elem_coords(dim, coords, elems + (long)(dim + 1) * e, X);
        scalar_element(T, dim, X, diff * (kdiff ? kdiff[e] : 1.0), mass * (kmass ? kmass[e] : 1.0), gphi, prod, Ae);
        const int* l2g = elem2dof + (long)e * n;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (!csr_add(ia, ja, val, l2g[i], l2g[j], Ae[i * n + j])) return -2;
    }
    return 0;
}

// Threaded stand-in for Feel++'s MPI assembly (a CPU baseline, not a parity
// path): nthreads workers each own a contiguous block of rows, as an MPI rank
// owns its dofs, and assemble every element that touches those rows (the
// ghost layer is computed twice), adding only into their own rows.  Same
// element matrices and the same per-row addition order as orc_assemble_scalar.
int orc_assemble_scalar_mt(int dim, int order, int qorder, int ne, const double* coords, const int* elems,
                           const int* elem2dof, double diff, const double* kdiff, double mass,
                           const double* kmass, long nrows, const int* ia, const int* ja, double* val,
                           int nthreads) {
    Tab T;
    if (!tabulate(dim, order, qorder, T)) return -1;
    const int n = T.n, Q = T.Q;
    if (nthreads < 1) nthreads = 1;
    int bad = 0;
#pragma omp parallel num_threads(nthreads) reduction(| : bad)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        // rows split by equal nonzero counts
        const long nnz = ia[nrows];
        long r0 = 0, r1 = 0;
        {
            const long lo = nnz * t / nt, hi = nnz * (t + 1) / nt;
            r0 = std::lower_bound(ia, ia + nrows, (int)lo) - ia;
            r1 = t == nt - 1 ? nrows : std::lower_bound(ia, ia + nrows, (int)hi) - ia;
        }
        std::vector<double> gphi(n * Q * dim), prod(Q), Ae(n * n);
        double X[12];
        for (long e = 0; e < ne && !bad; ++e) {
            const int* l2g = elem2dof + (long)e * n;
            bool mine = false;
            for (int i = 0; i < n; ++i) mine |= l2g[i] >= r0 && l2g[i] < r1;
            if (!mine) continue;
            elem_coords(dim, coords, elems + (long)(dim + 1) * e, X);
            scalar_element(T, dim, X, diff * (kdiff ? kdiff[e] : 1.0), mass * (kmass ? kmass[e] : 1.0), gphi, prod,
                           Ae);
            for (int i = 0; i < n; ++i) {
                if (l2g[i] < r0 || l2g[i] >= r1) continue;
                for (int j = 0; j < n; ++j)
                    if (!csr_add(ia, ja, val, l2g[i], l2g[j], Ae[i * n + j])) bad = 1;
            }
        }
    }
    return bad ? -2 : 0;
}

// Isotropic linear elasticity a = integrate(elements, inner(sigma(u), grad(v)))
// with sigma = lambda*tr(eps)I + 2*mu*eps (quickstart/qs_elasticity.cpp:41-63),
// lambda = E nu/((1+nu)(1-2nu)), mu = E/(2(1+nu)); vector space Pchv (3
// components, local index nloc*c + l).  Entry (test (c,i), trial (d,j)):
//   lambda dj_d di_c + mu (delta_cd grad(j).grad(i) + dj_c di_d).
int orc_assemble_elasticity(int order, int qorder, int ne, const double* coords, const int* elems,
                            const int* elem2dof, const double* E, const double* nu, const int* ia,
                            const int* ja, double* val) {
    const int dim = 3;
    Tab T;
    if (!tabulate(dim, order, qorder, T)) return -1;
    const int n = T.n, Q = T.Q, N = 3 * n;
    std::vector<double> gphi(n * Q * dim), prod(Q), Ae(N * N);
    double X[12];
    for (long e = 0; e < ne; ++e) {
        elem_coords(dim, coords, elems + 4 * e, X);
        Geo g = geomap(dim, X);
        for (int q = 0; q < Q; ++q) prod[q] = T.w[q] * g.J;
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < Q; ++q)
                for (int r = 0; r < dim; ++r) {
                    double s = 0;
                    for (int c = 0; c < dim; ++c) s += g.B[r][c] * T.dphi[(i * Q + q) * dim + c];
                    gphi[(i * Q + q) * dim + r] = s;
                }
        double lam = E[e] * nu[e] / ((1 + nu[e]) * (1 - 2 * nu[e]));
        double mu = E[e] / (2 * (1 + nu[e]));
        for (int d = 0; d < 3; ++d)
            for (int j = 0; j < n; ++j)
                for (int c = 0; c < 3; ++c)
                    for (int i = 0; i < n; ++i) {
                        double s = 0;
                        for (int q = 0; q < Q; ++q) {
                            const double* gj = &gphi[(j * Q + q) * dim];
                            const double* gi = &gphi[(i * Q + q) * dim];
                            double ex = lam * gj[d] * gi[c] + mu * gj[c] * gi[d];
                            if (c == d) ex += mu * (gj[0] * gi[0] + gj[1] * gi[1] + gj[2] * gi[2]);
                            s += prod[q] * ex;
                        }
                        Ae[(c * n + i) * N + (d * n + j)] = s;
                    }
        const int* l2g = elem2dof + (long)e * N;
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b)
                if (!csr_add(ia, ja, val, l2g[a], l2g[b], Ae[a * N + b])) return -2;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Linear form l = integrate(elements, f*id(v)) (linearformcontext.hpp:294-305,
// VecSetValues ADD: vectorpetsc.cpp:187).  fkind 0: f = fp[0] (constant);
// fkind 1: f = fp[0]*sin(pi x) sin(pi y); fkind 2: vector f = (fp[0],fp[1],fp[2])
// on a 3-component space.  x_q = sum_v x_v phi_geo,v(q).
// ---------------------------------------------------------------------------
int orc_assemble_rhs(int dim, int order, int ncomp, int qorder, int ne, const double* coords,
                     const int* elems, const int* elem2dof, int fkind, const double* fp, double* b) {
    Tab T;
    if (!tabulate(dim, order, qorder, T)) return -1;
    Tab G;   // P1 geometry basis at the quadrature points
    if (!tabulate(dim, 1, qorder, G)) return -1;
    const int n = T.n, Q = T.Q;
    double X[12];
    std::vector<double> be(n * ncomp);
    for (long e = 0; e < ne; ++e) {
        elem_coords(dim, coords, elems + (long)(dim + 1) * e, X);
        Geo g = geomap(dim, X);
        std::fill(be.begin(), be.end(), 0.0);
        for (int q = 0; q < Q; ++q) {
            double xq[3] = {0, 0, 0};
            for (int v = 0; v <= dim; ++v)
                for (int c = 0; c < dim; ++c) xq[c] += X[v * dim + c] * G.phi[v * Q + q];
            double wJ = T.w[q] * g.J;
            for (int c = 0; c < ncomp; ++c) {
                double f;
                if (fkind == 0) f = fp[0];
                else if (fkind == 1) f = fp[0] * std::sin(M_PI * xq[0]) * std::sin(M_PI * xq[1]);
                else f = fp[c];
                for (int i = 0; i < n; ++i) be[c * n + i] += wJ * f * T.phi[i * Q + q];
            }
        }
        const int* l2g = elem2dof + (long)e * n * ncomp;
        for (int a = 0; a < n * ncomp; ++a) b[l2g[a]] += be[a];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Boundary dofs for on(_range=boundaryfaces(mesh)) / markedfaces: the local
// dofs of each boundary face (IntegratorOnExpr::onFaces, feelvf/on.hpp:457-701,
// faceLocalDof :648): the face's vertices, edges and face-interior dofs.
// A face is on the boundary when exactly one element holds it.  sel: 0 = all
// boundary faces, 1 = boundary faces with every vertex at x == 0 (clamped end).
// Returns the number of marked dofs (all components).
// ---------------------------------------------------------------------------
long orc_boundary_dofs(int dim, int order, int ncomp, int ne, const double* coords, const int* elems,
                       const int* elem2dof, long ndof, int sel, unsigned char* isbnd) {
    const int nv = dim + 1, nl = nloc_scalar(dim, order);
    const int nde = dof_per_edge(order), ndf = dim == 3 ? dof_per_face(order) : 0;
    std::map<std::vector<int>, std::pair<long, int>> faces;   // sorted vertices -> (elem, local face), count
    std::map<std::vector<int>, int> count;
    for (long e = 0; e < ne; ++e) {
        const int* el = elems + (long)nv * e;
        for (int f = 0; f < nv; ++f) {
            std::vector<int> key;
            for (int v = 0; v < nv; ++v)
                if (v != f) key.push_back(el[v]);   // local face f is opposite vertex f
            std::sort(key.begin(), key.end());
            if (++count[key] == 1) faces[key] = {e, f};
        }
    }
    std::fill(isbnd, isbnd + ndof, 0);
    long nmark = 0;
    for (auto& [key, c] : count) {
        if (c != 1) continue;
        if (sel == 1) {
            bool ok = true;
            for (int v : key) ok = ok && coords[(long)v * dim] == 0.0;
            if (!ok) continue;
        }
        auto [e, f] = faces[key];
        const int* l2g = elem2dof + (long)e * nl * ncomp;
        std::vector<int> loc;
        for (int v = 0; v < nv; ++v)
            if (v != f) loc.push_back(v);
        int base = nv;
        for (int i = 0; i < nedges(dim); ++i) {
            int a = dim == 2 ? TRI_E2P[i][0] : TET_E2P[i][0];
            int b = dim == 2 ? TRI_E2P[i][1] : TET_E2P[i][1];
            if (a != f && b != f)
                for (int l = 0; l < nde; ++l) loc.push_back(base + i * nde + l);
        }
        base += nedges(dim) * nde;
        if (dim == 3)
            for (int l = 0; l < ndf; ++l) loc.push_back(base + f * ndf + l);
        for (int l : loc)
            for (int c = 0; c < ncomp; ++c) {
                int gd = l2g[nl * c + l];
                if (!isbnd[gd]) { isbnd[gd] = 1; ++nmark; }
            }
    }
    return nmark;
}

// ---------------------------------------------------------------------------
// Dirichlet elimination, MatrixPetsc::zeroRows (feelalg/matrixpetsc.cpp:1764-1850)
// with value_on_diagonal (on.hpp:347); the pattern is kept.
//   mode 0 "elimination" (MatZeroRows):            rows zeroed, diag set,
//                                                  b_i = diag*g_i
//   mode 1 "elimination_symmetric" (MatZeroRowsColumns): rows and columns zeroed,
//                                                  b_j -= A_ji g_i for kept rows,
//                                                  b_i = diag*g_i
// ---------------------------------------------------------------------------
int orc_dirichlet(int mode, long n, const int* ia, const int* ja, double* val, double* b,
                  const unsigned char* isbnd, const double* g, double diag) {
    if (mode == 1) {
        for (long r = 0; r < n; ++r) {
            if (isbnd[r]) continue;
            for (int p = ia[r]; p < ia[r + 1]; ++p)
                if (isbnd[ja[p]]) {
                    b[r] -= val[p] * (g ? g[ja[p]] : 0.0);
                    val[p] = 0.0;
                }
        }
    }
    for (long r = 0; r < n; ++r) {
        if (!isbnd[r]) continue;
        for (int p = ia[r]; p < ia[r + 1]; ++p) val[p] = (ja[p] == r) ? diag : 0.0;
        b[r] = diag * (g ? g[r] : 0.0);
    }
    return 0;
}

// normL2(_range=elements(mesh), _expr=idv(u) - f) = sqrt(|sum_e sum_q w_q J (u_h - f)^2|)
// (feelvf/norml2.hpp:37-57) and the H1 seminorm with grad(u_h) - grad(f)
// (feelvf/normsemih1.hpp), per component c of a Pch/Pchv space (local dof
// nloc*c + l).  f per component: kind 0 zero, 1 c0 + c1 x + c2 y + c3 z,
// 2 c0 sin(pi x) sin(pi y).  x_q = sum_v lambda_v(q) X_v on the [-1,1] simplex.
double orc_norm(int dim, int order, int ncomp, int qorder, int ne, const double* coords, const int* elems,
                const int* elem2dof, const double* u, const int* kinds, const double* params, int semi) {
    Tab T;
    if (!tabulate(dim, order, qorder, T)) return -1.0;
    const int n = T.n, Q = T.Q, nldof = n * ncomp;
    double X[12];
    double total = 0.0;
    for (long e = 0; e < ne; ++e) {
        elem_coords(dim, coords, elems + (long)(dim + 1) * e, X);
        Geo g = geomap(dim, X);
        for (int q = 0; q < Q; ++q) {
            double lam[4], x[3] = {0, 0, 0}, sum = 0;
            for (int c = 0; c < dim; ++c) { lam[c + 1] = 0.5 * (1.0 + T.pts[q * dim + c]); sum += lam[c + 1]; }
            lam[0] = 1.0 - sum;
            for (int d = 0; d < dim; ++d)
                for (int v = 0; v <= dim; ++v) x[d] += lam[v] * X[v * dim + d];
            const double wJ = T.w[q] * g.J;
            for (int c = 0; c < ncomp; ++c) {
                const int* l2g = elem2dof + (long)e * nldof + c * n;
                const double* pc = params + 4 * c;
                double fv = 0.0, gf[3] = {0, 0, 0};
                if (kinds[c] == 1) {
                    fv = pc[0];
                    for (int d = 0; d < dim; ++d) { fv += pc[d + 1] * x[d]; gf[d] = pc[d + 1]; }
                } else if (kinds[c] == 2) {
                    fv = pc[0] * std::sin(M_PI * x[0]) * std::sin(M_PI * x[1]);
                    gf[0] = pc[0] * M_PI * std::cos(M_PI * x[0]) * std::sin(M_PI * x[1]);
                    gf[1] = pc[0] * M_PI * std::sin(M_PI * x[0]) * std::cos(M_PI * x[1]);
                }
                if (!semi) {
                    double uh = 0.0;
                    for (int l = 0; l < n; ++l) uh += T.phi[l * Q + q] * u[l2g[l]];
                    total += wJ * (uh - fv) * (uh - fv);
                } else {
                    double gu[3] = {0, 0, 0};
                    for (int l = 0; l < n; ++l)
                        for (int r = 0; r < dim; ++r) {
                            double s2 = 0;
                            for (int a = 0; a < dim; ++a) s2 += g.B[r][a] * T.dphi[(l * Q + q) * dim + a];
                            gu[r] += u[l2g[l]] * s2;
                        }
                    for (int r = 0; r < dim; ++r) total += wJ * (gu[r] - gf[r]) * (gu[r] - gf[r]);
                }
            }
        }
    }
    return std::sqrt(std::fabs(total));
}

// y = A x (MatMult_SeqAIJ)
void orc_spmv(long n, const int* ia, const int* ja, const double* val, const double* x, double* y) {
    for (long r = 0; r < n; ++r) {
        double s = 0;
        for (int p = ia[r]; p < ia[r + 1]; ++p) s += val[p] * x[ja[p]];
        y[r] = s;
    }
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned CG with PETSc KSPCG semantics as Feel++ configures it
// (SolverLinearPetsc::solve, feelalg/solverlinearpetsc.cpp:575-640;
// KSPSetTolerances(rtol, atol, dtol, maxit) :592; KSPConvergedDefaultSetUIRNorm
// :610; KSPCG :872; PCJACOBI).  PETSc (third party, not vendored): default CG
// norm = preconditioned, ||B r_k|| <= max(rtol*||B r_0||, atol) converges,
// ||B r_k|| >= dtol*||B r_0|| diverges; x0 = x on entry (nonzero guess).
// its = number of A*p products; reason > 0 converged (2 rtol, 3 atol),
// -3 max iterations, -4 dtol, -8 indefinite preconditioner ((r,z) changes sign,
// KSP_DIVERGED_INDEFINITE_PC), -10 indefinite matrix ((p,Ap) zero or changes sign,
// KSP_DIVERGED_INDEFINITE_MAT), -9 NaN (KSP_DIVERGED_NANORINF).
// ---------------------------------------------------------------------------
int orc_pcg(long n, const int* ia, const int* ja, const double* val, const double* b, double* x,
            double rtol, double atol, double dtol, int maxit, int* its_out, double* resid_out,
            int* reason_out) {
    std::vector<double> r(n), z(n), p(n), w(n), dinv(n);
    for (long i = 0; i < n; ++i) {
        double d = 0;
        for (int q = ia[i]; q < ia[i + 1]; ++q)
            if (ja[q] == i) d = val[q];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    orc_spmv(n, ia, ja, val, x, w.data());
    for (long i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double dp = 0;
    for (long i = 0; i < n; ++i) { z[i] = dinv[i] * r[i]; dp += z[i] * z[i]; }
    dp = std::sqrt(dp);
    double ttol = std::max(rtol * dp, atol), rnorm0 = dp;
    int reason = 0, its = 0;
    if (dp <= ttol) reason = dp < atol ? 3 : 2;
    double beta = 0, betaold = 0, dpi = 0, dpiold = 0;
    for (long i = 0; i < n; ++i) beta += z[i] * r[i];
    for (int i = 0; !reason && i < maxit; ++i) {
        its = i + 1;
        if (beta == 0.0) { reason = 3; break; }
        if (i > 0 && beta * betaold < 0.0) { reason = -8; break; }
        if (i == 0) {
            p = z;
        } else {
            double bb = beta / betaold;
            for (long k = 0; k < n; ++k) p[k] = z[k] + bb * p[k];
        }
        dpiold = dpi;
        orc_spmv(n, ia, ja, val, p.data(), w.data());
        dpi = 0;
        for (long k = 0; k < n; ++k) dpi += p[k] * w[k];
        if (dpi == 0.0 || (i > 0 && dpi * dpiold <= 0.0)) { reason = -10; break; }
        double a = beta / dpi;
        for (long k = 0; k < n; ++k) { x[k] += a * p[k]; r[k] -= a * w[k]; }
        dp = 0;
        for (long k = 0; k < n; ++k) { z[k] = dinv[k] * r[k]; dp += z[k] * z[k]; }
        dp = std::sqrt(dp);
        if (dp != dp) { reason = -9; break; }
        if (dp <= ttol) { reason = dp < atol ? 3 : 2; break; }
        if (dp >= dtol * rnorm0) { reason = -4; break; }
        betaold = beta;
        beta = 0;
        for (long k = 0; k < n; ++k) beta += z[k] * r[k];
        if (i + 1 >= maxit) reason = -3;
    }
    if (!reason && its >= maxit) reason = -3;
    if (its_out) *its_out = its;
    if (resid_out) *resid_out = dp;
    if (reason_out) *reason_out = reason;
    return 0;
}

// orc_pcg with its loops threaded over rows (OpenMP, nthreads workers): a CPU
// baseline standing in for PETSc's MPI KSPCG, not a parity path (the sums are
// reassociated, so iterates differ from orc_pcg in the last bits).
int orc_pcg_mt(long n, const int* ia, const int* ja, const double* val, const double* b, double* x,
               double rtol, double atol, double dtol, int maxit, int nthreads, int* its_out, double* resid_out,
               int* reason_out) {
    std::vector<double> r(n), z(n), p(n), w(n), dinv(n);
    if (nthreads < 1) nthreads = 1;
    double dp = 0, beta = 0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : dp, beta)
    for (long i = 0; i < n; ++i) {
        double d = 0, s = 0;
        for (int q = ia[i]; q < ia[i + 1]; ++q) {
            if (ja[q] == i) d = val[q];
            s += val[q] * x[ja[q]];
        }
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
        r[i] = b[i] - s;
        z[i] = dinv[i] * r[i];
        dp += z[i] * z[i];
        beta += z[i] * r[i];
    }
    dp = std::sqrt(dp);
    const double ttol = std::max(rtol * dp, atol), rnorm0 = dp;
    int reason = 0, its = 0;
    if (dp <= ttol) reason = dp < atol ? 3 : 2;
    double betaold = 0, dpi = 0, dpiold = 0;
    for (int it = 0; !reason && it < maxit; ++it) {
        its = it + 1;
        if (beta == 0.0) { reason = 3; break; }
        if (it > 0 && beta * betaold < 0.0) { reason = -8; break; }
        const double bb = it == 0 ? 0.0 : beta / betaold;
        double pw = 0;
#pragma omp parallel num_threads(nthreads)
        {
#pragma omp for schedule(static)
            for (long k = 0; k < n; ++k) p[k] = it == 0 ? z[k] : z[k] + bb * p[k];
#pragma omp for schedule(static) reduction(+ : pw)
            for (long k = 0; k < n; ++k) {
                double s = 0;
                for (int q = ia[k]; q < ia[k + 1]; ++q) s += val[q] * p[ja[q]];
                w[k] = s;
                pw += p[k] * s;
            }
        }
        dpiold = dpi;
        dpi = pw;
        if (dpi == 0.0 || (it > 0 && dpi * dpiold <= 0.0)) { reason = -10; break; }
        const double a = beta / dpi;
        double zz = 0, zr = 0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : zz, zr)
        for (long k = 0; k < n; ++k) {
            x[k] += a * p[k];
            r[k] -= a * w[k];
            z[k] = dinv[k] * r[k];
            zz += z[k] * z[k];
            zr += z[k] * r[k];
        }
        dp = std::sqrt(zz);
        if (dp != dp) { reason = -9; break; }
        if (dp <= ttol) { reason = dp < atol ? 3 : 2; break; }
        if (dp >= dtol * rnorm0) { reason = -4; break; }
        betaold = beta;
        beta = zr;
        if (it + 1 >= maxit) reason = -3;
    }
    if (!reason && its >= maxit) reason = -3;
    if (its_out) *its_out = its;
    if (resid_out) *resid_out = dp;
    if (reason_out) *reason_out = reason;
    return 0;
}

}  // extern "C"
