// Implementation of the Feel++-shaped C++ surface (feel.hpp).
#include "feel.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <random>

namespace feelhip {

namespace {

std::shared_ptr<feel_hip_vec> makeVec(feel_hip_ctx* ctx, int64_t n) {
    feel_hip_vec* v = nullptr;
    check(feel_hip_vec_create(ctx, n, &v));
    return std::shared_ptr<feel_hip_vec>(v, [](feel_hip_vec* p) { feel_hip_vec_destroy(p); });
}

node_ptr mk(Node::Op op, std::vector<node_ptr> kids = {}, double c = 0) {
    auto n = std::make_shared<Node>();
    n->op = op;
    n->kids = std::move(kids);
    n->c = c;
    return n;
}

// ---------------------------------------------------------------- evaluation
struct Val {
    int r = 1, c = 1;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double& operator()(int i, int j) { return a[i * c + j]; }
    double operator()(int i, int j) const { return a[i * c + j]; }
    bool scalar() const { return r == 1 && c == 1; }
};
Val scal(double x) { Val v; v.a[0] = x; return v; }

struct Probe {
    Val u, gu, v, gv;   // trial / test value and gradient
    double x[3] = {0, 0, 0};
    int64_t elem = 0;
};

Val eval(Node const& n, Probe const& P) {
    auto kid = [&](int i) { return eval(*n.kids[i], P); };
    switch (n.op) {
        case Node::CST: return scal(n.c);
        case Node::CSTV: {
            Val o;
            o.r = 3; o.c = 1;
            for (int i = 0; i < 3; ++i) o.a[i] = n.cv[i];
            return o;
        }
        case Node::IDT: return P.u;
        case Node::ID: return P.v;
        case Node::GRADT: return P.gu;
        case Node::GRAD: return P.gv;
        case Node::IDV: {
            auto const& f = *n.f.values;
            if (P.elem < 0 || P.elem >= (int64_t)f.size()) throw std::invalid_argument("P0 field shorter than the mesh");
            return scal(f[P.elem]);
        }
        case Node::PX: return scal(P.x[0]);
        case Node::PY: return scal(P.x[1]);
        case Node::PZ: return scal(P.x[2]);
        case Node::SIN: {
            Val a = kid(0);
            if (!a.scalar()) throw std::invalid_argument("sin of a non-scalar");
            return scal(std::sin(a.a[0]));
        }
        case Node::NEG: {
            Val a = kid(0);
            for (double& x : a.a) x = -x;
            return a;
        }
        case Node::ADD:
        case Node::SUB: {
            Val a = kid(0), b = kid(1);
            if (a.r != b.r || a.c != b.c) throw std::invalid_argument("shape mismatch in +/-");
            for (int i = 0; i < 9; ++i) a.a[i] = n.op == Node::ADD ? a.a[i] + b.a[i] : a.a[i] - b.a[i];
            return a;
        }
        case Node::MUL: {
            Val a = kid(0), b = kid(1);
            if (a.scalar()) { for (double& x : b.a) x *= a.a[0]; return b; }
            if (b.scalar()) { for (double& x : a.a) x *= b.a[0]; return a; }
            if (a.c != b.r) throw std::invalid_argument("shape mismatch in *");
            Val o;
            o.r = a.r; o.c = b.c;
            for (int i = 0; i < o.r; ++i)
                for (int j = 0; j < o.c; ++j) {
                    double s = 0;
                    for (int k = 0; k < a.c; ++k) s += a(i, k) * b(k, j);
                    o(i, j) = s;
                }
            return o;
        }
        case Node::INNER: {
            Val a = kid(0), b = kid(1);
            if (a.r != b.r || a.c != b.c) throw std::invalid_argument("shape mismatch in inner");
            double s = 0;
            for (int i = 0; i < a.r * a.c; ++i) s += a.a[i] * b.a[i];
            return scal(s);
        }
        case Node::TRANS: {
            Val a = kid(0), o;
            o.r = a.c; o.c = a.r;
            for (int i = 0; i < a.r; ++i)
                for (int j = 0; j < a.c; ++j) o(j, i) = a(i, j);
            return o;
        }
        case Node::SYM: {
            Val a = kid(0), o = a;
            if (a.r != a.c) throw std::invalid_argument("sym of a non-square");
            for (int i = 0; i < a.r; ++i)
                for (int j = 0; j < a.c; ++j) o(i, j) = 0.5 * (a(i, j) + a(j, i));
            return o;
        }
        case Node::TRACE: {
            Val a = kid(0);
            if (a.r != a.c) throw std::invalid_argument("trace of a non-square");
            double s = 0;
            for (int i = 0; i < a.r; ++i) s += a(i, i);
            return scal(s);
        }
        case Node::EYE: {
            Val o;
            o.r = o.c = 3;
            o(0, 0) = o(1, 1) = o(2, 2) = 1;
            return o;
        }
    }
    throw std::logic_error("bad node");
}

bool hasOp(Node const& n, Node::Op op) {
    if (n.op == op) return true;
    for (auto& k : n.kids)
        if (hasOp(*k, op)) return true;
    return false;
}

// polynomial order of an expression for quadrature deduction (expr.hpp:1058-1076,
// operators.hpp:362-364): id/idt k, grad/gradt k-1, P0 fields and constants 0, x 1
int polyOrder(Node const& n, int k) {
    switch (n.op) {
        case Node::IDT: case Node::ID: return k;
        case Node::GRADT: case Node::GRAD: return std::max(0, k - 1);
        case Node::PX: case Node::PY: case Node::PZ: return 1;
        case Node::SIN: return 2;   // non-polynomial: Feel++ uses its default integration order; 2 here
        case Node::MUL: case Node::INNER: {
            int s = 0;
            for (auto& c : n.kids) s += polyOrder(*c, k);
            return s;
        }
        default: {
            int m = 0;
            for (auto& c : n.kids) m = std::max(m, polyOrder(*c, k));
            return m;
        }
    }
}

Val zeros(int r, int c) { Val v; v.r = r; v.c = c; return v; }

}  // namespace

// ---------------------------------------------------------------- expression constructors
Expr cst(double c) { return {mk(Node::CST, {}, c)}; }
Expr vec(double a, double b, double c) {
    auto n = std::make_shared<Node>();
    n->op = Node::CSTV;
    n->cv[0] = a; n->cv[1] = b; n->cv[2] = c;
    return {n};
}
Expr idt(Element const& u) { auto n = std::make_shared<Node>(); n->op = Node::IDT; n->trial = u.id(); return {n}; }
Expr id(Element const& v) { auto n = std::make_shared<Node>(); n->op = Node::ID; n->test = v.id(); return {n}; }
Expr gradt(Element const& u) { auto n = std::make_shared<Node>(); n->op = Node::GRADT; n->trial = u.id(); return {n}; }
Expr grad(Element const& v) { auto n = std::make_shared<Node>(); n->op = Node::GRAD; n->test = v.id(); return {n}; }
Expr idv(P0Field const& k) {
    if (!k.values) throw std::invalid_argument("idv of an empty field");
    auto n = std::make_shared<Node>();
    n->op = Node::IDV;
    n->f = k;
    return {n};
}
Expr Px() { return {mk(Node::PX)}; }
Expr Py() { return {mk(Node::PY)}; }
Expr Pz() { return {mk(Node::PZ)}; }
Expr sin(Expr const& a) { return {mk(Node::SIN, {a.n})}; }
Expr inner(Expr const& a, Expr const& b) { return {mk(Node::INNER, {a.n, b.n})}; }
Expr trans(Expr const& a) { return {mk(Node::TRANS, {a.n})}; }
Expr sym(Expr const& a) { return {mk(Node::SYM, {a.n})}; }
Expr trace(Expr const& a) { return {mk(Node::TRACE, {a.n})}; }
Expr eye3() { return {mk(Node::EYE)}; }
Expr operator+(Expr const& a, Expr const& b) { return {mk(Node::ADD, {a.n, b.n})}; }
Expr operator-(Expr const& a, Expr const& b) { return {mk(Node::SUB, {a.n, b.n})}; }
Expr operator-(Expr const& a) { return {mk(Node::NEG, {a.n})}; }
Expr operator*(Expr const& a, Expr const& b) { return {mk(Node::MUL, {a.n, b.n})}; }
Expr operator*(double a, Expr const& b) { return cst(a) * b; }
Expr operator*(Expr const& a, double b) { return a * cst(b); }

// ---------------------------------------------------------------- mesh / spaces / elements
mesh_ptr loadMesh(std::string const& path, int dim) {
    std::ifstream is(path);
    if (!is) throw std::runtime_error("cannot open " + path);
    return std::make_shared<Mesh>(Mesh{feelpp_amd::readMsh22(is, dim)});
}

FunctionSpace::FunctionSpace(mesh_ptr mesh, int order, int ncomp, std::shared_ptr<Backend> be)
    : M_mesh(std::move(mesh)), M_backend(std::move(be)) {
    // DofTable::build on the device; the host copy serves boundary dof lists and export
    auto const& m = M_mesh->m;
    check(feel_hip_space_create_from_mesh(M_backend->ctx(), m.dim, order, ncomp, m.nv, m.coords.data(), m.ne,
                                          m.elems.data(), &M_space));
    M_dt.dim = m.dim;
    M_dt.order = order;
    M_dt.ncomp = ncomp;
    M_dt.nloc = feelpp_amd::nLocalDof(m.dim, order);
    M_dt.nldof = M_dt.nloc * ncomp;
    M_dt.ne = m.ne;
    M_dt.ndof = feel_hip_space_ndof(M_space);
    M_dt.elem2dof.resize((size_t)m.ne * M_dt.nldof);
    check(feel_hip_space_elem2dof(M_space, M_dt.elem2dof.data()));
}

std::shared_ptr<feel_hip_graph> FunctionSpace::graph() {
    // one pattern per space, shared by its forms (StencilManager cache, backend.hpp:423)
    if (!M_graph) {
        feel_hip_graph* g = nullptr;
        check(feel_hip_graph_create_from_space(M_space, &g));
        M_graph = std::shared_ptr<feel_hip_graph>(g, [](feel_hip_graph* p) { feel_hip_graph_destroy(p); });
    }
    return M_graph;
}

Element FunctionSpace::element() {
    return Element(shared_from_this(), makeVec(M_backend->ctx(), M_dt.ndof), M_next_id++);
}

std::vector<double> Element::values() const {
    std::vector<double> out(feel_hip_vec_size(M_vec.get()));
    check(feel_hip_vec_get(M_vec.get(), out.data()));
    return out;
}

void Element::setValues(std::vector<double> const& v) {
    if ((int64_t)v.size() != feel_hip_vec_size(M_vec.get())) throw std::invalid_argument("size mismatch");
    check(feel_hip_vec_set(M_vec.get(), v.data()));
}

// ---------------------------------------------------------------- linear form
LinearForm::LinearForm(space_ptr test) : M_test(std::move(test)) {
    M_vec = makeVec(M_test->backendPtr()->ctx(), M_test->nDof());
}

std::vector<double> LinearForm::values() const {
    std::vector<double> out(feel_hip_vec_size(M_vec.get()));
    check(feel_hip_vec_get(M_vec.get(), out.data()));
    return out;
}

void LinearForm::assemble(IntegralExpr const& ie, bool add) {
    if (ie.range.mesh != M_test->mesh()) throw std::invalid_argument("integration range is not the test space's mesh");
    Node const& root = *ie.expr.n;
    if (hasOp(root, Node::IDT) || hasOp(root, Node::GRADT)) throw std::invalid_argument("linear form contains the trial function");
    if (hasOp(root, Node::GRAD) || hasOp(root, Node::IDV))
        throw std::invalid_argument("linear form outside the supported subset (f*id(v), f constant or sin(pi x)sin(pi y))");
    const int k = M_test->order(), nc = M_test->nComponents();
    feel_hip_linear_desc d{};
    d.quad_order = ie.quad >= 0 ? ie.quad : polyOrder(root, k);
    Probe P;
    if (nc == 1) {
        P.v = scal(1.0);
        auto f = [&](double x, double y, double z) { P.x[0] = x; P.x[1] = y; P.x[2] = z; return eval(root, P); };
        Val c = f(0.5, 0.5, 0.5);
        if (!c.scalar()) throw std::invalid_argument("linear form is not scalar");
        std::mt19937_64 g(7);
        std::uniform_real_distribution<double> U(0.05, 0.95);
        bool constant = true, sinsin = true;
        for (int t = 0; t < 8; ++t) {
            const double x = U(g), y = U(g), z = U(g);
            const double v = f(x, y, z).a[0];
            constant = constant && std::fabs(v - c.a[0]) <= 1e-14 * std::max(1.0, std::fabs(c.a[0]));
            const double ss = c.a[0] * std::sin(M_PI * x) * std::sin(M_PI * y);
            sinsin = sinsin && std::fabs(v - ss) <= 1e-12 * std::max(1.0, std::fabs(c.a[0]));
        }
        // linearity in v
        P.v = scal(2.0);
        const double v2 = f(0.5, 0.5, 0.5).a[0];
        if (std::fabs(v2 - 2 * c.a[0]) > 1e-12 * std::max(1.0, std::fabs(c.a[0])))
            throw std::invalid_argument("linear form is not linear in the test function");
        if (constant) { d.kind = FEEL_HIP_SRC_CONST; d.params[0] = c.a[0]; }
        else if (sinsin) { d.kind = FEEL_HIP_SRC_SINSIN; d.params[0] = c.a[0]; }
        else throw std::invalid_argument("source outside the supported subset (constant, c*sin(pi x)sin(pi y))");
    } else {
        d.kind = FEEL_HIP_SRC_VECTOR;
        for (int comp = 0; comp < nc; ++comp) {
            P.v = zeros(nc, 1);
            P.v.a[comp] = 1.0;
            Val c = eval(root, P);
            if (!c.scalar()) throw std::invalid_argument("linear form is not scalar");
            d.params[comp] = c.a[0];
        }
        if (hasOp(root, Node::PX) || hasOp(root, Node::PY) || hasOp(root, Node::PZ) || hasOp(root, Node::SIN))
            throw std::invalid_argument("vector sources must be constant");
    }
    check(feel_hip_assemble_linear(M_test->handle(), &d, M_vec.get(), add ? 1 : 0));
}

// ---------------------------------------------------------------- bilinear form
BilinearForm::BilinearForm(space_ptr trial, space_ptr test, int precision) : M_trial(std::move(trial)), M_test(std::move(test)) {
    if (M_trial != M_test) throw std::invalid_argument("trial and test spaces must be the same space");
    // Backend::newMatrix -> Stencil::computeGraph -> MatrixSparse::init (backend.hpp:349-440)
    M_graph = M_test->graph();
    feel_hip_mat* A = nullptr;
    check(feel_hip_mat_create(M_graph.get(), precision, &A));
    auto keepGraph = M_graph;
    M_mat = std::shared_ptr<feel_hip_mat>(A, [keepGraph](feel_hip_mat* p) { feel_hip_mat_destroy(p); });
}

std::vector<double> BilinearForm::values() const {
    std::vector<double> out(feel_hip_graph_nnz(M_graph.get()));
    check(feel_hip_mat_get_values(M_mat.get(), FEEL_HIP_FP64, out.data()));
    return out;
}

void BilinearForm::assemble(IntegralExpr const& ie, bool add) {
    if (ie.range.mesh != M_test->mesh()) throw std::invalid_argument("integration range is not the space's mesh");
    Node const& root = *ie.expr.n;
    const int k = M_test->order(), nc = M_test->nComponents(), dim = M_test->mesh()->dim();
    const int64_t ne = M_test->mesh()->numElements();
    auto const ctx = M_test->backendPtr()->ctx();
    const bool perElement = hasOp(root, Node::IDV);
    if (hasOp(root, Node::PX) || hasOp(root, Node::PY) || hasOp(root, Node::PZ) || hasOp(root, Node::SIN))
        throw std::invalid_argument("bilinear coefficients must be constants or P0 fields");
    feel_hip_bilinear_desc d{};
    d.quad_order = ie.quad >= 0 ? ie.quad : polyOrder(root, k);
    d.precision = 0;
    Probe P;
    auto evalAt = [&](int64_t e) { P.elem = e; return eval(root, P); };
    std::vector<double> c0(ne), c1(ne);
    std::shared_ptr<feel_hip_vec> v0, v1;
    if (nc == 1) {
        // b(u,v) = alpha u v + beta grad u . grad v; verify on random probes
        auto setProbe = [&](double u, std::array<double, 3> gu, double v, std::array<double, 3> gv) {
            P.u = scal(u); P.v = scal(v);
            P.gu = zeros(1, dim); P.gv = zeros(1, dim);
            for (int i = 0; i < dim; ++i) { P.gu.a[i] = gu[i]; P.gv.a[i] = gv[i]; }
        };
        const int64_t nEval = perElement ? ne : 1;
        for (int64_t e = 0; e < nEval; ++e) {
            setProbe(1, {0, 0, 0}, 1, {0, 0, 0});
            c0[e] = evalAt(e).a[0];
            setProbe(0, {1, 0, 0}, 0, {1, 0, 0});
            c1[e] = evalAt(e).a[0];
        }
        if (!perElement) { std::fill(c0.begin(), c0.end(), c0[0]); std::fill(c1.begin(), c1.end(), c1[0]); }
        std::mt19937_64 g(11);
        std::uniform_real_distribution<double> U(-1, 1);
        for (int t = 0; t < 6; ++t) {
            const int64_t e = perElement ? (int64_t)(t * 7919 % ne) : 0;
            std::array<double, 3> gu{U(g), U(g), U(g)}, gv{U(g), U(g), U(g)};
            if (dim == 2) gu[2] = gv[2] = 0;
            const double u = U(g), v = U(g);
            setProbe(u, gu, v, gv);
            Val r = evalAt(e);
            if (!r.scalar()) throw std::invalid_argument("bilinear form is not scalar");
            const double expect = c0[e] * u * v + c1[e] * (gu[0] * gv[0] + gu[1] * gv[1] + gu[2] * gv[2]);
            if (std::fabs(r.a[0] - expect) > 1e-12 * (std::fabs(c0[e]) + std::fabs(c1[e]) + 1e-300))
                throw std::invalid_argument("bilinear form outside the supported subset (mass, isotropic diffusion)");
        }
        bool hasMass = false, hasDiff = false;
        for (int64_t e = 0; e < ne; ++e) { hasMass |= c0[e] != 0; hasDiff |= c1[e] != 0; }
        d.kind = (hasMass ? FEEL_HIP_EXPR_MASS : 0) | (hasDiff ? FEEL_HIP_EXPR_DIFFUSION : 0);
        if (!d.kind) throw std::invalid_argument("bilinear form is identically zero");
        d.mass_scale = 1.0;
        d.diff_scale = 1.0;
        if (hasMass) { v0 = makeVec(ctx, ne); check(feel_hip_vec_set(v0.get(), c0.data())); d.mass_coef = v0.get(); }
        if (hasDiff) { v1 = makeVec(ctx, ne); check(feel_hip_vec_set(v1.get(), c1.data())); d.diff_coef = v1.get(); }
        M_recognised = std::string(hasMass ? "mass" : "") + (hasMass && hasDiff ? "+" : "") + (hasDiff ? "diffusion" : "");
    } else {
        // b(u,v) = lambda tr(grad u) tr(grad v) + mu (grad u + grad u^T) : grad v
        auto setProbe = [&](Val gu, Val gv) {
            P.u = zeros(3, 1); P.v = zeros(3, 1);
            P.gu = gu; P.gv = gv;
        };
        auto unit = [](int i, int j) { Val m = zeros(3, 3); m(i, j) = 1; return m; };
        const int64_t nEval = perElement ? ne : 1;
        for (int64_t e = 0; e < nEval; ++e) {
            setProbe(unit(0, 0), unit(1, 1));
            c0[e] = evalAt(e).a[0];   // lambda
            setProbe(unit(0, 1), unit(0, 1));
            c1[e] = evalAt(e).a[0];   // mu
        }
        if (!perElement) { std::fill(c0.begin(), c0.end(), c0[0]); std::fill(c1.begin(), c1.end(), c1[0]); }
        std::mt19937_64 g(13);
        std::uniform_real_distribution<double> U(-1, 1);
        for (int t = 0; t < 6; ++t) {
            const int64_t e = perElement ? (int64_t)(t * 7919 % ne) : 0;
            Val gu = zeros(3, 3), gv = zeros(3, 3), uu = zeros(3, 1), vv = zeros(3, 1);
            for (int i = 0; i < 9; ++i) { gu.a[i] = U(g); gv.a[i] = U(g); }
            for (int i = 0; i < 3; ++i) { uu.a[i] = U(g); vv.a[i] = U(g); }
            P.u = uu; P.v = vv; P.gu = gu; P.gv = gv;
            Val r = evalAt(e);
            if (!r.scalar()) throw std::invalid_argument("bilinear form is not scalar");
            double tu = gu(0, 0) + gu(1, 1) + gu(2, 2), tv = gv(0, 0) + gv(1, 1) + gv(2, 2), s = 0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) s += (gu(i, j) + gu(j, i)) * gv(i, j);
            const double expect = c0[e] * tu * tv + c1[e] * s;
            if (std::fabs(r.a[0] - expect) > 1e-12 * (std::fabs(c0[e]) + std::fabs(c1[e]) + 1e-300))
                throw std::invalid_argument("bilinear form outside the supported subset (isotropic elasticity)");
        }
        // lambda, mu -> E, nu
        std::vector<double> E(ne), nu(ne);
        for (int64_t e = 0; e < ne; ++e) {
            const double lam = c0[e], mu = c1[e];
            E[e] = mu * (3 * lam + 2 * mu) / (lam + mu);
            nu[e] = lam / (2 * (lam + mu));
        }
        d.kind = FEEL_HIP_EXPR_ELASTICITY;
        v0 = makeVec(ctx, ne); check(feel_hip_vec_set(v0.get(), E.data())); d.young = v0.get();
        v1 = makeVec(ctx, ne); check(feel_hip_vec_set(v1.get(), nu.data())); d.poisson = v1.get();
        M_recognised = "elasticity";
    }
    check(feel_hip_assemble_bilinear(M_mat.get(), &d, add ? 1 : 0));
    check(feel_hip_ctx_synchronize(ctx));   // coefficient vectors are released on return
}

BilinearForm& BilinearForm::operator+=(OnExpr const& o) {
    auto const& m = M_test->mesh()->m;
    auto const& dt = M_test->dofTable();
    if (o.range.mesh != M_test->mesh()) throw std::invalid_argument("on() range is not the space's mesh");
    std::vector<uint8_t> flags = feelpp_amd::boundaryDofs(m, dt, o.range.sel);
    std::vector<int32_t> rows;
    for (int64_t i = 0; i < dt.ndof; ++i)
        if (flags[i]) rows.push_back((int32_t)i);
    // g at the dof nodes (interpolant of _expr); one placement per dof is enough
    std::vector<double> gval(dt.ndof, 0.0);
    const Node& root = *o.expr.n;
    if (root.op != Node::CST || root.c != 0.0) {
        auto nodes = feelpp_amd::lagrangeNodes(m.dim, dt.order);
        const int nvx = m.dim + 1;
        Probe P;
        for (int64_t e = 0; e < m.ne; ++e)
            for (int l = 0; l < dt.nloc; ++l) {
                double lam[4], s = 0;
                for (int c = 0; c < m.dim; ++c) { lam[c + 1] = 0.5 * (1 + nodes[l * m.dim + c]); s += lam[c + 1]; }
                lam[0] = 1 - s;
                for (int c = 0; c < 3; ++c) P.x[c] = 0;
                for (int v = 0; v < nvx; ++v)
                    for (int c = 0; c < m.dim; ++c) P.x[c] += lam[v] * m.coords[(size_t)m.elems[e * nvx + v] * m.dim + c];
                P.elem = e;
                Val gv = eval(root, P);
                for (int comp = 0; comp < dt.ncomp; ++comp)
                    gval[dt.elem2dof[e * dt.nldof + dt.nloc * comp + l]] = gv.scalar() ? gv.a[0] : gv.a[comp];
            }
    }
    auto ctx = M_test->backendPtr()->ctx();
    feel_hip_rows* r = nullptr;
    check(feel_hip_rows_create(ctx, dt.ndof, (int64_t)rows.size(), rows.data(), &r));
    std::shared_ptr<feel_hip_rows> keep(r, [](feel_hip_rows* p) { feel_hip_rows_destroy(p); });
    auto gv = makeVec(ctx, dt.ndof);
    check(feel_hip_vec_set(gv.get(), gval.data()));
    const int mode = M_onType == "elimination_symmetric" ? FEEL_HIP_ON_ELIMINATION_SYMMETRIC : FEEL_HIP_ON_ELIMINATION;
    check(feel_hip_mat_zero_rows(M_mat.get(), r, gv.get(), o.rhs.vec(), mode, 1.0));   // value_on_diagonal (on.hpp:347)
    check(feel_hip_ctx_synchronize(ctx));
    return *this;
}

SolveResult BilinearForm::solveImpl(LinearForm& rhs, Element& sol, double rtol, int maxit, int prec) {
    feel_hip_solve_opts o;
    feel_hip_solve_opts_default(&o);
    o.rtol = rtol;
    o.maxit = maxit;
    o.precision = prec;
    feel_hip_solve_data sd;
    check(feel_hip_solve_pcg(M_mat.get(), rhs.vec(), sol.vec(), &o, &sd));
    SolveResult r;
    r.converged = sd.converged != 0;
    r.iterations = sd.its;
    r.residual = sd.residual;
    r.reason = sd.reason;
    r.outer = sd.outer_its;
    return r;
}

}  // namespace feelhip
