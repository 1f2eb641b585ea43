// Feel++-shaped C++ surface over the device C-ABI:
//
//   auto mesh = kuhnBox(...);                      // or loadMesh(path, 3)
//   auto Vh = Pch<2>(mesh);  auto u = Vh->element(), v = Vh->element();
//   auto l = form1(_test = Vh);
//   l = integrate(_range = elements(mesh), _expr = cst(1.) * id(v));
//   auto a = form2(_trial = Vh, _test = Vh);
//   a = integrate(_range = elements(mesh), _expr = inner(idv(k) * gradt(u), grad(v)));
//   a += on(_range = boundaryfaces(mesh), _rhs = l, _element = u, _expr = cst(0.));
//   a.solve(_rhs = l, _solution = u);
//
// mirroring feelvf/form.hpp:92-146, feelvf/integrate.hpp:64-98,
// feelvf/bilinearform.hpp:1437-1468, feelvf/on.hpp, feelvf/bilinearformbase.hpp:446-466.
// Expressions are small runtime trees; the supported subset (mass,
// diffusion with a P0 coefficient, isotropic elasticity with P0 lambda/mu,
// constant or sin(pi x)sin(pi y) sources) is recognised by evaluating the
// expression on probe values of the trial/test functions, so inner(k*gradt(u),grad(v))
// and k*gradt(u)*trans(grad(v)) are the same form.  Anything else throws.
#pragma once
#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "feelpp_hip.h"
#include "feelpp_host.h"
#include "fe.hpp"
#include "space.hpp"

namespace feelhip {

using feelpp_amd::SimplexMesh;

inline void check(int rc) {
    if (rc != FEEL_HIP_OK) throw std::runtime_error(std::string("feelpp_hip: ") + feel_hip_last_error());
}

// ------------------------------------------------------------------ named arguments
template <class Tag, class T>
struct Arg {
    T value;
};
template <class Tag>
struct Key {
    template <class T>
    Arg<Tag, std::decay_t<T>> operator=(T&& v) const { return {std::forward<T>(v)}; }
};
struct test_t; struct trial_t; struct range_t; struct expr_t; struct quad_t; struct rhs_t; struct element_t;
struct solution_t; struct rtol_t; struct maxit_t; struct precision_t; struct type_t;
inline constexpr Key<test_t> _test{};
inline constexpr Key<trial_t> _trial{};
inline constexpr Key<range_t> _range{};
inline constexpr Key<expr_t> _expr{};
inline constexpr Key<quad_t> _quad{};
inline constexpr Key<rhs_t> _rhs{};
inline constexpr Key<element_t> _element{};
inline constexpr Key<solution_t> _solution{};
inline constexpr Key<rtol_t> _rtol{};
inline constexpr Key<maxit_t> _maxit{};
inline constexpr Key<precision_t> _precision{};
inline constexpr Key<type_t> _type{};

namespace detail {
template <class Tag, class Default>
Default pick(Default d) { return d; }
template <class Tag, class Default, class T, class... Rest>
auto pick(Default d, Arg<Tag, T> const& a, Rest const&...) { return a.value; }
template <class Tag, class Default, class First, class... Rest>
auto pick(Default d, First const&, Rest const&... rest) { return pick<Tag>(d, rest...); }
struct none_t {};
}  // namespace detail

// ------------------------------------------------------------------ backend / mesh / spaces
class Backend {
   public:
    explicit Backend(int device = 0) { check(feel_hip_ctx_create(device, &M_ctx)); }
    ~Backend() { feel_hip_ctx_destroy(M_ctx); }
    Backend(Backend const&) = delete;
    Backend& operator=(Backend const&) = delete;
    feel_hip_ctx* ctx() const { return M_ctx; }
    // Backend::build("hip") (feelalg/backend.cpp:231-245): only the hip kind exists here
    static std::shared_ptr<Backend> build(std::string const& kind = "hip", int device = 0) {
        if (kind != "hip") throw std::invalid_argument("backend kind '" + kind + "' is not provided by feelpp_amd");
        return std::make_shared<Backend>(device);
    }

   private:
    feel_hip_ctx* M_ctx = nullptr;
};
inline std::shared_ptr<Backend>& backend() {
    static std::shared_ptr<Backend> b = Backend::build("hip");
    return b;
}

struct Mesh {
    SimplexMesh m;
    int dim() const { return m.dim; }
    int64_t numElements() const { return m.ne; }
};
using mesh_ptr = std::shared_ptr<Mesh>;
inline mesh_ptr unitSquare(int n, double jitter = 0.0, uint64_t seed = 0) {
    return std::make_shared<Mesh>(Mesh{feelpp_amd::unitSquareMesh(n, jitter, seed)});
}
inline mesh_ptr kuhnBox(int nx, int ny, int nz, double lx = 1, double ly = 1, double lz = 1, double jitter = 0.0,
                        uint64_t seed = 0) {
    return std::make_shared<Mesh>(Mesh{feelpp_amd::kuhnBoxMesh(nx, ny, nz, lx, ly, lz, jitter, seed)});
}
mesh_ptr loadMesh(std::string const& path, int dim);

struct ElementsRange { mesh_ptr mesh; };
struct FacesRange { mesh_ptr mesh; int sel; };
inline ElementsRange elements(mesh_ptr m) { return {m}; }
inline FacesRange boundaryfaces(mesh_ptr m) { return {m, 0}; }
// markedfaces(mesh, "x=0"): the boundary faces on the plane x = 0 (the clamped end of the beam)
inline FacesRange markedfaces(mesh_ptr m, std::string const& marker) {
    if (marker != "x=0") throw std::invalid_argument("only the marker \"x=0\" is defined on generated meshes");
    return {m, 1};
}

class FunctionSpace;
using space_ptr = std::shared_ptr<FunctionSpace>;

// A finite element function: device vector of dofs (VectorHip)
class Element {
   public:
    Element() = default;
    Element(space_ptr s, std::shared_ptr<feel_hip_vec> v, int id) : M_space(std::move(s)), M_vec(std::move(v)), M_id(id) {}
    space_ptr functionSpace() const { return M_space; }
    feel_hip_vec* vec() const { return M_vec.get(); }
    int id() const { return M_id; }
    std::vector<double> values() const;
    void setValues(std::vector<double> const& v);
    void zero() { check(feel_hip_vec_zero(M_vec.get())); }

   private:
    space_ptr M_space;
    std::shared_ptr<feel_hip_vec> M_vec;
    int M_id = -1;
};

class FunctionSpace : public std::enable_shared_from_this<FunctionSpace> {
   public:
    FunctionSpace(mesh_ptr mesh, int order, int ncomp, std::shared_ptr<Backend> be);
    ~FunctionSpace() { feel_hip_space_destroy(M_space); }
    Element element();
    int64_t nDof() const { return M_dt.ndof; }
    int order() const { return M_dt.order; }
    int nComponents() const { return M_dt.ncomp; }
    mesh_ptr mesh() const { return M_mesh; }
    feelpp_amd::DofTable const& dofTable() const { return M_dt; }
    feel_hip_space* handle() const { return M_space; }
    std::shared_ptr<Backend> backendPtr() const { return M_backend; }
    // the COUPLED sparsity pattern of the space, built on first use
    std::shared_ptr<feel_hip_graph> graph();

   private:
    mesh_ptr M_mesh;
    feelpp_amd::DofTable M_dt;
    std::shared_ptr<Backend> M_backend;
    feel_hip_space* M_space = nullptr;
    std::shared_ptr<feel_hip_graph> M_graph;
    int M_next_id = 0;
};
// Reference functions of the error norms (per component)
inline feel_hip_function linearFn(double c0, double cx = 0, double cy = 0, double cz = 0) {
    return feel_hip_function{FEEL_HIP_FN_LINEAR, {c0, cx, cy, cz}};
}
inline feel_hip_function sinsinFn(double c0) { return feel_hip_function{FEEL_HIP_FN_SINSIN, {c0, 0, 0, 0}}; }

// normL2(_range=elements(mesh), _expr=idv(u) - f) (feelvf/norml2.hpp:37-57); f per
// component, empty for f = 0; quad < 0 takes the order of the squared integrand
inline double normL2(ElementsRange const&, Element const& u, std::vector<feel_hip_function> const& f = {},
                     int quad = -1) {
    double out = 0;
    check(feel_hip_norm_l2(u.functionSpace()->handle(), u.vec(), f.empty() ? nullptr : f.data(), quad, &out));
    return out;
}
// H1 seminorm of u - f (feelvf/normsemih1.hpp)
inline double normSemiH1(ElementsRange const&, Element const& u, std::vector<feel_hip_function> const& f = {},
                         int quad = -1) {
    double out = 0;
    check(feel_hip_norm_semi_h1(u.functionSpace()->handle(), u.vec(), f.empty() ? nullptr : f.data(), quad, &out));
    return out;
}

// Gmsh export of a field: the mesh and an $ElementNodeData block with the
// values at element vertices (feelfilters/exportergmsh_impl.hpp:1210-1324)
inline void exportGmsh(std::string const& path, Element const& u, std::string const& name = "u", double time = 0.0,
                       int step = 0) {
    auto Vh = u.functionSpace();
    auto const& m = Vh->mesh()->m;
    auto const& dt = Vh->dofTable();
    const auto vals = u.values();
    check(feel_mesh_write_msh22_solution(path.c_str(), m.dim, m.nv, m.coords.data(), m.ne, m.elems.data(), dt.order,
                                         dt.ncomp, dt.elem2dof.data(), vals.data(), name.c_str(), time, step));
}

template <int K>
space_ptr Pch(mesh_ptr mesh, std::shared_ptr<Backend> be = backend()) {
    return std::make_shared<FunctionSpace>(mesh, K, 1, be);
}
inline space_ptr Pch(mesh_ptr mesh, int k, std::shared_ptr<Backend> be = backend()) {
    return std::make_shared<FunctionSpace>(mesh, k, 1, be);
}
inline space_ptr Pchv(mesh_ptr mesh, int k, std::shared_ptr<Backend> be = backend()) {
    return std::make_shared<FunctionSpace>(mesh, k, mesh->dim(), be);
}
template <int K>
space_ptr Pchv(mesh_ptr mesh, std::shared_ptr<Backend> be = backend()) {
    return std::make_shared<FunctionSpace>(mesh, K, mesh->dim(), be);
}

// P0 (element-wise) coefficient field, as Pdh<0> elements are used with idv()
struct P0Field {
    std::shared_ptr<std::vector<double>> values;
};
inline P0Field p0(std::vector<double> v) { return {std::make_shared<std::vector<double>>(std::move(v))}; }

// ------------------------------------------------------------------ expressions
struct Node;
using node_ptr = std::shared_ptr<const Node>;
struct Expr {
    node_ptr n;
};
struct Node {
    enum Op { CST, CSTV, IDT, ID, GRADT, GRAD, IDV, PX, PY, PZ, SIN, ADD, SUB, MUL, NEG, INNER, TRANS, SYM, TRACE, EYE };
    Op op;
    double c = 0;
    double cv[3] = {0, 0, 0};   // CSTV
    int field = -1;       // IDV: index into the expression's field list
    int trial = -1, test = -1;
    P0Field f;
    std::vector<node_ptr> kids;
};

Expr cst(double c);
// constant column vector, as expr<3,1>("{a,b,c}")
Expr vec(double a, double b, double c);
Expr idt(Element const& u);
Expr id(Element const& v);
Expr gradt(Element const& u);
Expr grad(Element const& v);
Expr idv(P0Field const& k);
Expr Px();
Expr Py();
Expr Pz();
Expr sin(Expr const& a);
Expr inner(Expr const& a, Expr const& b);
Expr trans(Expr const& a);
Expr sym(Expr const& a);
Expr trace(Expr const& a);
Expr eye3();
Expr operator+(Expr const& a, Expr const& b);
Expr operator-(Expr const& a, Expr const& b);
Expr operator-(Expr const& a);
Expr operator*(Expr const& a, Expr const& b);
Expr operator*(double a, Expr const& b);
Expr operator*(Expr const& a, double b);

struct IntegralExpr {
    ElementsRange range;
    Expr expr;
    int quad = -1;   // -1: deduced from the polynomial order (expr.hpp:1058-1076)
};
template <class... A>
IntegralExpr integrate(A const&... a) {
    IntegralExpr ie;
    ie.range = detail::pick<range_t>(ElementsRange{}, a...);
    ie.expr = detail::pick<expr_t>(Expr{}, a...);
    ie.quad = detail::pick<quad_t>(-1, a...);
    if (!ie.range.mesh || !ie.expr.n) throw std::invalid_argument("integrate needs _range and _expr");
    return ie;
}

// ------------------------------------------------------------------ forms
class LinearForm {
   public:
    LinearForm() = default;
    explicit LinearForm(space_ptr test);
    LinearForm& operator=(IntegralExpr const& ie) { assemble(ie, false); return *this; }
    LinearForm& operator+=(IntegralExpr const& ie) { assemble(ie, true); return *this; }
    feel_hip_vec* vec() const { return M_vec.get(); }
    space_ptr test() const { return M_test; }
    std::vector<double> values() const;

   private:
    void assemble(IntegralExpr const& ie, bool add);
    space_ptr M_test;
    std::shared_ptr<feel_hip_vec> M_vec;
};

struct OnExpr {
    FacesRange range;
    LinearForm rhs;     // shares the device vector of the caller's form
    Element element;
    Expr expr;
};
template <class... A>
OnExpr on(A const&... a) {
    OnExpr o;
    o.range = detail::pick<range_t>(FacesRange{}, a...);
    o.rhs = detail::pick<rhs_t>(LinearForm{}, a...);
    o.element = detail::pick<element_t>(Element{}, a...);
    o.expr = detail::pick<expr_t>(cst(0.0), a...);
    if (!o.range.mesh || !o.rhs.vec()) throw std::invalid_argument("on needs _range and _rhs");
    return o;
}

struct SolveResult {
    bool converged = false;
    int iterations = 0;
    double residual = 0;
    int reason = 0;
    int outer = 0;
};

class BilinearForm {
   public:
    BilinearForm(space_ptr trial, space_ptr test, int precision = FEEL_HIP_FP64_FP32);
    BilinearForm& operator=(IntegralExpr const& ie) { assemble(ie, false); return *this; }
    BilinearForm& operator+=(IntegralExpr const& ie) { assemble(ie, true); return *this; }
    BilinearForm& operator+=(OnExpr const& o);
    // a.solve(_rhs = l, _solution = u [, _rtol, _maxit, _precision]) -> Backend::solve with
    // ksp-type=cg, pc-type=jacobi (bilinearformbase.hpp:446-466)
    template <class... A>
    SolveResult solve(A const&... a) {
        LinearForm rhs = detail::pick<rhs_t>(LinearForm{}, a...);
        Element sol = detail::pick<solution_t>(Element{}, a...);
        const double rtol = detail::pick<rtol_t>(1e-8, a...);
        const int maxit = detail::pick<maxit_t>(1000, a...);
        const int prec = detail::pick<precision_t>((int)FEEL_HIP_FP64, a...);
        if (!rhs.vec() || !sol.vec()) throw std::invalid_argument("solve needs _rhs and _solution");
        return solveImpl(rhs, sol, rtol, maxit, prec);
    }
    feel_hip_mat* matrix() const { return M_mat.get(); }
    std::vector<double> values() const;
    std::string const& recognised() const { return M_recognised; }
    // on.type: elimination (default, options.cpp:338) or elimination_symmetric
    void setOnType(std::string const& t) { M_onType = t; }

   private:
    void assemble(IntegralExpr const& ie, bool add);
    SolveResult solveImpl(LinearForm& rhs, Element& sol, double rtol, int maxit, int prec);
    space_ptr M_trial, M_test;
    std::shared_ptr<feel_hip_graph> M_graph;
    std::shared_ptr<feel_hip_mat> M_mat;
    std::string M_recognised;
    std::string M_onType = "elimination";
};

template <class... A>
BilinearForm form2(A const&... a) {
    space_ptr test = detail::pick<test_t>(space_ptr{}, a...);
    space_ptr trial = detail::pick<trial_t>(test, a...);
    if (!test) test = trial;
    if (!test) throw std::invalid_argument("form2 needs _test and/or _trial");
    return BilinearForm(trial, test);
}
template <class... A>
LinearForm form1(A const&... a) {
    space_ptr test = detail::pick<test_t>(space_ptr{}, a...);
    if (!test) throw std::invalid_argument("form1 needs _test");
    return LinearForm(test);
}

}  // namespace feelhip
