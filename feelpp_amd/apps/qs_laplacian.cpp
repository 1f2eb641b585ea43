// Quickstart Laplacian on the MI355X backend, following cgLaplacian
// (feelpp/feel/feelpde/cg_laplacian.hpp:94-147): l = integrate(f*id(v));
// a = integrate(inner(k*gradt(u),grad(v))) [+ mass]; a += on(boundary, g = 0);
// a.solve(); normL2 / normSemiH1 of the solution (and its L2 error when the
// exact solution is known).  Prints one JSON line; --out PREFIX writes
// PREFIX.sol (raw doubles).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "../csrc/host/feel.hpp"

using namespace feelhip;

int main(int argc, char** argv) {
    int dim = 3, n = 8, order = 2, seed = 2024;
    double jitter = 0.15, rtol = 1e-10, mass = 0.0;
    std::string exportPath;
    std::string kappa = "lognormal", source = "one", precision = "fp64", ontype = "elimination_symmetric", out, mesh_path;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "--dim") dim = std::stoi(next());
        else if (a == "--n") n = std::stoi(next());
        else if (a == "--order") order = std::stoi(next());
        else if (a == "--seed") seed = std::stoi(next());
        else if (a == "--jitter") jitter = std::stod(next());
        else if (a == "--rtol") rtol = std::stod(next());
        else if (a == "--mass") mass = std::stod(next());
        else if (a == "--kappa") kappa = next();
        else if (a == "--source") source = next();
        else if (a == "--precision") precision = next();
        else if (a == "--on.type") ontype = next();
        else if (a == "--mesh") mesh_path = next();
        else if (a == "--out") out = next();
        else if (a == "--export") exportPath = next();
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    try {
        auto t0 = std::chrono::steady_clock::now();
        auto toc = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
        mesh_ptr mesh = !mesh_path.empty() ? loadMesh(mesh_path, dim)
                        : dim == 2 ? unitSquare(n, jitter, seed) : kuhnBox(n, n, n, 1, 1, 1, jitter, seed);
        auto Vh = Pch(mesh, order);
        auto u = Vh->element();
        auto v = Vh->element();
        P0Field k = p0(kappa == "lognormal" ? feelpp_amd::lognormalField(mesh->numElements(), 0.0, 1.0, seed + 1)
                                            : std::vector<double>(mesh->numElements(), 1.0));
        double tSetup = toc();
        auto l = form1(_test = Vh);
        Expr f = source == "sinsin" ? sin(M_PI * Px()) * sin(M_PI * Py()) : cst(1.0);
        l = integrate(_range = elements(mesh), _expr = f * id(v), _quad = source == "sinsin" ? 2 : order);
        auto a = form2(_trial = Vh, _test = Vh);
        a.setOnType(ontype);
        Expr e = inner(idv(k) * gradt(u), grad(v));
        if (mass != 0.0) e = e + mass * idt(u) * id(v);
        a = integrate(_range = elements(mesh), _expr = e);
        a += on(_range = boundaryfaces(mesh), _rhs = l, _element = u, _expr = cst(0.));
        double tAssemble = toc() - tSetup;
        SolveResult r = a.solve(_rhs = l, _solution = u, _rtol = rtol, _maxit = 100000,
                                _precision = precision == "fp32" ? (int)FEEL_HIP_FP32 : (int)FEEL_HIP_FP64);
        double tSolve = toc() - tSetup - tAssemble;
        auto x = u.values();
        if (!out.empty()) {
            std::ofstream os(out + ".sol", std::ios::binary);
            os.write(reinterpret_cast<const char*>(x.data()), sizeof(double) * x.size());
        }
        if (!exportPath.empty()) exportGmsh(exportPath, u, "u");
        double nrm = 0;
        for (double xi : x) nrm += xi * xi;
        // post-solve evaluation: normL2 and H1 seminorm of u; with k = 1 and the
        // sin(pi x) sin(pi y) source in 2D the exact solution is f / (2 pi^2 + mass)
        const double l2 = normL2(elements(mesh), u), h1 = normSemiH1(elements(mesh), u);
        double l2err = -1.0;
        if (dim == 2 && source == "sinsin" && kappa == "one")
            l2err = normL2(elements(mesh), u, {sinsinFn(1.0 / (2.0 * M_PI * M_PI + mass))});
        std::printf("{\"app\": \"qs_laplacian\", \"dim\": %d, \"order\": %d, \"elements\": %lld, \"ndof\": %lld, "
                    "\"form\": \"%s\", \"converged\": %d, \"its\": %d, \"outer\": %d, \"residual\": %.6e, "
                    "\"norm\": %.17g, \"norm_l2\": %.17g, \"norm_semi_h1\": %.17g, \"l2_error\": %.17g, "
                    "\"setup_s\": %.3f, \"assemble_s\": %.4f, \"solve_s\": %.4f}\n",
                    dim, order, (long long)mesh->numElements(), (long long)Vh->nDof(), a.recognised().c_str(),
                    (int)r.converged, r.iterations, r.outer, r.residual, std::sqrt(nrm), l2, h1, l2err, tSetup,
                    tAssemble, tSolve);
        return r.converged ? 0 : 1;
    } catch (std::exception const& ex) {
        std::fprintf(stderr, "qs_laplacian: %s\n", ex.what());
        return 3;
    }
}
