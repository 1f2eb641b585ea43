// Quickstart linear elasticity on the MI355X backend (quickstart/qs_elasticity.cpp:41-88):
// a = integrate(inner(sigmat, grad(v))), sigmat = lambda*trace(deft)*Id + 2*mu*deft,
// deft = sym(gradt(u)); l = integrate(inner(f, id(v))); clamped at x = 0.
// Beam 1 x 0.25 x 0.25 (Kuhn), E log-uniform in [1,100], nu uniform in [0.2,0.35].
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>

#include "../csrc/host/feel.hpp"

using namespace feelhip;

int main(int argc, char** argv) {
    int nx = 40, ny = 10, nz = 10, order = 2, seed = 7;
    double rtol = 1e-10, jitter = 0.0;
    std::string precision = "fp64", out, exportPath;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "--nx") nx = std::stoi(next());
        else if (a == "--ny") ny = std::stoi(next());
        else if (a == "--nz") nz = std::stoi(next());
        else if (a == "--order") order = std::stoi(next());
        else if (a == "--seed") seed = std::stoi(next());
        else if (a == "--jitter") jitter = std::stod(next());
        else if (a == "--rtol") rtol = std::stod(next());
        else if (a == "--precision") precision = next();
        else if (a == "--out") out = next();
        else if (a == "--export") exportPath = next();
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    try {
        auto t0 = std::chrono::steady_clock::now();
        auto toc = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
        auto mesh = kuhnBox(nx, ny, nz, 1.0, 0.25, 0.25, jitter, seed);
        auto Vh = Pchv(mesh, order);
        auto u = Vh->element();
        auto v = Vh->element();
        const int64_t ne = mesh->numElements();
        auto E = feelpp_amd::uniformField(ne, 1.0, 100.0, true, seed + 1);
        auto nu = feelpp_amd::uniformField(ne, 0.2, 0.35, false, seed + 2);
        std::vector<double> lam(ne), mu(ne);
        for (int64_t e = 0; e < ne; ++e) {
            lam[e] = E[e] * nu[e] / ((1 + nu[e]) * (1 - 2 * nu[e]));
            mu[e] = E[e] / (2 * (1 + nu[e]));
        }
        auto lambda = p0(lam), mue = p0(mu);
        auto deft = sym(gradt(u));
        auto sigmat = idv(lambda) * trace(deft) * eye3() + 2.0 * idv(mue) * deft;
        double tSetup = toc();
        auto l = form1(_test = Vh);
        l = integrate(_range = elements(mesh), _expr = inner(vec(0., 0., -1.), id(v)), _quad = 2 * order);
        auto a = form2(_trial = Vh, _test = Vh);
        a.setOnType("elimination_symmetric");
        a = integrate(_range = elements(mesh), _expr = inner(sigmat, grad(v)), _quad = 2 * order);
        a += on(_range = markedfaces(mesh, "x=0"), _rhs = l, _element = u, _expr = cst(0.));
        double tAssemble = toc() - tSetup;
        SolveResult r = a.solve(_rhs = l, _solution = u, _rtol = rtol, _maxit = 200000,
                                _precision = precision == "fp32" ? (int)FEEL_HIP_FP32 : (int)FEEL_HIP_FP64);
        double tSolve = toc() - tSetup - tAssemble;
        auto x = u.values();
        if (!out.empty()) {
            std::ofstream os(out + ".sol", std::ios::binary);
            os.write(reinterpret_cast<const char*>(x.data()), sizeof(double) * x.size());
        }
        if (!exportPath.empty()) exportGmsh(exportPath, u, "displacement");
        std::printf("{\"app\": \"qs_elasticity\", \"order\": %d, \"elements\": %lld, \"ndof\": %lld, \"form\": \"%s\", "
                    "\"converged\": %d, \"its\": %d, \"outer\": %d, \"residual\": %.6e, \"setup_s\": %.3f, "
                    "\"assemble_s\": %.4f, \"solve_s\": %.4f}\n",
                    order, (long long)ne, (long long)Vh->nDof(), a.recognised().c_str(), (int)r.converged, r.iterations,
                    r.outer, r.residual, tSetup, tAssemble, tSolve);
        return r.converged ? 0 : 1;
    } catch (std::exception const& ex) {
        std::fprintf(stderr, "qs_elasticity: %s\n", ex.what());
        return 3;
    }
}
