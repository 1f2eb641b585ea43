// Transient heat equation driven the way the heat toolbox drives assembly and
// solve (toolboxes/feel/feelmodels/heat/heatassembly.hpp:115-131, BDF1):
//   a = integrate(rhoCp/dt * idt(T) * id(v) + inner(k * gradt(T), grad(v)))  once,
//   a += on(boundaryfaces, T = 0)                                            once,
//   each step: l = integrate(q * id(v)) + integrate(rhoCp * idv(T^n / dt) * id(v)),
//              on() re-applied to l, a.solve(l, T) warm-started from T^n.
// The transient rhs term is the mass matrix times T^n / dt.  Prints one JSON
// line; --out PREFIX writes PREFIX.sol, --export PATH a Gmsh file.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>

#include "../csrc/host/feel.hpp"

using namespace feelhip;

int main(int argc, char** argv) {
    int dim = 2, n = 32, order = 2, seed = 2024, steps = 10;
    double jitter = 0.1, rtol = 1e-10, dt = 0.01, rhoCp = 1.0, q = 1.0;
    std::string kappa = "one", precision = "fp64", out, exportPath;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "--dim") dim = std::stoi(next());
        else if (a == "--n") n = std::stoi(next());
        else if (a == "--order") order = std::stoi(next());
        else if (a == "--seed") seed = std::stoi(next());
        else if (a == "--steps") steps = std::stoi(next());
        else if (a == "--jitter") jitter = std::stod(next());
        else if (a == "--rtol") rtol = std::stod(next());
        else if (a == "--dt") dt = std::stod(next());
        else if (a == "--rhocp") rhoCp = std::stod(next());
        else if (a == "--q") q = std::stod(next());
        else if (a == "--kappa") kappa = next();
        else if (a == "--precision") precision = next();
        else if (a == "--out") out = next();
        else if (a == "--export") exportPath = next();
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    try {
        auto t0 = std::chrono::steady_clock::now();
        auto toc = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
        mesh_ptr mesh = dim == 2 ? unitSquare(n, jitter, seed) : kuhnBox(n, n, n, 1, 1, 1, jitter, seed);
        auto Vh = Pch(mesh, order);
        auto T = Vh->element();
        auto v = Vh->element();
        auto tmp = Vh->element();
        P0Field k = p0(kappa == "lognormal" ? feelpp_amd::lognormalField(mesh->numElements(), 0.0, 1.0, seed + 1)
                                            : std::vector<double>(mesh->numElements(), 1.0));
        auto l = form1(_test = Vh);
        auto f = form1(_test = Vh);
        f = integrate(_range = elements(mesh), _expr = cst(q) * id(v));
        auto m = form2(_trial = Vh, _test = Vh);
        m = integrate(_range = elements(mesh), _expr = idt(T) * id(v));
        auto a = form2(_trial = Vh, _test = Vh);
        a = integrate(_range = elements(mesh),
                      _expr = cst(rhoCp / dt) * idt(T) * id(v) + inner(idv(k) * gradt(T), grad(v)));
        a += on(_range = boundaryfaces(mesh), _rhs = l, _element = T, _expr = cst(0.));
        // boundary rows for re-applying on() to each step's right-hand side
        auto flags = feelpp_amd::boundaryDofs(mesh->m, Vh->dofTable(), 0);
        std::vector<int32_t> rows;
        for (size_t i = 0; i < flags.size(); ++i)
            if (flags[i]) rows.push_back((int32_t)i);
        feel_hip_rows* r = nullptr;
        check(feel_hip_rows_create(Vh->backendPtr()->ctx(), Vh->nDof(), (int64_t)rows.size(), rows.data(), &r));
        std::shared_ptr<feel_hip_rows> keep(r, [](feel_hip_rows* p) { feel_hip_rows_destroy(p); });
        const int prec = precision == "fp32" ? (int)FEEL_HIP_FP32 : (int)FEEL_HIP_FP64;
        const double tSetup = toc();
        long totalIts = 0;
        bool ok = true;
        for (int s = 0; s < steps; ++s) {
            check(feel_hip_vec_copy(l.vec(), f.vec()));
            check(feel_hip_mat_mult(m.matrix(), T.vec(), tmp.vec()));
            check(feel_hip_vec_axpy(l.vec(), rhoCp / dt, tmp.vec()));
            check(feel_hip_vec_set_rows(l.vec(), r, nullptr));
            SolveResult res = a.solve(_rhs = l, _solution = T, _rtol = rtol, _maxit = 100000, _precision = prec);
            totalIts += res.iterations;
            ok = ok && res.converged;
        }
        const double tLoop = toc() - tSetup;
        auto x = T.values();
        if (!out.empty()) {
            std::ofstream os(out + ".sol", std::ios::binary);
            os.write(reinterpret_cast<const char*>(x.data()), sizeof(double) * x.size());
        }
        if (!exportPath.empty()) exportGmsh(exportPath, T, "temperature", dt * steps, steps);
        std::printf("{\"app\": \"heat\", \"dim\": %d, \"order\": %d, \"elements\": %lld, \"ndof\": %lld, "
                    "\"form\": \"%s\", \"steps\": %d, \"converged\": %d, \"its_total\": %ld, \"norm_l2\": %.17g, "
                    "\"setup_s\": %.3f, \"s_per_step\": %.5f}\n",
                    dim, order, (long long)mesh->numElements(), (long long)Vh->nDof(), a.recognised().c_str(), steps,
                    (int)ok, totalIts, normL2(elements(mesh), T), tSetup, tLoop / std::max(steps, 1));
        return ok ? 0 : 1;
    } catch (std::exception const& ex) {
        std::fprintf(stderr, "heat: %s\n", ex.what());
        return 3;
    }
}
