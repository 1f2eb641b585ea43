// Post-solve evaluation: L2 norm and H1 seminorm of u - f over elements(mesh)
// (feelvf/norml2.hpp:37-57, feelvf/normsemih1.hpp), as a caller verifies a
// solution (python/pyfeelpp/tests/test_forms.py:44-53).
#include <cmath>
#include <vector>

#include "../host/fe.hpp"
#include "common.hpp"
#include "kernels.hpp"

using namespace feelpp_amd;

namespace {

// w[Q] | phi[nloc*Q] | dphi[(l*Q+q)*dim+a] | lambda[(dim+1)*Q], cached per quadrature order
int normTable(feel_hip_space* s, int qorder, const double** out, int* Q) {
    auto it = s->ntab.find(qorder);
    if (it != s->ntab.end()) { *out = it->second.first; *Q = s->nnpts[qorder]; return FEEL_HIP_OK; }
    QuadratureRule rule = quadrature(s->dim, qorder);
    Tabulation T = tabulate(s->dim, s->order, rule.pts);
    const int q = rule.npts, d = s->dim, nl = s->nloc;
    std::vector<double> h((size_t)q * (1 + nl + (size_t)nl * d + d + 1));
    double* w = h.data();
    double* phi = w + q;
    double* dphi = phi + (size_t)nl * q;
    double* lam = dphi + (size_t)nl * q * d;
    for (int k = 0; k < q; ++k) w[k] = rule.w[k];
    for (size_t i = 0; i < (size_t)nl * q; ++i) phi[i] = T.phi[i];
    for (size_t i = 0; i < (size_t)nl * q * d; ++i) dphi[i] = T.dphi[i];
    for (int k = 0; k < q; ++k) {
        double sum = 0;
        for (int c = 0; c < d; ++c) {
            lam[(c + 1) * q + k] = 0.5 * (1.0 + rule.pts[(size_t)k * d + c]);
            sum += lam[(c + 1) * q + k];
        }
        lam[k] = 1.0 - sum;
    }
    double* dptr = nullptr;
    if (s->ctx->alloc((void**)&dptr, sizeof(double) * h.size()) != FEEL_HIP_OK) return FEEL_HIP_ENOMEM;
    FEEL_HIP_CHECK(hipMemcpy(dptr, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    s->ntab[qorder] = {dptr, h.size()};
    s->nnpts[qorder] = q;
    *out = dptr;
    *Q = q;
    return FEEL_HIP_OK;
}

int norm(feel_hip_space* s, const feel_hip_vec* u, const feel_hip_function* f, int qorder, int semi, double* out) {
    if (!s || !u || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    if (u->n != s->ndof) return setError(FEEL_HIP_EINVAL, "vector size must equal the space's dofs");
    NormFn fn{};
    bool sinsin = false;
    for (int c = 0; c < s->ncomp; ++c) {
        fn.kind[c] = f ? f[c].kind : FEEL_HIP_FN_ZERO;
        if (fn.kind[c] < FEEL_HIP_FN_ZERO || fn.kind[c] > FEEL_HIP_FN_SINSIN)
            return setError(FEEL_HIP_EUNSUPPORTED, "unknown reference function kind");
        for (int k = 0; k < 4; ++k) fn.c[c][k] = f ? f[c].c[k] : 0.0;
        sinsin |= fn.kind[c] == FEEL_HIP_FN_SINSIN;
    }
    if (qorder < 0) qorder = 2 * s->order + (sinsin ? 2 : 0);
    if (qorder > 10) return setError(FEEL_HIP_EUNSUPPORTED, "quadrature order above 10");
    feel_hip_ctx* c = s->ctx;
    const double* tab = nullptr;
    int Q = 0;
    int rc = normTable(s, qorder, &tab, &Q);
    if (rc != FEEL_HIP_OK) return rc;
    double* part = c->part + 2 * kReduceBlocks;
    if (s->dim == 3)
        hipLaunchKernelGGL(kElemNorm<3>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, s->ne, s->nloc, s->ncomp, Q,
                           s->coords, s->elems, s->e2d, u->d, tab, fn, semi, part);
    else
        hipLaunchKernelGGL(kElemNorm<2>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, s->ne, s->nloc, s->ncomp, Q,
                           s->coords, s->elems, s->e2d, u->d, tab, fn, semi, part);
    FEEL_HIP_CHECK(hipGetLastError());
    std::vector<double> h(kReduceBlocks);
    FEEL_HIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * kReduceBlocks, hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    double a = 0.0;
    for (double v : h) a += v;   // fixed order
    *out = std::sqrt(std::fabs(a));
    return FEEL_HIP_OK;
}

}  // namespace

extern "C" {

int feel_hip_norm_l2(feel_hip_space* s, const feel_hip_vec* u, const feel_hip_function* f, int quad_order, double* out) {
    return norm(s, u, f, quad_order, 0, out);
}

int feel_hip_norm_semi_h1(feel_hip_space* s, const feel_hip_vec* u, const feel_hip_function* f, int quad_order,
                          double* out) {
    return norm(s, u, f, quad_order, 1, out);
}

}  // extern "C"
