#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace feelpp_amd {

// Condition number of D^{-1} A estimated from the coefficients of k CG
// iterations: the Lanczos tridiagonal T_k (diagonal 1/a_j + b_{j-1}/a_{j-1},
// off-diagonal sqrt(b_j)/a_j with a_j = gamma_j / pAp_j, b_j = gamma_{j+1} /
// gamma_j) has extreme eigenvalues that approach those of D^{-1} A from
// inside, so the ratio is a lower bound that grows with k.  The eigenvalues
// come from Sturm-sequence bisection, each to the relative precision rel.
double cgConditionEstimate(const double* hist, int k, double rel) {
    if (k < 2) return 1.0;
    std::vector<double> d(k), e(k, 0.0);
    for (int j = 0; j < k; ++j) {
        const int q = j;
        const double a = hist[2 * q] / hist[2 * q + 1];
        d[j] = 1.0 / a;
        if (j > 0) {
            const double ap = hist[2 * (q - 1)] / hist[2 * (q - 1) + 1];
            const double bp = hist[2 * q] / hist[2 * (q - 1)];
            d[j] += bp / ap;
            e[j - 1] = std::sqrt(bp) / ap;
        }
    }
    double lo = 1e300, hi = -1e300;
    for (int j = 0; j < k; ++j) {
        const double rad = (j > 0 ? std::fabs(e[j - 1]) : 0.0) + std::fabs(e[j]);
        lo = std::min(lo, d[j] - rad);
        hi = std::max(hi, d[j] + rad);
    }
    auto below = [&](double x) {   // eigenvalues of T_k below x
        int cnt = 0;
        double q = d[0] - x;
        for (int j = 0;; ) {
            if (q < 0) ++cnt;
            if (++j == k) break;
            if (q == 0) q = 1e-300;
            q = d[j] - x - e[j - 1] * e[j - 1] / q;
        }
        return cnt;
    };
    auto bisect = [&](int target) {   // smallest x with below(x) >= target
        double a = lo, b = hi;
        for (int i = 0; i < 200 && b - a > rel * std::max(std::fabs(a), std::fabs(b)); ++i) {
            const double m = 0.5 * (a + b);
            if (below(m) >= target) b = m; else a = m;
        }
        return 0.5 * (a + b);
    };
    const double lmin = bisect(1), lmax = bisect(k);
    return lmin > 0 ? lmax / lmin : 1e300;
}

}  // namespace feelpp_amd
