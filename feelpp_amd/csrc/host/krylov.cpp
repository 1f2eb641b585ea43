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
namespace {

// The Lanczos tridiagonal T_k of a CG run (diagonal d, off-diagonal e), its
// Gershgorin interval, and Sturm counts of its eigenvalues.
struct Lanczos {
    int k;
    std::vector<double> d, e;
    double lo = 1e300, hi = -1e300;
    Lanczos(const double* hist, int k_) : k(k_), d(k_), e(k_, 0.0) {
        for (int j = 0; j < k; ++j) {
            const double a = hist[2 * j] / hist[2 * j + 1];
            d[j] = 1.0 / a;
            if (j > 0) {
                const double ap = hist[2 * (j - 1)] / hist[2 * (j - 1) + 1];
                const double bp = hist[2 * j] / hist[2 * (j - 1)];
                d[j] += bp / ap;
                e[j - 1] = std::sqrt(bp) / ap;
            }
        }
        for (int j = 0; j < k; ++j) {
            const double rad = (j > 0 ? std::fabs(e[j - 1]) : 0.0) + std::fabs(e[j]);
            lo = std::min(lo, d[j] - rad);
            hi = std::max(hi, d[j] + rad);
        }
    }
    int below(double x) const {   // eigenvalues of T_k below x
        int cnt = 0;
        double q = d[0] - x;
        for (int j = 0;;) {
            if (q < 0) ++cnt;
            if (++j == k) break;
            if (q == 0) q = 1e-300;
            q = d[j] - x - e[j - 1] * e[j - 1] / q;
        }
        return cnt;
    }
    double bisect(int target, double rel) const {   // smallest x with below(x) >= target
        double a = lo, b = hi;
        for (int i = 0; i < 200 && b - a > rel * std::max(std::fabs(a), std::fabs(b)); ++i) {
            const double m = 0.5 * (a + b);
            if (below(m) >= target) b = m; else a = m;
        }
        return 0.5 * (a + b);
    }
};

}  // namespace

double cgConditionEstimate(const double* hist, int k, double rel) {
    if (k < 2) return 1.0;
    const Lanczos T(hist, k);
    const double lmin = T.bisect(1, rel), lmax = T.bisect(k, rel);
    return lmin > 0 ? lmax / lmin : 1e300;
}

// Whether the same estimate exceeds limit, without resolving the smallest
// eigenvalue: the largest to 1 % (a few Sturm counts from the Gershgorin
// interval), then one count at lmax / limit.  Agrees with
// cgConditionEstimate(...) > limit except for estimates within about 1 % of
// the limit.
bool cgConditionExceeds(const double* hist, int k, double limit) {
    if (k < 2) return 1.0 > limit;
    const Lanczos T(hist, k);
    const double lmax = T.bisect(k, 1e-2);
    return T.below(lmax / limit) >= 1;
}

}  // namespace feelpp_amd
