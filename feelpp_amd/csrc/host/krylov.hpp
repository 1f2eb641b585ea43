// Krylov diagnostics shared by the device solvers (host code).
#pragma once

namespace feelpp_amd {

// Condition number estimate of the Jacobi-preconditioned operator D^{-1} A
// from k CG iterations; hist holds k + 1 pairs (gamma_j = (r_j, z_j),
// pAp_j = (p_j, A p_j)), the last pAp unused.  Returns 1 for k < 2.
double cgConditionEstimate(const double* hist, int k, double rel = 1e-14);
bool cgConditionExceeds(const double* hist, int k, double limit);

}  // namespace feelpp_amd
