// C entry points around the reference's Eigen-free numerics, compiled together
// with src/Planck.cpp and src/Legendre.cpp from the reference tree into
// oracle/_ref/libref_parts.so.  Used only by tests to pin the oracle's
// restatement of these functions (same pivot sequence, same bits).
#include <exception>

#include "BandSolver.hpp"
#include "Legendre.hpp"
#include "Planck.hpp"

extern "C" {

int ref_planck2(double lo, double hi, double t, double* out) {
  try {
    *out = disortpp::planckFunction2(lo, hi, t);
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

int ref_planck_deriv2(double lo, double hi, double t, double* out) {
  try {
    *out = disortpp::planckFunctionDeriv2(lo, hi, t);
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

void ref_legendre_flat(int nmu, int m, int maxmu, int twonm1, const double* mu, double* ylm) {
  disortpp::legendrePolynomialsFlat(nmu, m, maxmu, twonm1, mu, ylm);
}

void ref_band_factor(double* abd, int lda, int n, int ml, int mu, int* ipvt) {
  disortpp::detail::bandFactor(abd, lda, n, ml, mu, ipvt);
}

void ref_band_solve(const double* abd, int lda, int n, int ml, int mu, const int* ipvt, double* b) {
  disortpp::detail::bandSolve(abd, lda, n, ml, mu, ipvt, b);
}
}
