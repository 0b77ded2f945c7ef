// CPU FP64 oracle for the DISORT flux path -- TEST INFRASTRUCTURE ONLY.
//
// This directory is a from-scratch restatement of the DisORT++ algorithm
// (NewStrangeWorlds/DisORT, GPL-3.0; no reference text is copied).  It exists
// to check the GPU library: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and never as the thing measured
// or shipped.  Every function cites the reference file:line whose behaviour
// it restates.
//
// Parity status of each function (what pins it):
//   planck_band / planck_band_deriv : bit-compared against the reference
//        src/Planck.cpp compiled into oracle/_ref (tests/test_oracle_numerics.py)
//        plus golden vectors in tests/golden/planck.json.
//   legendre_flat   : closed forms P0..P2 (tests/test_utils.cpp:150-210) and
//        the reference src/Legendre.cpp compiled into oracle/_ref.
//   gauss_quadrature: numpy.polynomial.legendre.leggauss.
//   band_factor / band_solve : brute force dense solve on small systems and
//        the reference src/BandSolver.hpp compiled into oracle/_ref
//        (same pivot sequence).
//   sym_eigen (Jacobi), cholesky, lu_solve : numpy.linalg on random SPD /
//        general matrices; invariants (orthogonality, reconstruction).
//   flux_solve      : DisORT++ self-regression values (13 digits,
//        tests/test_solver.cpp:5238-5434), C-DISORT disotest tables
//        (9a/9b, 6b/6c, thermal T1-T4), closed forms (Beer's law,
//        transparent medium), energy conservation at omega=1.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace oracle {

// ---- numerics (restating src/Planck.cpp, src/Legendre.cpp, ...) ----------
double planck_band(double wnumlo, double wnumhi, double temp);        // planckFunction2
double planck_band_deriv(double wnumlo, double wnumhi, double temp);  // planckFunctionDeriv2
void legendre_flat(int nmu, int m, int maxmu, int twonm1, const double* mu, double* ylm);
void gauss_quadrature(int nn, double* cmu /*[2nn]*/, double* cwt /*[2nn]*/);
double chapman(int lc, double tau_frac, const double* level_altitudes,
               const double* dtau_lyr, int num_layers, double r, double mu0);

// LINPACK-semantics band LU (column-major, lda rows, 1-based ipvt).
void band_factor(double* abd, int lda, int n, int ml, int mu, int* ipvt);
void band_solve(const double* abd, int lda, int n, int ml, int mu, const int* ipvt, double* b);

// Dense helpers, column-major n x n.
bool cholesky_lower(int n, const double* a, double* l);  // a = l l^T; false if not SPD
void sym_eigen_jacobi(int n, const double* a, double* evals, double* evecs); // ascending
// Partial-pivot LU of a (in place) + solve of one right-hand side.
struct DenseLU {
  int n = 0;
  std::vector<double> lu;  // column-major
  std::vector<int> piv;
  void factor(int n, const double* a);
  void solve(double* b) const;
  double frob_norm_of_input = 0.0;
};

// ---- flux solver (restating src/FluxSolver.cpp DisortFluxSolver<N>) -------
struct FluxConfig {
  int nlyr = 0, nstr = 0, nmom = 0;           // pmom rows hold max(nmom,nstr)+1 entries
  bool thermal = false, spherical = false, deltam_plus = false;
  bool diffusion_bc = false, index_from_bottom = false, jacobian = false;
  double beam_flux = 0, beam_mu = 0.5, iso_top = 0, iso_bot = 0;
  double temp_top = 0, temp_bot = 0, emis_top = 1, albedo = 0;
  double wn_lo = 0, wn_hi = 0, radius = 0;
  std::vector<double> dtau, ssa, temper, alt;
  std::vector<double> pmom;                   // [nlyr][pstride]
  int pstride() const { return (nmom > nstr ? nmom : nstr) + 1; }
  double pm(int k, int lc) const { return pmom[(size_t)lc * pstride() + k]; }
};

struct FluxOutput {
  // flux_direct_beam, flux_down, flux_up, flux_tau_divergence, mean_intensity,
  // mean_intensity_down, mean_intensity_up, mean_intensity_direct_beam
  std::vector<double> f[8];
  // Temperature Jacobians [6][(nlyr+1)*(nlyr+2)] when requested.
  std::vector<double> jac[6];
};

// Throws std::invalid_argument / std::runtime_error with the reference's messages.
void validate_flux_config(const FluxConfig& c);
FluxOutput flux_solve(const FluxConfig& c);

}  // namespace oracle
